// Footprint filter matrices of the collapsed-cone image (gfx950).
//
// box_matrix (citrus/scatter_solver/scatter_image.py:952-1007): B[c][j] = |[j, j+1] n W_c| / |W_c| with the window
// W_c = [c + 1/2 - w/2, c + 1/2 + w/2] n [0, N] of width w (clamped below at 1e-3) centred on pixel c, clipped to the
// panel and renormalised there. One thread per matrix entry; the blur itself is a batched GEMM of the images with
// these matrices.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "citrus_hip.h"
#include "common.h"

namespace citrus {
namespace {

__global__ __launch_bounds__(256) void box_matrix_kernel(const float* __restrict__ w, int k, int N,
                                                         float* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t total = (int64_t)k * N * N;
  if (idx >= total) return;
  const int j = (int)(idx % N);
  const int c = (int)((idx / N) % N);
  const int m = (int)(idx / ((int64_t)N * N));
  const float half = 0.5f * fmaxf(w[m], 1e-3f);
  const float nf = (float)N;
  const float center = (float)c + 0.5f;
  const float lo = fminf(fmaxf(center - half, 0.0f), nf);
  const float hi = fminf(fmaxf(center + half, 0.0f), nf);
  const float jf = (float)j;
  out[idx] = fmaxf(fminf(hi, jf + 1.0f) - fmaxf(lo, jf), 0.0f) / (hi - lo);
}

}  // namespace
}  // namespace citrus

extern "C" int citrus_box_matrix(const float* widths, int k, int N, float* out, void* stream) {
  using namespace citrus;
  if (!widths || !out) return fail(CITRUS_ERR_ARG, "NULL widths or output");
  if (k < 0 || N < 1) return fail(CITRUS_ERR_ARG, "k >= 0 and N >= 1 required");
  const int64_t total = (int64_t)k * N * N;
  if (total == 0) return CITRUS_OK;
  hipLaunchKernelGGL(box_matrix_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), widths, k, N, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("box matrix: ") + hipGetErrorString(e));
  return CITRUS_OK;
}
