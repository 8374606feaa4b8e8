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

// B[c][j] for a window of width w centred on pixel c of N
__device__ __forceinline__ float box_value(float w, int N, int c, int j) {
  const float half = 0.5f * fmaxf(w, 1e-3f);
  const float nf = (float)N;
  const float center = (float)c + 0.5f;
  const float lo = fminf(fmaxf(center - half, 0.0f), nf);
  const float hi = fminf(fmaxf(center + half, 0.0f), nf);
  const float jf = (float)j;
  return fmaxf(fminf(hi, jf + 1.0f) - fmaxf(lo, jf), 0.0f) / (hi - lo);
}

__global__ __launch_bounds__(256) void box_matrix_kernel(const float* __restrict__ w, int k, int N,
                                                         float* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t total = (int64_t)k * N * N;
  if (idx >= total) return;
  const int j = (int)(idx % N);
  const int c = (int)((idx / N) % N);
  const int m = (int)(idx / ((int64_t)N * N));
  out[idx] = box_value(w[m], N, c, j);
}

// Emission-weighted mean panel height of each direction from the emission kernel's per-block sums
// partial [K][nb][2] = (sum e, sum e h): one workgroup per direction, a fixed-order tree over the blocks
// (scatter_image.py:1033-1060, the footprint's hbar_k); the box-centre height where the emission vanishes.
__global__ __launch_bounds__(256) void footprint_heights_kernel(const float* __restrict__ partial, int nb,
                                                                float center_height, float* __restrict__ heights) {
  __shared__ float s0[256], s1[256];
  const int k = blockIdx.x, t = threadIdx.x;
  float a = 0.0f, b = 0.0f;
  for (int i = t; i < nb; i += 256) {
    a += partial[((int64_t)k * nb + i) * 2];
    b += partial[((int64_t)k * nb + i) * 2 + 1];
  }
  s0[t] = a;
  s1[t] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s0[t] += s0[t + w];
      s1[t] += s1[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    const float den = s0[0], num = s1[0];
    heights[k] = den > 0.0f ? num / fmaxf(den, 1.17549435e-38f) : center_height;
  }
}

// Both footprint filters of a chunk from the mean heights: widths (h cell_width[k][a]) scale_a in pixels; the
// channel filter transposed, channel_t[k][j][c] = B_c[c][j] w[k][c] (the right operand of the channel blur, its
// columns scaled by the optional channel weights w), the row filter rows[k][c][j] = B_r[c][j].
__global__ __launch_bounds__(256) void footprint_filters_kernel(const float* __restrict__ heights,
                                                                const float* __restrict__ cell_width, float sc,
                                                                float sr, int K, int Nc, int Nr,
                                                                const float* __restrict__ channel_weight,
                                                                float* __restrict__ channel_t,
                                                                float* __restrict__ rows) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nc = (int64_t)K * Nc * Nc, nr = (int64_t)K * Nr * Nr;
  if (idx < nc) {
    const int c = (int)(idx % Nc);            // column of the transposed matrix: the window's pixel
    const int j = (int)((idx / Nc) % Nc);     // row: the pixel covered
    const int m = (int)(idx / ((int64_t)Nc * Nc));
    const float b = box_value(heights[m] * cell_width[2 * m] * sc, Nc, c, j);
    channel_t[idx] = channel_weight ? b * channel_weight[(int64_t)m * Nc + c] : b;
  } else if (idx < nc + nr) {
    const int64_t i = idx - nc;
    const int j = (int)(i % Nr);
    const int c = (int)((i / Nr) % Nr);
    const int m = (int)(i / ((int64_t)Nr * Nr));
    rows[i] = box_value(heights[m] * cell_width[2 * m + 1] * sr, Nr, c, j);
  }
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

extern "C" int citrus_footprint_filters(const float* partial, int K, int nblocks, float center_height,
                                        const float* cell_width, float scale_channel, float scale_row, int Nc, int Nr,
                                        const float* channel_weight, float* heights,
                                        float* channel_filter_t, float* row_filter, void* stream) {
  using namespace citrus;
  if (!partial || !cell_width || !heights || !channel_filter_t || !row_filter)
    return fail(CITRUS_ERR_ARG, "NULL input or output");
  if (K < 0 || K > 65535 || nblocks < 1 || Nc < 1 || Nr < 1)
    return fail(CITRUS_ERR_ARG, "0 <= K <= 65535, nblocks >= 1, Nc >= 1 and Nr >= 1 required");
  if (K == 0) return CITRUS_OK;
  const int64_t total = (int64_t)K * Nc * Nc + (int64_t)K * Nr * Nr;
  if (!grid_fits((total + 255) / 256, 256)) return fail(CITRUS_ERR_ARG, "filters too large for one launch");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(footprint_heights_kernel, dim3((unsigned)K), dim3(256), 0, s, partial, nblocks, center_height,
                     heights);
  hipLaunchKernelGGL(footprint_filters_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, heights,
                     cell_width, scale_channel, scale_row, K, Nc, Nr, channel_weight, channel_filter_t,
                     row_filter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("footprint filters: ") + hipGetErrorString(e));
  return CITRUS_OK;
}
