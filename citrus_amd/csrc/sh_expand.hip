// Spherical-harmonic expansion of cell moments at a set of directions (gfx950):
//   C[k][j] = sum_n Y[k][n] B[n][j]        (k: directions, n: SH moments, j: (group, cell) columns)
// the multiple-collision source of the collapsed-cone image at a chunk of cone directions, from the cell means of
// the scattering-source moments (citrus/scatter_solver/scatter_image.py:1225,1246, `torch.mm(harmonics, moments_by_cell)`).
//
// K is tiny (<= 64 per chunk), nsh <= 36 and the column count is G x cells (4.2 M at 64^3 x 16), so the product is
// HBM-bound (read B once, write C once). One thread per column keeps its nsh moments in registers and produces the
// K outputs one after another (each accumulated in n order, float32); Y is the same for every lane and is read from
// LDS as a broadcast. Loads of B and stores of C are coalesced along the columns.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "citrus_hip.h"
#include "common.h"

namespace citrus {
namespace {

constexpr int kThreads = 256;

template <int NSH>
__global__ __launch_bounds__(kThreads) void sh_expand_kernel(const float* __restrict__ Y, int K,
                                                             const float* __restrict__ B, int64_t ncol,
                                                             float* __restrict__ C) {
  extern __shared__ float ys[];  // Y [K][NSH], read by every lane at the same address (LDS broadcast)
  for (int i = threadIdx.x; i < K * NSH; i += kThreads) ys[i] = Y[i];
  __syncthreads();
  const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (j >= ncol) return;
  float b[NSH];
#pragma unroll
  for (int n = 0; n < NSH; ++n) b[n] = B[(int64_t)n * ncol + j];
  // one direction at a time (4 in flight): the output store of each is coalesced along the columns
#pragma unroll 4
  for (int k = 0; k < K; ++k) {
    const float* y = ys + k * NSH;
    float a = 0.0f;
#pragma unroll
    for (int n = 0; n < NSH; ++n) a = fmaf(y[n], b[n], a);
    C[(int64_t)k * ncol + j] = a;
  }
}

template <int NSH>
hipError_t launch(const float* Y, int K, const float* B, int64_t ncol, float* C, hipStream_t s) {
  hipLaunchKernelGGL(sh_expand_kernel<NSH>, dim3((unsigned)((ncol + kThreads - 1) / kThreads)), dim3(kThreads),
                     sizeof(float) * (size_t)K * NSH, s, Y, K, B, ncol, C);
  return hipGetLastError();
}

}  // namespace
}  // namespace citrus

extern "C" int citrus_sh_expand(const float* Y, int K, int nsh, const float* B, int64_t ncol, float* C,
                                void* stream) {
  using namespace citrus;
  if (!Y || !B || !C) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (K < 1 || ncol < 1) return fail(CITRUS_ERR_ARG, "K and ncol must be >= 1");
  if (K * nsh > 16384) return fail(CITRUS_ERR_ARG, "K * nsh must be <= 16384 (the directions' harmonics sit in LDS)");
  if ((ncol + kThreads - 1) / kThreads > 0x7fffffffLL) return fail(CITRUS_ERR_ARG, "too many columns");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipError_t e;
  switch (nsh) {
    case 1: e = launch<1>(Y, K, B, ncol, C, s); break;
    case 4: e = launch<4>(Y, K, B, ncol, C, s); break;
    case 9: e = launch<9>(Y, K, B, ncol, C, s); break;
    case 16: e = launch<16>(Y, K, B, ncol, C, s); break;
    case 25: e = launch<25>(Y, K, B, ncol, C, s); break;
    case 36: e = launch<36>(Y, K, B, ncol, C, s); break;
    case 49: e = launch<49>(Y, K, B, ncol, C, s); break;
    case 64: e = launch<64>(Y, K, B, ncol, C, s); break;
    default: return fail(CITRUS_ERR_UNSUPPORTED, "nsh must be (L+1)^2 with L <= 7");
  }
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("sh expand: ") + hipGetErrorString(e));
  return CITRUS_OK;
}
