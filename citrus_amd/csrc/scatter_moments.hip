// Scattering-source moments T[go][n][k] = sum_gi mu_s_l[gi][go][l(n)] phi[gi][n][k] for every cell (gfx950).
//
// One thread per (view, cell, moment n, LD moment k): consecutive threads walk the contiguous [n][k] of a cell, so
// every read of phi[gi] and write of T[go] is a coalesced stream and phi is read from HBM once. Output groups are
// accumulated in registers in chunks of 8; the transfer coefficient of a (gi, go, l) triple is read from the dense
// volume or formed from the densities and the per-material table held in LDS.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "citrus_hip.h"
#include "common.h"

namespace citrus {
namespace {

constexpr int kThreads = 256;
constexpr int kGoChunk = 8;
constexpr int kMaxMaterials = 8;

__device__ __forceinline__ int degree_of(int n) {
  int l = (int)sqrtf((float)n);
  while (l * l > n) --l;
  while ((l + 1) * (l + 1) <= n) ++l;
  return l;
}

template <bool FACTORED>
__global__ __launch_bounds__(kThreads) void scatter_moments_kernel(const float* __restrict__ transfer,
                                                                   const float* __restrict__ rho, int M,
                                                                   int Lt, const float* __restrict__ phi,
                                                                   float* __restrict__ T, int G, int nsh,
                                                                   int64_t N, int64_t total, int lower_only) {
  extern __shared__ float sigma_l[];  // factored: [M][G][G][Lt+1]
  if (FACTORED) {
    const int ns = M * G * G * (Lt + 1);
    for (int i = threadIdx.x; i < ns; i += blockDim.x) sigma_l[i] = transfer[i];
    __syncthreads();
  }
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int64_t per_view = N * nsh * 4;
  const int64_t v = t / per_view;
  const int64_t r = t - v * per_view;  // (cell, n, k) within the view
  const int64_t cell = r / (nsh * 4);
  const int n = (int)((r / 4) % nsh);
  const int l = degree_of(n);
  const float* phi_v = phi + v * G * per_view;
  float* T_v = T + v * G * per_view;
  float rho_c[kMaxMaterials];
  if (FACTORED) {
    for (int m = 0; m < M; ++m) rho_c[m] = rho[(int64_t)m * N + cell];
  }
  for (int go0 = 0; go0 < G; go0 += kGoChunk) {
    float acc[kGoChunk];
#pragma unroll
    for (int j = 0; j < kGoChunk; ++j) acc[j] = 0.0f;
    const int gi_end = lower_only ? (go0 + kGoChunk < G ? go0 + kGoChunk : G) : G;
    for (int gi = 0; gi < gi_end; ++gi) {
      const float p = phi_v[(int64_t)gi * per_view + r];
#pragma unroll
      for (int j = 0; j < kGoChunk; ++j) {
        const int go = go0 + j;
        if (go >= G || (lower_only && gi > go)) continue;
        float mu;
        if (FACTORED) {
          mu = 0.0f;
          for (int m = 0; m < M; ++m) mu = fmaf(rho_c[m], sigma_l[((m * G + gi) * G + go) * (Lt + 1) + l], mu);
        } else {
          mu = transfer[(((int64_t)gi * G + go) * (Lt + 1) + l) * N + cell];
        }
        acc[j] = fmaf(mu, p, acc[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < kGoChunk; ++j)
      if (go0 + j < G) T_v[(int64_t)(go0 + j) * per_view + r] = acc[j];
  }
}

int check(int G, int nsh, int Lt, int64_t N, int V) {
  if (G < 1 || N < 1 || V < 1) return fail(CITRUS_ERR_ARG, "G, N and V must be >= 1");
  int L = 0;
  while ((L + 1) * (L + 1) < nsh) ++L;
  if ((L + 1) * (L + 1) != nsh) return fail(CITRUS_ERR_ARG, "nsh must be a square (L+1)^2");
  if (Lt < L) return fail(CITRUS_ERR_ARG, "transfer Legendre order below the moments' degree");
  return CITRUS_OK;
}

}  // namespace
}  // namespace citrus

extern "C" int citrus_scatter_moments_dense(const float* transfer, int Lt, const float* phi, float* T, int G, int nsh,
                                            int64_t N, int V, void* stream) {
  using namespace citrus;
  if (int rc = check(G, nsh, Lt, N, V)) return rc;
  if (!transfer || !phi || !T) return fail(CITRUS_ERR_ARG, "NULL buffer");
  const int64_t total = (int64_t)V * N * nsh * 4;
  const int64_t blocks = (total + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(scatter_moments_kernel<false>, dim3((unsigned)blocks), dim3(kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), transfer, nullptr, 0, Lt, phi, T, G, nsh, N, total, 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("scatter moments: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

extern "C" int citrus_scatter_moments_factored(const float* rho, int num_materials, const float* sigma, int Lt,
                                               int downscatter_only, const float* phi, float* T, int G, int nsh,
                                               int64_t N, int V, void* stream) {
  using namespace citrus;
  if (int rc = check(G, nsh, Lt, N, V)) return rc;
  if (!rho || !sigma || !phi || !T) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (num_materials < 1 || num_materials > kMaxMaterials) return fail(CITRUS_ERR_ARG, "1..8 materials supported");
  const size_t lds = sizeof(float) * (size_t)num_materials * G * G * (Lt + 1);
  if (lds > 64 * 1024) return fail(CITRUS_ERR_UNSUPPORTED, "transfer table exceeds 64 KiB of LDS");
  const int64_t total = (int64_t)V * N * nsh * 4;
  const int64_t blocks = (total + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(scatter_moments_kernel<true>, dim3((unsigned)blocks), dim3(kThreads), lds,
                     reinterpret_cast<hipStream_t>(stream), sigma, rho, num_materials, Lt, phi, T, G, nsh, N, total,
                     downscatter_only ? 1 : 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("scatter moments: ") + hipGetErrorString(e));
  return CITRUS_OK;
}
