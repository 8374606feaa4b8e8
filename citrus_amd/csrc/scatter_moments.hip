// Scattering-source moments T[go][n][k] = sum_gi mu_s_l[gi][go][l(n)] phi[gi][n][k] for every cell (gfx950).
//
// A workgroup owns kCells cells of one view. It first forms each cell's transfer matrices mu[l][gi][go] for the
// degrees l <= L in LDS (from the dense volume, or from the densities and the per-material table), once per cell
// instead of once per moment; then one thread per (cell, n, k) element keeps phi[.][n][k] of every input group in
// registers and accumulates all output groups from the matrix of its degree. phi and T are streamed once, coalesced
// along the contiguous [n][k] of each cell.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "citrus_hip.h"
#include "common.h"

namespace citrus {
namespace {

constexpr int kCells = 2;
constexpr int kMaxMaterials = 8;

__device__ __forceinline__ int degree_of(int n) {
  int l = 0;
  while ((l + 1) * (l + 1) <= n) ++l;
  return l;
}

// row stride of one degree's [GMAX][GMAX] matrix in LDS: one word of padding puts the rows of different degrees,
// which the lanes of one wave read at the same (gi, go), in different banks
template <int GMAX>
constexpr int degree_stride() { return GMAX * GMAX + 1; }

// LDS: mu [kCells][L+1][GMAX][GMAX] (degree stride padded), then the material table [M][G][G][Lt+1] (factored form)
template <int GMAX, bool FACTORED>
__global__ __launch_bounds__(kCells * 256) void scatter_moments_kernel(
    const float* __restrict__ transfer, const float* __restrict__ rho, int M, int Lt, const float* __restrict__ phi,
    float* __restrict__ T, int G, int nsh, int L, int N, int V, int lower_only, float* __restrict__ means) {
  extern __shared__ float lds[];
  constexpr int DS = degree_stride<GMAX>();
  const int per_cell = (L + 1) * DS + 1;  // odd: the two cells' matrices also fall in different banks
  float* mu = lds;
  float* table = lds + kCells * per_cell;
  const int nthreads = blockDim.x;
  const int tid = threadIdx.x;
  const int cells_per_view = (N + kCells - 1) / kCells;
  // workgroups go to the 8 XCDs round robin; give each XCD a contiguous run of cells instead, so that the 4-byte
  // cell-mean stores of neighbouring cells (4 bytes apart, 16 MB apart across moments) meet in one L2
  const unsigned nb = gridDim.x;
  const unsigned b = (nb % 8 == 0) ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;
  const int v = (int)(b / cells_per_view);
  const int c0 = (int)(b % cells_per_view) * kCells;
  if (FACTORED) {
    const int nt = M * G * G * (Lt + 1);
    for (int i = tid; i < nt; i += nthreads) table[i] = transfer[i];
    __syncthreads();
  }
  // per-cell transfer matrices, entries (l, gi, go) with compile-time strides
  const int entries = (L + 1) * GMAX * GMAX;
  for (int cc = 0; cc < kCells; ++cc) {
    const int cell = c0 + cc;
    for (int r = tid; r < entries; r += nthreads) {
      const int l = r / (GMAX * GMAX);
      const int gi = (r / GMAX) % GMAX;
      const int go = r % GMAX;
      float m_val = 0.0f;
      if (cell < N && gi < G && go < G && !(lower_only && gi > go)) {
        if (FACTORED) {
#pragma unroll
          for (int m = 0; m < kMaxMaterials; ++m)
            if (m < M)
              m_val = fmaf(rho[(int64_t)m * N + cell], table[((m * G + gi) * G + go) * (Lt + 1) + l], m_val);
        } else {
          m_val = transfer[(((int64_t)gi * G + go) * (Lt + 1) + l) * N + cell];
        }
      }
      mu[cc * per_cell + l * DS + gi * GMAX + go] = m_val;
    }
  }
  __syncthreads();
  const int elems = nsh * 4;
  const int cc = tid / elems;
  const int r = tid - cc * elems;
  const int cell = c0 + cc;
  if (cc >= kCells || cell >= N) return;
  const int l = degree_of(r / 4);
  const int64_t view_stride = (int64_t)G * N * elems;
  const int64_t group_stride = (int64_t)N * elems;
  const int64_t off = (int64_t)v * view_stride + (int64_t)cell * elems + r;
  float p[GMAX];
#pragma unroll
  for (int gi = 0; gi < GMAX; ++gi) p[gi] = gi < G ? phi[off + gi * group_stride] : 0.0f;
  const float* mcell = mu + cc * per_cell + l * DS;
#pragma unroll
  for (int go = 0; go < GMAX; ++go) {
    if (go >= G) break;
    float acc = 0.0f;
#pragma unroll
    for (int gi = 0; gi < GMAX; ++gi)
      if (gi < G && !(lower_only && gi > go)) acc = fmaf(mcell[gi * GMAX + go], p[gi], acc);
    T[off + go * group_stride] = acc;
    // the cell means again, moment-major [V][nsh][G][N]: the operand of the image's multiple-collision expansion
    if (means && (r & 3) == 0) means[(((int64_t)v * nsh + (r >> 2)) * G + go) * N + cell] = acc;
  }
}

int check(int G, int nsh, int Lt, int64_t N, int V, int& L) {
  if (G < 1 || G > 32) return fail(CITRUS_ERR_ARG, "1..32 groups supported");
  if (N < 1 || V < 1 || N > 0x7fffffff) return fail(CITRUS_ERR_ARG, "N and V must be >= 1 (N < 2^31)");
  L = 0;
  while ((L + 1) * (L + 1) < nsh) ++L;
  if ((L + 1) * (L + 1) != nsh || nsh > 64) return fail(CITRUS_ERR_ARG, "nsh must be (L+1)^2 with L <= 7");
  if (Lt < L) return fail(CITRUS_ERR_ARG, "transfer Legendre order below the moments' degree");
  return CITRUS_OK;
}

template <bool FACTORED>
int launch(const float* transfer, const float* rho, int M, int Lt, const float* phi, float* T, int G, int nsh,
           int64_t N, int V, int lower_only, float* means, void* stream) {
  int L = 0;
  if (int rc = check(G, nsh, Lt, N, V, L)) return rc;
  const int gmax = G <= 8 ? 8 : (G <= 16 ? 16 : 32);
  const size_t per_cell = (size_t)(L + 1) * (gmax * gmax + 1) + 1;
  const size_t lds = sizeof(float) * ((size_t)kCells * per_cell + (FACTORED ? (size_t)M * G * G * (Lt + 1) : 0));
  if (lds > 64 * 1024) return fail(CITRUS_ERR_UNSUPPORTED, "transfer tables exceed 64 KiB of LDS");
  const int threads = kCells * nsh * 4;
  const int64_t blocks = (int64_t)V * ((N + kCells - 1) / kCells);
  if (blocks > 0x7fffffffLL) return fail(CITRUS_ERR_ARG, "too many cells");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)blocks), block((unsigned)threads);
  if (G <= 8)
    hipLaunchKernelGGL((scatter_moments_kernel<8, FACTORED>), grid, block, lds, s, transfer, rho, M, Lt, phi, T, G,
                       nsh, L, (int)N, V, lower_only, means);
  else if (G <= 16)
    hipLaunchKernelGGL((scatter_moments_kernel<16, FACTORED>), grid, block, lds, s, transfer, rho, M, Lt, phi, T, G,
                       nsh, L, (int)N, V, lower_only, means);
  else
    hipLaunchKernelGGL((scatter_moments_kernel<32, FACTORED>), grid, block, lds, s, transfer, rho, M, Lt, phi, T, G,
                       nsh, L, (int)N, V, lower_only, means);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("scatter moments: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

// ---- bfloat16 operand of the source contraction -------------------------------------------------------------------
// T [V][G][N][nsh][4] float32 -> T16 [V][G][N][4][KP] bfloat16 (KP = nsh rounded up to 16, zero padded), each
// (view, group) scaled by a power of two 2^-e with max|T| in [2^e, 2^(e+1)), so every group's largest value is in
// [1, 2); inv_scale[v][g] = 2^e undoes it after the contraction.

constexpr int kAmaxBlock = 256;

__global__ __launch_bounds__(kAmaxBlock) void group_amax_kernel(const float* __restrict__ T, int64_t per_group,
                                                                int blocks_per_group, float* __restrict__ partial) {
  __shared__ float red[kAmaxBlock];
  const int vg = blockIdx.y, b = blockIdx.x;
  const int64_t chunk = (per_group + blocks_per_group - 1) / blocks_per_group;
  const int64_t lo = (int64_t)b * chunk, hi = lo + chunk < per_group ? lo + chunk : per_group;
  const float* src = T + (int64_t)vg * per_group;
  float m = 0.0f;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kAmaxBlock) m = fmaxf(m, fabsf(src[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = kAmaxBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(int64_t)vg * blocks_per_group + b] = red[0];
}

__global__ void group_scale_kernel(const float* __restrict__ partial, int blocks_per_group, int n,
                                   float* __restrict__ scale, float* __restrict__ inv_scale) {
  const int vg = blockIdx.x * blockDim.x + threadIdx.x;
  if (vg >= n) return;
  float m = 0.0f;
  for (int b = 0; b < blocks_per_group; ++b) m = fmaxf(m, partial[(int64_t)vg * blocks_per_group + b]);
  const int e = m > 0.0f ? ilogbf(m) : 0;
  scale[vg] = ldexpf(1.0f, -e);
  inv_scale[vg] = ldexpf(1.0f, e);
}

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  const uint32_t b = __float_as_uint(f);
  return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

__global__ void to_bf16_operand_kernel(const float* __restrict__ T, int nsh, int kp, int64_t N, int64_t n_vg,
                                       const float* __restrict__ scale, uint16_t* __restrict__ T16) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (vg, cell, comp, n), n fastest
  const int64_t total = n_vg * N * 4 * kp;
  if (i >= total) return;
  const int n = (int)(i % kp);
  const int comp = (int)((i / kp) % 4);
  const int64_t cell_vg = i / (4 * (int64_t)kp);  // vg * N + cell
  const int64_t vg = cell_vg / N;
  T16[i] = n < nsh ? bf16_rne(T[(cell_vg * nsh + n) * 4 + comp] * scale[vg]) : (uint16_t)0;
}

}  // namespace
}  // namespace citrus

extern "C" int64_t citrus_bf16_operand_workspace_floats(int G, int V) {
  if (G < 1 || V < 1) return -1;
  return (int64_t)G * V * (64 + 1);
}

extern "C" int citrus_bf16_operand(const float* T, int G, int nsh, int64_t N, int V, uint16_t* T16, float* inv_scale,
                                   float* workspace, void* stream) {
  using namespace citrus;
  if (!T || !T16 || !inv_scale || !workspace) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (G < 1 || V < 1 || N < 1 || nsh < 1 || nsh > 64) return fail(CITRUS_ERR_ARG, "bad sizes");
  const int kp = (nsh + 15) / 16 * 16;
  const int blocks_per_group = 64;
  const int n_vg = G * V;
  float* partial = workspace;
  float* scale = workspace + (int64_t)n_vg * blocks_per_group;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(group_amax_kernel, dim3(blocks_per_group, n_vg), dim3(kAmaxBlock), 0, s, T, N * nsh * 4,
                     blocks_per_group, partial);
  hipLaunchKernelGGL(group_scale_kernel, dim3((n_vg + 63) / 64), dim3(64), 0, s, partial, blocks_per_group, n_vg,
                     scale, inv_scale);
  const int64_t total = (int64_t)n_vg * N * 4 * kp;
  if ((total + 255) / 256 > 0x7fffffffLL) return fail(CITRUS_ERR_ARG, "operand too large");
  hipLaunchKernelGGL(to_bf16_operand_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, T, nsh, kp, N,
                     (int64_t)n_vg, scale, T16);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("bf16 operand: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

extern "C" int citrus_scatter_moments_dense(const float* transfer, int Lt, const float* phi, float* T, int G, int nsh,
                                            int64_t N, int V, void* stream) {
  using namespace citrus;
  if (!transfer || !phi || !T) return fail(CITRUS_ERR_ARG, "NULL buffer");
  return launch<false>(transfer, nullptr, 0, Lt, phi, T, G, nsh, N, V, 0, nullptr, stream);
}

extern "C" int citrus_scatter_moments_factored(const float* rho, int num_materials, const float* sigma, int Lt,
                                               int downscatter_only, const float* phi, float* T, int G, int nsh,
                                               int64_t N, int V, float* cell_means, void* stream) {
  using namespace citrus;
  if (!rho || !sigma || !phi || !T) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (num_materials < 1 || num_materials > kMaxMaterials) return fail(CITRUS_ERR_ARG, "1..8 materials supported");
  return launch<true>(sigma, rho, num_materials, Lt, phi, T, G, nsh, N, V, downscatter_only ? 1 : 0, cell_means,
                      stream);
}
