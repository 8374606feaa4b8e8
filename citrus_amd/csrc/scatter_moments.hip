// Scattering-source moments T[go][n][k] = sum_gi mu_s_l[gi][go][l(n)] phi[gi][n][k] for every cell (gfx950).
//
// A workgroup owns CELLS cells of one view (4 up to 16 groups, else 2). It first forms each cell's transfer matrices mu[l][gi][go] for the
// degrees l <= L in LDS (from the dense volume, or from the densities and the per-material table), once per cell
// instead of once per moment; then one thread per (cell, n, k) element keeps phi[.][n][k] of every input group in
// registers and accumulates all output groups from the matrix of its degree. phi and T are streamed once, coalesced
// along the contiguous [n][k] of each cell.

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>

#include "citrus_hip.h"
#include "common.h"

namespace citrus {
namespace {

constexpr int kMaxMaterials = 8;
constexpr int kTableShare = 8;  // table words per thread held in registers while phi loads (4096 at 512 threads)
constexpr int kExactRun = 4;    // groups of cells per workgroup in the exact (view step) form

__device__ __forceinline__ int degree_of(int n) {
  int l = 0;
  while ((l + 1) * (l + 1) <= n) ++l;
  return l;
}

// row stride of one degree's [GMAX][GMAX] matrix in LDS: one word of padding puts the rows of different degrees,
// which the lanes of one wave read at the same (gi, go), in different banks
template <int GMAX>
constexpr int degree_stride() { return GMAX * GMAX + 1; }

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  const uint32_t b = __float_as_uint(f);
  return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

// LDS: mu [CELLS][L+1][GMAX][GMAX] (degree stride padded), then the material table [M][G][G][Lt+1] (factored form).
// MIXED: the result goes out as the mixed-precision operand of the sweep instead of float32 T: degrees l <= 1 to
// T (float32 [V][G][N][4 comp][4]), degrees l >= 2 to T16 (bfloat16 [V][G][N][4 comp][kh], zero padded).
// EXACT: G == GMAX and downscatter only (lower_only), the view step's form: every group and triangle bound is a
// compile-time constant, so the matrix formation and the contraction carry no per-entry tests.
// STORED: phi is a stored accumulator instead (the state store's format: q [V][G][nsh][N] float16 and scale
// [V][G][nsh], the cell means q scale with zero LD slopes, decoded as citrus_state_unpack decodes them), so T equals
// that of the unpacked accumulator bit for bit without the float32 accumulator being written and read.
// RUN: a workgroup takes RUN consecutive groups of CELLS cells in turn: the material table is staged in LDS once per
// workgroup, and the next group's phi is in flight while the current group is contracted. With RUN > 1 the cell
// means of the run are gathered in LDS ([nsh][GMAX][CELLS RUN] after the table) and leave as one run of CELLS RUN
// cells per (moment, group): as 16-byte pieces per group of cells they reached memory as partly written sectors.
template <int GMAX, int CELLS, bool FACTORED, bool MIXED, bool EXACT, int RUN, bool STORED>
__global__ __launch_bounds__(CELLS * 256) void scatter_moments_kernel(
    const float* __restrict__ transfer, const float* __restrict__ rho, int M, int Lt, const float* __restrict__ phi,
    float* __restrict__ T, int G, int nsh, int L, int N, int V, int lower_only, float* __restrict__ means,
    uint16_t* __restrict__ T16, int kh, const __half* __restrict__ qs, const float* __restrict__ qscale) {
  extern __shared__ float lds[];
  constexpr int DS = degree_stride<GMAX>();
  const int per_cell = (L + 1) * DS + 1;  // odd: the two cells' matrices also fall in different banks
  float* mu = lds;
  float* table = lds + CELLS * per_cell;
  constexpr int RC = CELLS * RUN;  // cells of the run
  float* mstage = table + (FACTORED ? M * G * G * (Lt + 1) : 0);  // RUN > 1: [nsh][GMAX][RC]
  const int nthreads = blockDim.x;
  const int tid = threadIdx.x;
  const int runs_per_view = (N + CELLS * RUN - 1) / (CELLS * RUN);
  // workgroups go to the 8 XCDs round robin; give each XCD a contiguous run of cells instead, so that the 4-byte
  // cell-mean stores of neighbouring cells (4 bytes apart, 16 MB apart across moments) meet in one L2
  const unsigned nb = gridDim.x;
  const unsigned b = (nb % 8 == 0) ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;
  const int v = (int)(b / runs_per_view);
  const int run0 = (int)(b % runs_per_view) * CELLS * RUN;
  // this thread's element (cell of the group, n, k); its phi of every input group is requested before the transfer
  // matrices are formed
  const int elems = nsh * 4;
  const int cc = tid / elems;
  const int r = tid - cc * elems;
  const int64_t view_stride = (int64_t)G * N * elems;
  const int64_t group_stride = (int64_t)N * elems;
  // unconditional loads (inactive threads and groups past G read a valid element and discard it): no branches
  // between the loads, so all of them are in flight together
  auto load_phi = [&](const int c0, float (&p)[GMAX]) __attribute__((always_inline)) {
    const int cell = c0 + cc;
    const bool on = cc < CELLS && cell < N;
    if constexpr (STORED) {  // the cell mean (component 0) decoded, the slopes zero
      const int sc = on ? cell : min(c0, N - 1), sn = on ? (r >> 2) : 0;
#pragma unroll
      for (int gi = 0; gi < GMAX; ++gi) {
        const int64_t gs = ((int64_t)v * G + (gi < G ? gi : G - 1)) * nsh + sn;
        const float x = __half2float(qs[gs * N + sc]) * qscale[gs];
        p[gi] = gi < G && (r & 3) == 0 ? x : 0.0f;
      }
      return;
    }
    const int64_t safe = on ? (int64_t)v * view_stride + (int64_t)cell * elems + r
                            : (int64_t)v * view_stride + (int64_t)min(c0, N - 1) * elems;
#pragma unroll
    for (int gi = 0; gi < GMAX; ++gi) {
      const float x = phi[safe + (int64_t)(gi < G ? gi : G - 1) * group_stride];
      p[gi] = gi < G ? x : 0.0f;  // inactive threads never use theirs
    }
  };
  // the factored table's share of this thread (at most kTableShare words, else a loop after the phi loads), loaded
  // before phi so that storing it to LDS waits for its own loads only
  const int nt = FACTORED ? M * G * G * (Lt + 1) : 0;
  const bool share_fits = nt <= kTableShare * nthreads;
  float tv[kTableShare];
  if (FACTORED && share_fits) {
#pragma unroll
    for (int j = 0; j < kTableShare; ++j) {
      const int i = tid + j * nthreads;
      tv[j] = transfer[i < nt ? i : 0];
    }
  }
  float p[GMAX];
  load_phi(run0, p);
  if (FACTORED) {
    if (share_fits) {
#pragma unroll
      for (int j = 0; j < kTableShare; ++j) {
        const int i = tid + j * nthreads;
        if (i < nt) table[i] = tv[j];
      }
    } else {
      for (int i = tid; i < nt; i += nthreads) table[i] = transfer[i];
    }
    __syncthreads();
  }
  const int entries = (L + 1) * GMAX * GMAX;
  const int l = degree_of(r / 4);
  const int n = r >> 2, k = r & 3;
  for (int it = 0; it < RUN; ++it) {
    const int c0 = run0 + it * CELLS;
    if (c0 >= N) break;  // uniform over the workgroup
    float pn[GMAX];
    if (RUN > 1 && it + 1 < RUN) load_phi(min(c0 + CELLS, N - 1), pn);
    if (it > 0) __syncthreads();  // the previous group's contraction has read its matrices
    // per-cell transfer matrices, entries (l, gi, go) with compile-time strides
    if constexpr (EXACT && FACTORED) {
      // densities of the group's cells once per thread (uniform loads), then every entry as a material sum
      float rc[CELLS][kMaxMaterials];
#pragma unroll
      for (int mc = 0; mc < CELLS; ++mc)
#pragma unroll
        for (int m = 0; m < kMaxMaterials; ++m)
          rc[mc][m] = (m < M && c0 + mc < N) ? rho[(int64_t)m * N + c0 + mc] : 0.0f;
      for (int e = tid; e < entries; e += nthreads) {
        const int le = e / (GMAX * GMAX);
        const int gi = (e / GMAX) % GMAX;
        const int go = e % GMAX;
        const float* tm = table + (gi * GMAX + go) * (Lt + 1) + le;
        float m_val[CELLS];
#pragma unroll
        for (int mc = 0; mc < CELLS; ++mc) m_val[mc] = 0.0f;
        if (gi <= go) {
          for (int m = 0; m < M; ++m) {
            const float t = tm[m * GMAX * GMAX * (Lt + 1)];
#pragma unroll
            for (int mc = 0; mc < CELLS; ++mc) {
              // the densities of material m (a register array indexed by a loop variable: select it)
              float rr = rc[mc][0];
#pragma unroll
              for (int q = 1; q < kMaxMaterials; ++q) rr = m == q ? rc[mc][q] : rr;
              m_val[mc] = fmaf(rr, t, m_val[mc]);
            }
          }
        }
#pragma unroll
        for (int mc = 0; mc < CELLS; ++mc) mu[mc * per_cell + le * DS + gi * GMAX + go] = m_val[mc];
      }
    } else {
      for (int mc = 0; mc < CELLS; ++mc) {
        const int mcell_index = c0 + mc;
        for (int e = tid; e < entries; e += nthreads) {
          const int le = e / (GMAX * GMAX);
          const int gi = (e / GMAX) % GMAX;
          const int go = e % GMAX;
          float m_val = 0.0f;
          if (mcell_index < N && gi < G && go < G && !(lower_only && gi > go)) {
            if (FACTORED) {
#pragma unroll
              for (int m = 0; m < kMaxMaterials; ++m)
                if (m < M)
                  m_val = fmaf(rho[(int64_t)m * N + mcell_index], table[((m * G + gi) * G + go) * (Lt + 1) + le],
                               m_val);
            } else {
              m_val = transfer[(((int64_t)gi * G + go) * (Lt + 1) + le) * N + mcell_index];
            }
          }
          mu[mc * per_cell + le * DS + gi * GMAX + go] = m_val;
        }
      }
    }
    __syncthreads();
    const int cell = c0 + cc;
    if (cc < CELLS && cell < N) {
      const int64_t off = (int64_t)v * view_stride + (int64_t)cell * elems + r;
      const float* mcell = mu + cc * per_cell + l * DS;
#pragma unroll
      for (int go = 0; go < GMAX; ++go) {
        if (!EXACT && go >= G) break;
        float acc = 0.0f;
#pragma unroll
        for (int gi = 0; gi < GMAX; ++gi) {
          if (EXACT) {
            if (gi <= go) acc = fmaf(mcell[gi * GMAX + go], p[gi], acc);
          } else if (gi < G && !(lower_only && gi > go)) {
            acc = fmaf(mcell[gi * GMAX + go], p[gi], acc);
          }
        }
        if (MIXED) {
          const int64_t row = ((((int64_t)v * G + go) * N + cell) * 4 + k);  // (view, group, cell, component)
          if (n < 4) {
            T[row * 4 + n] = acc;
          } else {
            T16[row * kh + (n - 4)] = bf16_rne(acc);
            if (n - 4 < kh - (nsh - 4)) T16[row * kh + (nsh - 4) + (n - 4)] = 0;  // padding of the row
          }
        } else {
          T[off + go * group_stride] = acc;
        }
        // the cell means again, moment-major [V][nsh][G][N]: the operand of the image's multiple-collision expansion
        if (means && k == 0) {
          if (RUN > 1)
            mstage[(n * GMAX + go) * RC + it * CELLS + cc] = acc;
          else
            means[(((int64_t)v * nsh + n) * G + go) * N + cell] = acc;
        }
      }
    }
    if (RUN > 1) {
#pragma unroll
      for (int gi = 0; gi < GMAX; ++gi) p[gi] = pn[gi];
    }
  }
  if (RUN > 1 && means) {
    __syncthreads();
    const int ncells = min(RC, N - run0);  // cells of the run inside the grid
    for (int row = tid; row < nsh * G; row += nthreads) {
      const int n = row / G, go = row - n * G;
      const float* src = mstage + (n * GMAX + go) * RC;
      float* dst = means + (((int64_t)v * nsh + n) * G + go) * N + run0;
      if (ncells == RC && (N & 3) == 0 && (reinterpret_cast<uintptr_t>(means) & 15) == 0) {
#pragma unroll
        for (int j = 0; j < RC; j += 4)
          *reinterpret_cast<float4*>(dst + j) = float4{src[j], src[j + 1], src[j + 2], src[j + 3]};
      } else {
        for (int j = 0; j < ncells; ++j) dst[j] = src[j];
      }
    }
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

template <bool FACTORED, bool MIXED = false>
int launch(const float* transfer, const float* rho, int M, int Lt, const float* phi, float* T, int G, int nsh,
           int64_t N, int V, int lower_only, float* means, void* stream, uint16_t* T16 = nullptr,
           const __half* qs = nullptr, const float* qscale = nullptr) {
  int L = 0;
  if (int rc = check(G, nsh, Lt, N, V, L)) return rc;
  const int gmax = G <= 8 ? 8 : (G <= 16 ? 16 : 32);
  // 4 cells per workgroup (their transfer matrices and the material table in LDS, one table staging per 4 cells) up
  // to 16 groups; 2 at 32 groups, whose matrices are 4x larger
  const int cells = gmax <= 16 ? 4 : 2;
  const size_t per_cell = (size_t)(L + 1) * (gmax * gmax + 1) + 1;
  // the exact form: G fills the kernel's group width and only downscatter enters (the view step's transfer); it
  // takes kExactRun groups of cells per workgroup and stages their cell means
  const bool exact = G == gmax && lower_only != 0;
  const int run = exact ? kExactRun : 1;
  const size_t lds = sizeof(float) * ((size_t)cells * per_cell + (FACTORED ? (size_t)M * G * G * (Lt + 1) : 0) +
                                      (run > 1 ? (size_t)nsh * gmax * cells * run : 0));
  if (lds > 80 * 1024) return fail(CITRUS_ERR_UNSUPPORTED, "transfer tables exceed 80 KiB of LDS");
  const int threads = cells * nsh * 4;
  const int64_t blocks = (int64_t)V * ((N + cells * run - 1) / (cells * run));
  if (!grid_fits(blocks, threads)) return fail(CITRUS_ERR_ARG, "too many cells for one launch (2^32 work-items)");
  const int kh = MIXED ? 16 * ((nsh - 4 + 15) / 16) : 0;
  if (MIXED && (nsh < 16 || kh - (nsh - 4) > nsh - 4))
    return fail(CITRUS_ERR_UNSUPPORTED, "the mixed-precision operand is built for P3 and above");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)blocks), block((unsigned)threads);
  auto go = [&](auto gm) {
    constexpr int GM = decltype(gm)::value;
    constexpr int CELLS = GM <= 16 ? 4 : 2;
    if (qs != nullptr) {  // a stored accumulator (the factored float32 form only)
      if constexpr (FACTORED && !MIXED) {
        if (exact)
          hipLaunchKernelGGL((scatter_moments_kernel<GM, CELLS, true, false, true, kExactRun, true>), grid, block,
                             lds, s, transfer, rho, M, Lt, phi, T, G, nsh, L, (int)N, V, lower_only, means, T16, kh,
                             qs, qscale);
        else
          hipLaunchKernelGGL((scatter_moments_kernel<GM, CELLS, true, false, false, 1, true>), grid, block, lds, s,
                             transfer, rho, M, Lt, phi, T, G, nsh, L, (int)N, V, lower_only, means, T16, kh, qs,
                             qscale);
      }
      return;
    }
    if (exact)
      hipLaunchKernelGGL((scatter_moments_kernel<GM, CELLS, FACTORED, MIXED, true, kExactRun, false>), grid, block,
                         lds, s, transfer, rho, M, Lt, phi, T, G, nsh, L, (int)N, V, lower_only, means, T16, kh,
                         nullptr, nullptr);
    else
      hipLaunchKernelGGL((scatter_moments_kernel<GM, CELLS, FACTORED, MIXED, false, 1, false>), grid, block, lds, s,
                         transfer, rho, M, Lt, phi, T, G, nsh, L, (int)N, V, lower_only, means, T16, kh, nullptr,
                         nullptr);
  };
  if (G <= 8)
    go(std::integral_constant<int, 8>{});
  else if (G <= 16)
    go(std::integral_constant<int, 16>{});
  else
    go(std::integral_constant<int, 32>{});
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("scatter moments: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

// ---- mixed-precision operand of the source contraction ----------------------------------------------------------
// T [V][G][N][nsh][4] float32 -> T_low [V][G][N][4][4] float32 (degrees l <= 1, component-major) and
// T16 [V][G][N][4][kh] bfloat16 (degrees l >= 2, round to nearest even, zero padded to kh = nsh - 4 rounded up to 16).
// One thread per (view, group, cell, component, slot), slot < 4 the float32 part.

__global__ void mixed_operand_kernel(const float* __restrict__ T, int nsh, int kh, int64_t rows,
                                     float* __restrict__ T_low, uint16_t* __restrict__ T16) {
  const int width = 4 + kh;
  const int64_t total = rows * width;  // 4.8e9 at 8 views of 64^3 x 16 groups x P5: a stride loop
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / width;  // (view, group, cell) * 4 + component
    const int slot = (int)(i - row * width);
    const int comp = (int)(row & 3);
    const int64_t cell = row >> 2;
    if (slot < 4) {
      T_low[row * 4 + slot] = T[(cell * nsh + slot) * 4 + comp];
    } else {
      const int n = slot;
      T16[row * kh + (slot - 4)] = n < nsh ? bf16_rne(T[(cell * nsh + n) * 4 + comp]) : (uint16_t)0;
    }
  }
}

}  // namespace
}  // namespace citrus

extern "C" int citrus_bf16_operand(const float* T, int G, int nsh, int64_t N, int V, float* T_low, uint16_t* T16,
                                   void* stream) {
  using namespace citrus;
  if (!T || !T16 || !T_low) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (G < 1 || V < 1 || N < 1) return fail(CITRUS_ERR_ARG, "bad sizes");
  if (nsh != 16 && nsh != 36) return fail(CITRUS_ERR_UNSUPPORTED, "the mixed operand is built for P3 and P5");
  const int kh = 16 * ((nsh - 4 + 15) / 16);
  const int64_t rows = (int64_t)V * G * N * 4;
  const int64_t total = rows * (4 + kh);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(mixed_operand_kernel, dim3(stride_grid(total, 256)), dim3(256), 0, s, T, nsh, kh, rows, T_low,
                     T16);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("mixed operand: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

extern "C" int citrus_scatter_moments_mixed(const float* rho, int num_materials, const float* sigma, int Lt,
                                            int downscatter_only, const float* phi, float* T_low, void* T_bf16, int G,
                                            int nsh, int64_t N, int V, float* cell_means, void* stream) {
  using namespace citrus;
  if (!rho || !sigma || !phi || !T_low || !T_bf16) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (num_materials < 1 || num_materials > kMaxMaterials) return fail(CITRUS_ERR_ARG, "1..8 materials supported");
  if (nsh != 16 && nsh != 36) return fail(CITRUS_ERR_UNSUPPORTED, "the mixed operand is built for P3 and P5");
  return launch<true, true>(sigma, rho, num_materials, Lt, phi, T_low, G, nsh, N, V, downscatter_only ? 1 : 0,
                            cell_means, stream, static_cast<uint16_t*>(T_bf16));
}

extern "C" int citrus_scatter_moments_dense(const float* transfer, int Lt, const float* phi, float* T, int G, int nsh,
                                            int64_t N, int V, void* stream) {
  using namespace citrus;
  if (!transfer || !phi || !T) return fail(CITRUS_ERR_ARG, "NULL buffer");
  return launch<false>(transfer, nullptr, 0, Lt, phi, T, G, nsh, N, V, 0, nullptr, stream);
}

extern "C" int citrus_scatter_moments_stored(const float* rho, int num_materials, const float* sigma, int Lt,
                                             int downscatter_only, const void* q, const float* scale, float* T, int G,
                                             int nsh, int64_t N, int V, void* stream) {
  using namespace citrus;
  if (!rho || !sigma || !q || !scale || !T) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (num_materials < 1 || num_materials > kMaxMaterials) return fail(CITRUS_ERR_ARG, "1..8 materials supported");
  return launch<true>(sigma, rho, num_materials, Lt, nullptr, T, G, nsh, N, V, downscatter_only ? 1 : 0, nullptr,
                      stream, nullptr, static_cast<const __half*>(q), scale);
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
