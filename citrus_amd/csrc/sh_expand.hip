// Spherical-harmonic expansion of cell moments at a set of directions (gfx950):
//   C[k][j] = sum_n Y[k][n] B[n][j]        (k: directions, n: SH moments, j: (group, cell) columns)
// the multiple-collision source of the collapsed-cone image at a chunk of cone directions, from the cell means of
// the scattering-source moments (citrus/scatter_solver/scatter_image.py:1225,1246, `torch.mm(harmonics, moments_by_cell)`).
//
// K is small (64 per chunk), nsh <= 64 and the column count is G x cells (4.2 M at 64^3 x 16): the product reads B
// once and writes C once, 36 multiply-adds per output at P5. On vector FMAs that is issue-bound (3 TB/s); on the
// matrix cores it is HBM-bound. The product runs as C^T = B^T Y^T on v_mfma_f32_16x16x4f32: the A operand is a tile
// of 16 columns by 4 moments of B (lane l: column l%16, moment l/16, one dword load per lane, 64 contiguous bytes per
// moment row), the B operand 4 moments by 16 directions of Y (from LDS), and each lane's four results are four
// consecutive columns of one direction, written as one 16-byte store. A wave holds the A tiles of its 64 columns in
// registers (all moments) and loops over the direction tiles. Sums run in float32 four moments per step.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "citrus_hip.h"
#include "common.h"

namespace citrus {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr int kColTiles = 4;                            // 16-column tiles per wave
constexpr int kBlockCols = kWaves * kColTiles * 16;     // 256 columns per workgroup
constexpr int kMaxLdsBytes = 64 * 1024;

__host__ __device__ constexpr int ksteps(int nsh) { return (nsh + 3) / 4; }
__host__ __device__ constexpr int ypitch(int nsh) { return 4 * ksteps(nsh) + 1; }  // padded row: no bank repeats

template <int NSH>
__global__ __launch_bounds__(64 * kWaves) void sh_expand_kernel(const float* __restrict__ Y, int K,
                                                                const float* __restrict__ B, int64_t ncol,
                                                                float* __restrict__ C, int vec_store) {
  constexpr int KS = ksteps(NSH), YP = ypitch(NSH);
  extern __shared__ float ys[];  // Y [DT*16][YP], zero beyond K directions and NSH moments
  const int DT = (K + 15) / 16;
  for (int i = threadIdx.x; i < DT * 16 * YP; i += 64 * kWaves) {
    const int d = i / YP, n = i % YP;
    ys[i] = (d < K && n < NSH) ? Y[d * NSH + n] : 0.0f;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 15, q = lane >> 4;
  const int64_t col0 = (int64_t)blockIdx.x * kBlockCols + wave * (kColTiles * 16);
  // A tiles: column col0 + 16 t + j, moment 4 s + q
  float a[kColTiles][KS];
#pragma unroll
  for (int t = 0; t < kColTiles; ++t) {
    const int64_t col = col0 + 16 * t + j;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int n = 4 * s + q;
      a[t][s] = (n < NSH && col < ncol) ? B[(int64_t)n * ncol + col] : 0.0f;
    }
  }
  __syncthreads();
  for (int dt = 0; dt < DT; ++dt) {
    float y[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) y[s] = ys[(dt * 16 + j) * YP + 4 * s + q];
    const int dir = dt * 16 + j;
#pragma unroll
    for (int t = 0; t < kColTiles; ++t) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < KS; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][s], y[s], acc, 0, 0, 0);
      // acc[r] = C[dir][col0 + 16 t + 4 q + r]
      const int64_t c = col0 + 16 * t + 4 * q;
      if (dir < K) {
        float* dst = C + (int64_t)dir * ncol + c;
        if (vec_store && c + 3 < ncol) {
          *reinterpret_cast<f4*>(dst) = acc;
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (c + r < ncol) dst[r] = acc[r];
        }
      }
    }
  }
}

template <int NSH>
size_t lds_bytes(int K) {
  return sizeof(float) * (size_t)((K + 15) / 16) * 16 * ypitch(NSH);
}

template <int NSH>
hipError_t launch(const float* Y, int K, const float* B, int64_t ncol, float* C, hipStream_t s) {
  const int vec = (ncol % 4 == 0 && reinterpret_cast<uintptr_t>(C) % 16 == 0) ? 1 : 0;
  hipLaunchKernelGGL(sh_expand_kernel<NSH>, dim3((unsigned)((ncol + kBlockCols - 1) / kBlockCols)),
                     dim3(64 * kWaves), lds_bytes<NSH>(K), s, Y, K, B, ncol, C, vec);
  return hipGetLastError();
}

// ---- multiple-collision emission of a direction chunk, collapsed over the groups (one quadrature point per cell) --
// The signal-only image needs of the multiple-collision source only  sum_g w[k][g] exp(-tau_g) S_MC,g  (and the
// unweighted sum, for the footprint's emission-weighted height) per (direction, cell): the expansion, the attenuation
// and the group sum run in one pass and write 2 floats per (direction, cell) instead of G. A workgroup holds 64 cells
// (16 per wave) and 64 directions; per group g the wave expands its 16 cells at the 64 directions on the matrix cores
// (as above: lane l's results are cells c4 .. c4+3 of one direction), multiplies by exp(-tau_g) of those cells and
// accumulates, so the moments of a cell are read once per 64 directions.

constexpr int kDirTiles = 4;    // 64 directions per workgroup
constexpr int kMaxGroups = 32;

template <int NSH, int TC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void multiple_emission_kernel(const float* __restrict__ Y, int K,
                                                                const float* __restrict__ means, int G, int64_t N,
                                                                const float* __restrict__ tau,
                                                                const float* __restrict__ tau_mix,
                                                                const float* __restrict__ signal_weight,
                                                                float* __restrict__ out, int vec) {
  constexpr int KS = ksteps(NSH), YP = ypitch(NSH);
  __shared__ float ys[16 * kDirTiles * YP];        // Y of the 64 directions, zero beyond K and NSH
  __shared__ float ws[16 * kDirTiles * kMaxGroups];  // signal weights [direction][group]
  __shared__ float mixs[kMaxGroups * (TC > 0 ? TC : 1)];
  const int k0 = blockIdx.y * 16 * kDirTiles;
  const int KB = K - k0 < 16 * kDirTiles ? K - k0 : 16 * kDirTiles;
  for (int i = threadIdx.x; i < 16 * kDirTiles * YP; i += 256) {
    const int d = i / YP, n = i % YP;
    ys[i] = (d < KB && n < NSH) ? Y[(int64_t)(k0 + d) * NSH + n] : 0.0f;
  }
  for (int i = threadIdx.x; i < 16 * kDirTiles * kMaxGroups; i += 256) {
    const int d = i / kMaxGroups, g = i % kMaxGroups;
    ws[i] = (d < KB && g < G) ? signal_weight[(int64_t)(k0 + d) * G + g] : 0.0f;
  }
  if constexpr (TC > 0)
    for (int i = threadIdx.x; i < G * TC; i += 256) mixs[i] = tau_mix[i];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 15, q = lane >> 4;
  const int64_t cell0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
  const int64_t ca = cell0 + j;      // the lane's cell of the A operand
  const int64_t c4 = cell0 + 4 * q;  // the lane's result cells c4 .. c4 + 3
  auto load4 = [&](const float* row) -> f4 {  // elements c4 .. c4 + 3 of a row of N (0 beyond N)
    if (vec && c4 + 3 < N) return *reinterpret_cast<const f4*>(row + c4);
    f4 v;
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = c4 + r < N ? row[c4 + r] : 0.0f;
    return v;
  };
  auto store4 = [&](float* row, f4 v) {
    if (vec && c4 + 3 < N) {
      *reinterpret_cast<f4*>(row + c4) = v;
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (c4 + r < N) row[c4 + r] = v[r];
    }
  };
  // traced channels of the lane's cells at its direction of each tile
  f4 tc[kDirTiles][TC > 0 ? TC : 1];
  if constexpr (TC > 0) {
#pragma unroll
    for (int dt = 0; dt < kDirTiles; ++dt) {
      const int dir = k0 + dt * 16 + j;
#pragma unroll
      for (int c = 0; c < TC; ++c)
        tc[dt][c] = dir < K ? load4(tau + ((int64_t)dir * TC + c) * N) : f4{0.f, 0.f, 0.f, 0.f};
    }
  }
  // A-operand loads: wave-uniform row base (moment 4 s, group g) plus the lane's 32-bit byte offset (moment 4 s + q,
  // cell ca), clamped into the array. A clamped cell only feeds result row ca, which is not stored; a clamped moment
  // (nsh not a multiple of 4) meets a zero harmonic but is zeroed as well, in case the data holds an infinity.
  const int64_t cac = ca < N ? ca : N - 1;
  uint32_t off[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int n = 4 * s + q < NSH ? 4 * s + q : NSH - 1;
    off[s] = (uint32_t)((((int64_t)(n - 4 * s) * G) * N + cac) * (int64_t)sizeof(float));
  }
  auto load_a = [&](int g, float (&dst)[KS]) {
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const char* row = reinterpret_cast<const char*>(means + ((int64_t)4 * s * G + g) * N);
      const float v = *reinterpret_cast<const float*>(row + off[s]);
      if constexpr (NSH % 4 == 0)
        dst[s] = v;
      else
        dst[s] = 4 * s + q < NSH ? v : 0.0f;
    }
  };
  float a[KS], an[KS];
  load_a(0, a);
  f4 sw[kDirTiles], ss[kDirTiles];
#pragma unroll
  for (int dt = 0; dt < kDirTiles; ++dt) sw[dt] = ss[dt] = f4{0.f, 0.f, 0.f, 0.f};
  __syncthreads();
  for (int g = 0; g < G; ++g) {
    // the next group's moments are in flight during this group's products
    load_a(g + 1 < G ? g + 1 : g, an);
#pragma unroll
    for (int dt = 0; dt < kDirTiles; ++dt) {
      const int dl = dt * 16 + j;
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < KS; ++s)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], ys[dl * YP + 4 * s + q], acc, 0, 0, 0);
      // acc[r] = S_MC,g at direction k0 + dl, cell c4 + r
      f4 t;
      if constexpr (TC > 0) {
        t = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < TC; ++c) {
          const float m = mixs[g * TC + c];
#pragma unroll
          for (int r = 0; r < 4; ++r) t[r] = fmaf(m, tc[dt][c][r], t[r]);
        }
      } else {
        const int dir = k0 + dl;
        t = dir < K ? load4(tau + ((int64_t)dir * G + g) * N) : f4{0.f, 0.f, 0.f, 0.f};
      }
      const float w = ws[dl * kMaxGroups + g];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float e = attenuation(t[r]);
        if (TC > 0 && tc[dt][0][r] == INFINITY) e = 0.0f;  // no path to the panel
        const float v = acc[r] * e;
        ss[dt][r] += v;
        sw[dt][r] = fmaf(w, v, sw[dt][r]);
      }
    }
#pragma unroll
    for (int s = 0; s < KS; ++s) a[s] = an[s];
  }
#pragma unroll
  for (int dt = 0; dt < kDirTiles; ++dt) {
    const int dir = k0 + dt * 16 + j;
    if (dir < K) {
      store4(out + (int64_t)dir * 2 * N, sw[dt]);
      store4(out + ((int64_t)dir * 2 + 1) * N, ss[dt]);
    }
  }
}

template <int NSH, int TC>
hipError_t launch_emission(const float* Y, int K, const float* means, int G, int64_t N, const float* tau,
                           const float* tau_mix, const float* signal_weight, float* out, hipStream_t s) {
  const int vec = (N % 4 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(tau) % 16 == 0) ? 1 : 0;
  const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((K + 16 * kDirTiles - 1) / (16 * kDirTiles)));
  hipLaunchKernelGGL((multiple_emission_kernel<NSH, TC>), grid, dim3(256), 0, s, Y, K, means, G, N, tau, tau_mix,
                     signal_weight, out, vec);
  return hipGetLastError();
}

template <int TC>
hipError_t dispatch_emission(int nsh, const float* Y, int K, const float* means, int G, int64_t N, const float* tau,
                             const float* tau_mix, const float* signal_weight, float* out, hipStream_t s) {
  switch (nsh) {
    case 1: return launch_emission<1, TC>(Y, K, means, G, N, tau, tau_mix, signal_weight, out, s);
    case 4: return launch_emission<4, TC>(Y, K, means, G, N, tau, tau_mix, signal_weight, out, s);
    case 9: return launch_emission<9, TC>(Y, K, means, G, N, tau, tau_mix, signal_weight, out, s);
    case 16: return launch_emission<16, TC>(Y, K, means, G, N, tau, tau_mix, signal_weight, out, s);
    case 25: return launch_emission<25, TC>(Y, K, means, G, N, tau, tau_mix, signal_weight, out, s);
    case 36: return launch_emission<36, TC>(Y, K, means, G, N, tau, tau_mix, signal_weight, out, s);
    case 49: return launch_emission<49, TC>(Y, K, means, G, N, tau, tau_mix, signal_weight, out, s);
    default: return launch_emission<64, TC>(Y, K, means, G, N, tau, tau_mix, signal_weight, out, s);
  }
}

}  // namespace
}  // namespace citrus

extern "C" int citrus_multiple_emission(const float* Y, int K, int nsh, const float* means, int G, int64_t N,
                                        const float* tau, int tau_channels, const float* tau_mix,
                                        const float* signal_weight, float* out, void* stream) {
  using namespace citrus;
  if (!Y || !means || !tau || !signal_weight || !out) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (K < 1 || N < 1) return fail(CITRUS_ERR_ARG, "K and N must be >= 1");
  if (G < 1 || G > kMaxGroups) return fail(CITRUS_ERR_ARG, "1..32 groups supported");
  if (nsh != 1 && nsh != 4 && nsh != 9 && nsh != 16 && nsh != 25 && nsh != 36 && nsh != 49 && nsh != 64)
    return fail(CITRUS_ERR_UNSUPPORTED, "nsh must be (L+1)^2 with L <= 7");
  if (tau_mix && (tau_channels < 1 || tau_channels > 2))
    return fail(CITRUS_ERR_UNSUPPORTED, "1 or 2 traced channels with a mix table");
  if (!grid_fits((N + 63) / 64, 256) || (K + 63) / 64 > 65535) return fail(CITRUS_ERR_ARG, "too many cells or directions");
  // the lanes' moment offsets within four moment rows are 32-bit byte offsets
  if ((3 * (int64_t)G + 1) * N * (int64_t)sizeof(float) > 0xffffffffLL)
    return fail(CITRUS_ERR_UNSUPPORTED, "(3 G + 1) N x 4 bytes must stay below 4 GiB");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipError_t e;
  if (!tau_mix)
    e = dispatch_emission<0>(nsh, Y, K, means, G, N, tau, nullptr, signal_weight, out, s);
  else if (tau_channels == 1)
    e = dispatch_emission<1>(nsh, Y, K, means, G, N, tau, tau_mix, signal_weight, out, s);
  else
    e = dispatch_emission<2>(nsh, Y, K, means, G, N, tau, tau_mix, signal_weight, out, s);
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("multiple emission: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

extern "C" int citrus_sh_expand(const float* Y, int K, int nsh, const float* B, int64_t ncol, float* C,
                                void* stream) {
  using namespace citrus;
  if (!Y || !B || !C) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (K < 1 || ncol < 1) return fail(CITRUS_ERR_ARG, "K and ncol must be >= 1");
  if (nsh < 1 || nsh > 64) return fail(CITRUS_ERR_UNSUPPORTED, "nsh must be (L+1)^2 with L <= 7");
  if (sizeof(float) * (size_t)((K + 15) / 16) * 16 * (4 * ((nsh + 3) / 4) + 1) > (size_t)kMaxLdsBytes)
    return fail(CITRUS_ERR_ARG, "too many directions: the directions' harmonics (K x nsh, padded) must fit 64 KB of LDS");
  if (!grid_fits((ncol + kBlockCols - 1) / kBlockCols, 64 * kWaves)) return fail(CITRUS_ERR_ARG, "too many columns");
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
