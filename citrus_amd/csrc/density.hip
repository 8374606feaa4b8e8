// Interpolating density model for gfx950: the densities of the reference materials that reproduce an attenuation
// volume X (citrus/utilities.py:710-753, applied by DensityModel, citrus/citrus.py:466-509).
//
// Per voxel, with x = max(X, 0) and the materials in increasing nominal attenuation mu_s = rho_s^nom m_s:
//   j = #{s : mu_s <= x};  inside the bracket (1 <= j <= M-1): slots a = j-1, b = j, f = (x - mu_a) / (mu_b - mu_a),
//   rho_a = (1 - f) rho_a^nom, rho_b = f rho_b^nom;  outside: the first or last slot alone with rho = x / m_a.
// The arithmetic is the reference's, in float32 and in its order (no contraction), so the densities match it bit for
// bit. One thread per voxel, the M <= 8 slot parameters in registers; output [M][n] in library material order.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "citrus_hip.h"
#include "common.h"

#pragma clang fp contract(off)

namespace citrus {
namespace {

constexpr int kMaxMaterials = 8;

struct DensityParams {
  int M;
  int material[kMaxMaterials];  // library index of each slot (bracketing order)
  float mu[kMaxMaterials];      // nominal attenuation of each slot, 1/cm
  float nominal[kMaxMaterials]; // nominal density, g/cm^3
  float mass[kMaxMaterials];    // mass attenuation at the model's energy, cm^2/g
};

__global__ void densities_kernel(const float* __restrict__ X, int64_t n, DensityParams p, float* __restrict__ rho) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = fmaxf(X[i], 0.0f);
  int j = 0;
#pragma unroll
  for (int s = 0; s < kMaxMaterials; ++s) j += (s < p.M && p.mu[s] <= x) ? 1 : 0;
  const int a = j - 1 < 0 ? 0 : (j - 1 > p.M - 1 ? p.M - 1 : j - 1);
  const int b = j > p.M - 1 ? p.M - 1 : j;
  const bool inside = j >= 1 && j <= p.M - 1;
  float mu_a = 0.f, mu_b = 0.f, nom_a = 0.f, nom_b = 0.f, m_a = 1.f;
  int mat_a = 0, mat_b = 0;
#pragma unroll
  for (int s = 0; s < kMaxMaterials; ++s) {
    if (s == a) { mu_a = p.mu[s]; nom_a = p.nominal[s]; m_a = p.mass[s]; mat_a = p.material[s]; }
    if (s == b) { mu_b = p.mu[s]; nom_b = p.nominal[s]; mat_b = p.material[s]; }
  }
  float rho_a, rho_b;
  if (inside) {
    const float span = mu_b - mu_a;
    const float f = (x - mu_a) / span;
    rho_a = (1.0f - f) * nom_a;
    rho_b = f * nom_b;
  } else {
    rho_a = x / m_a;
    rho_b = 0.0f;
  }
#pragma unroll
  for (int m = 0; m < kMaxMaterials; ++m) {
    if (m >= p.M) break;
    // the reference adds rho_a then rho_b into zeros (scatter_add); a == b only outside, where rho_b is 0
    float v = 0.0f;
    if (m == mat_a) v = v + rho_a;
    if (m == mat_b) v = v + rho_b;
    rho[(int64_t)m * n + i] = v;
  }
}

}  // namespace
}  // namespace citrus

extern "C" int citrus_attenuation_to_densities(const float* X, int64_t n, int M, const int* material,
                                               const float* mu_nominal, const float* nominal_density,
                                               const float* mass_attenuation, float* rho, void* stream) {
  using namespace citrus;
  if (!X || !rho || !material || !mu_nominal || !nominal_density || !mass_attenuation)
    return fail(CITRUS_ERR_ARG, "NULL buffer or parameter");
  if (M < 1 || M > kMaxMaterials) return fail(CITRUS_ERR_ARG, "1..8 materials supported");
  if (n < 0) return fail(CITRUS_ERR_ARG, "n must be >= 0");
  DensityParams p{};
  p.M = M;
  for (int s = 0; s < M; ++s) {
    if (material[s] < 0 || material[s] >= M) return fail(CITRUS_ERR_ARG, "material index outside 0..M-1");
    if (s > 0 && !(mu_nominal[s] > mu_nominal[s - 1]))
      return fail(CITRUS_ERR_ARG, "nominal attenuations must increase along the slots");
    p.material[s] = material[s];
    p.mu[s] = mu_nominal[s];
    p.nominal[s] = nominal_density[s];
    p.mass[s] = mass_attenuation[s];
  }
  if (n == 0) return CITRUS_OK;
  if (!grid_fits((n + 255) / 256, 256)) return fail(CITRUS_ERR_ARG, "too many voxels for one launch (2^32)");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(densities_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, X, n, p, rho);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("density model launch: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

// ---- conservative resampling between grids ------------------------------------------------------------------------
// out[ch][j] = sum_i field[ch][i] |V_i n V_j| / |V_j| over source voxels i and target cells j (geometry.py:304). The
// overlap share of a target cell factors over the axes; along each axis the source cells a target cell overlaps are
// contiguous, given as (first index, count, shares padded to `width`). One thread per (channel, target cell) sums its
// count_r x count_c x count_s source voxels in a fixed order (slices innermost: contiguous reads), in float32.

namespace citrus {
namespace {

struct Band {
  const int* first;
  const int* count;
  const float* share;
  int width;
};

__global__ void resample_kernel(const float* __restrict__ src, int channels, int R, int C, int S, int Rt, int Ct,
                                int St, Band br, Band bc, Band bs, float* __restrict__ out) {
  const int64_t cells = (int64_t)Rt * Ct * St;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells * channels) return;
  const int ch = (int)(t / cells);
  const int64_t j = t - (int64_t)ch * cells;
  const int r = (int)(j / ((int64_t)Ct * St));
  const int c = (int)((j / St) % Ct);
  const int s = (int)(j % St);
  const float* f = src + (int64_t)ch * R * C * S;
  const int r0 = br.first[r], nr = min(br.count[r], br.width);
  const int c0 = bc.first[c], nc = min(bc.count[c], bc.width);
  const int s0 = bs.first[s], ns = min(bs.count[s], bs.width);
  float acc = 0.0f;
  for (int a = 0; a < nr; ++a) {
    const int ri = r0 + a;
    if (ri < 0 || ri >= R) continue;
    const float wr = br.share[(int64_t)r * br.width + a];
    for (int b = 0; b < nc; ++b) {
      const int ci = c0 + b;
      if (ci < 0 || ci >= C) continue;
      const float wrc = wr * bc.share[(int64_t)c * bc.width + b];
      const float* row = f + ((int64_t)ri * C + ci) * S;
      float line = 0.0f;
      for (int d = 0; d < ns; ++d) {
        const int si = s0 + d;
        if (si < 0 || si >= S) continue;
        line += row[si] * bs.share[(int64_t)s * bs.width + d];
      }
      acc += wrc * line;
    }
  }
  out[t] = acc;
}

}  // namespace
}  // namespace citrus

extern "C" int citrus_resample_conservative(const float* field, int channels, int R, int C, int S, int Rt, int Ct,
                                            int St, const int* row_first, const int* row_count,
                                            const float* row_share, int row_width, const int* col_first,
                                            const int* col_count, const float* col_share, int col_width,
                                            const int* slice_first, const int* slice_count,
                                            const float* slice_share, int slice_width, float* out, void* stream) {
  using namespace citrus;
  if (!field || !out || !row_first || !row_count || !row_share || !col_first || !col_count || !col_share ||
      !slice_first || !slice_count || !slice_share)
    return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (channels < 0 || R < 1 || C < 1 || S < 1 || Rt < 1 || Ct < 1 || St < 1)
    return fail(CITRUS_ERR_ARG, "grid sizes must be >= 1 and channels >= 0");
  if (row_width < 1 || col_width < 1 || slice_width < 1) return fail(CITRUS_ERR_ARG, "band widths must be >= 1");
  const int64_t total = (int64_t)Rt * Ct * St * channels;
  if (total == 0) return CITRUS_OK;
  if (!grid_fits((total + 255) / 256, 256)) return fail(CITRUS_ERR_ARG, "too many output values for one launch (2^32)");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const Band br{row_first, row_count, row_share, row_width}, bc{col_first, col_count, col_share, col_width},
      bs{slice_first, slice_count, slice_share, slice_width};
  hipLaunchKernelGGL(resample_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, field, channels, R, C,
                     S, Rt, Ct, St, br, bc, bs, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("resample launch: ") + hipGetErrorString(e));
  return CITRUS_OK;
}
