// First-collision source at the sweep directions, its sampling-correction sums and its LD cell moments (gfx950).
//
//   S_FC,g(r, w) = sum_{g' <= g} sum_m rho_m(r) mu_s,eta^{g'->g}(Omega_UC(r) . w) phi_UC,g'(r)
//
// (citrus/scatter_solver/uncollided.py:200-270, the transfer table interpolated linearly in eta at the position of
// cross_sections.py:905-945, uniform or through an adaptive map). Materials enter per point through up to four slots
// (material index, density), so a point costs what its nonzero materials cost. The table is node-major
// [n_eta][M][G][GP] (GP = 4, 8, 16 or 32, the kernel's group count; entries below the diagonal zeroed by the host): one point's interpolation
// at one eta reads two contiguous node rows as float4, which neighbouring points share in L1.
//
// Three kernels share the evaluation:
//   source   thread per (point, direction) -> out [K][G][P]                               (uncollided.py:200)
//   sums     thread per point, all directions -> out [5][G][P] = I0, I1, I2 (sum_k w_k eta^j S_k, FP32 in direction
//            order) and the exact moments E0, E1 (uncollided.py:357-384, :538-608; uncollided_kernels.py:637)
//   moments  thread per (cell, direction): corrected samples S scale (1 + beta eta) at the f^3 sub-points of a cell,
//            restricted to LD moments mean and mean(S u_a)/mean(u_a^2) (uncollided.py:311-352, :420-433, :464-535),
//            written [K][G][4][cells], cell-major [K][cells][G][4], group-major [G][cells][K][4] or, when f = 1 (the
//            slopes vanish), group-major cell means [G][cells][K].

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>

#include "citrus_hip.h"
#include "block_planes.h"
#include "common.h"

namespace citrus {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int kMaxSlots = 4;
constexpr int kBlock = 256;
constexpr int kMaxTauChannels = 8;

struct FcDev {
  int M, G, GP, n_eta, n_map, n_slots;
  const float* table;
  const float* eta_map;
  const int* slot_material;
  const float* slot_density;
  const float* phi_uc;
  int R, C, S, P;
  double x0, y0, z0, dx, dy, dz;
  double fx, fy, fz;
  int parallel;
  const float* omegas;
  int K;
};

// padded g_out width of the table rows: the kernel's compile-time group count
int group_pad(int G) { return G <= 4 ? 4 : G <= 8 ? 8 : G <= 16 ? 16 : 32; }

FcDev make_dev(const CitrusFcArgs& a) {
  FcDev d{};
  d.M = a.M; d.G = a.G; d.GP = group_pad(a.G); d.n_eta = a.n_eta; d.n_map = a.eta_map ? a.n_map : 0;
  d.n_slots = a.n_slots; d.table = a.table; d.eta_map = a.eta_map; d.slot_material = a.slot_material;
  d.slot_density = a.slot_density; d.phi_uc = a.phi_uc; d.R = a.R; d.C = a.C; d.S = a.S; d.P = a.R * a.C * a.S;
  d.x0 = a.lattice[0]; d.y0 = a.lattice[1]; d.z0 = a.lattice[2];
  d.dx = a.lattice[3]; d.dy = a.lattice[4]; d.dz = a.lattice[5];
  d.fx = a.focal[0]; d.fy = a.focal[1]; d.fz = a.focal[2]; d.parallel = a.parallel;
  d.omegas = a.omegas; d.K = a.K;
  return d;
}

// Omega_UC of point p = (r, c, s): (x - focal)/|x - focal| formed in FP64, used in FP32
__device__ __forceinline__ void primary_direction(const FcDev& d, int r, int c, int s, float& ux, float& uy, float& uz) {
  if (d.parallel) {
    ux = (float)d.fx; uy = (float)d.fy; uz = (float)d.fz;
    return;
  }
  const double px = d.x0 + c * d.dx - d.fx, py = d.y0 + r * d.dy - d.fy, pz = d.z0 + s * d.dz - d.fz;
  const double inv = 1.0 / sqrt(px * px + py * py + pz * pz);
  ux = (float)(px * inv); uy = (float)(py * inv); uz = (float)(pz * inv);
}

// lower node and weight of the upper node (cross_sections.py:905-945), in FP32 like the reference's working dtype
__device__ __forceinline__ void table_position(const FcDev& d, float eta, int& i0, float& w1) {
  eta = fminf(fmaxf(eta, -1.0f), 1.0f);
  const int n = d.n_eta;
  float pos;
  if (d.n_map == 0) {
    pos = (eta + 1.0f) / 2.0f * (float)(n - 1);
  } else {
    const int K = d.n_map;
    const float q = sqrtf(fmaxf((1.0f - eta) * 0.5f, 0.0f));
    const float p = q * (float)(K - 1);
    const float k0f = fminf(fmaxf(floorf(p), 0.0f), (float)(K - 2));
    const float a = fminf(fmaxf(p - k0f, 0.0f), 1.0f);
    const int k0 = (int)k0f;
    const float u0 = d.eta_map[k0];
    pos = (u0 + (d.eta_map[k0 + 1] - u0) * a) * (float)(n - 1);
  }
  int i = (int)floorf(pos);
  i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
  i0 = i;
  w1 = fminf(fmaxf(pos - (float)i, 0.0f), 1.0f);
}

// per-point data kept in registers: phi_UC of every group; the slots (material, density) are read per use (L1)
template <int GMAX>
struct PointData {
  int p;
  float ph[GMAX];
};

template <int GMAX>
__device__ __forceinline__ void load_point(const FcDev& d, int p, PointData<GMAX>& pd) {
  pd.p = p;
#pragma unroll
  for (int g = 0; g < GMAX; ++g) pd.ph[g] = g < d.G ? d.phi_uc[(int64_t)g * d.P + p] : 0.0f;
}

// The source is linear in the table, so the two nodes are contracted separately, acc0 += cv a and acc1 += cv b, and
// interpolated once at the end, acc0 + w1 (acc1 - acc0): two operations per table entry instead of three. Explicit
// fused multiply-adds, so that the per-channel and the all-channel forms round alike.
__device__ __forceinline__ f4 fma4(float c, f4 a, f4 acc) {
  return f4{fmaf(c, a.x, acc.x), fmaf(c, a.y, acc.y), fmaf(c, a.z, acc.z), fmaf(c, a.w, acc.w)};
}

__device__ __forceinline__ f4 interpolate4(f4 acc0, f4 acc1, float w1) {
  return f4{fmaf(w1, acc1.x - acc0.x, acc0.x), fmaf(w1, acc1.y - acc0.y, acc0.y), fmaf(w1, acc1.z - acc0.z, acc0.z),
            fmaf(w1, acc1.w - acc0.w, acc0.w)};
}

// S[go] at one eta; entries go < g' of the table are zero (downscatter only)
template <int GMAX>
__device__ __forceinline__ void eval_source(const FcDev& d, const PointData<GMAX>& pd, float eta, float S[GMAX]) {
  int i0;
  float w1;
  table_position(d, eta, i0, w1);
  constexpr int NCH = GMAX / 4;
  f4 acc0[NCH], acc1[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) acc0[ch] = acc1[ch] = f4{0.f, 0.f, 0.f, 0.f};
  // rows are GMAX wide (the host pads g_out to the kernel's group count); rows g' >= G are clamped and carry phi 0
  const int64_t node = (int64_t)d.M * d.G * GMAX;
  for (int s = 0; s < d.n_slots; ++s) {
    const float rho = d.slot_density[(int64_t)s * d.P + pd.p];
    // a material absent from the point adds exact zeros (the table is finite, and x + 0 == x for the sums, which
    // never hold -0): skip its gathers, which a wave then issues only for the lanes that hold the material -- none
    // outside the body, one slot in most cells
    if (rho == 0.0f) continue;
    const float* t0 = d.table + (int64_t)i0 * node + (int64_t)d.slot_material[(int64_t)s * d.P + pd.p] * d.G * GMAX;
#pragma unroll
    for (int gi = 0; gi < GMAX; ++gi) {
      const float cv = rho * pd.ph[gi];
      const int row = gi < d.G ? gi : d.G - 1;
      const f4* r0 = reinterpret_cast<const f4*>(t0 + row * GMAX);
      const f4* r1 = reinterpret_cast<const f4*>(t0 + node + row * GMAX);
#pragma unroll
      for (int ch = gi / 4; ch < NCH; ++ch) {
        acc0[ch] = fma4(cv, r0[ch], acc0[ch]);
        acc1[ch] = fma4(cv, r1[ch], acc1[ch]);
      }
    }
  }
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    const f4 v = interpolate4(acc0[ch], acc1[ch], w1);
    S[4 * ch + 0] = v.x;
    S[4 * ch + 1] = v.y;
    S[4 * ch + 2] = v.z;
    S[4 * ch + 3] = v.w;
  }
}

template <int GMAX>
__global__ __launch_bounds__(kBlock) void fc_source_kernel(FcDev d, float* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (int64_t)d.P * d.K) return;
  const int p = (int)(idx % d.P);
  const int k = (int)(idx / d.P);
  const int r = p / (d.C * d.S), c = (p / d.S) % d.C, s = p % d.S;
  float ux, uy, uz;
  primary_direction(d, r, c, s, ux, uy, uz);
  PointData<GMAX> pd;
  load_point(d, p, pd);
  const float eta = direction_cosine(ux, uy, uz, d.omegas[3 * k], d.omegas[3 * k + 1], d.omegas[3 * k + 2]);
  float S[GMAX];
  eval_source(d, pd, eta, S);
#pragma unroll
  for (int g = 0; g < GMAX; ++g)
    if (g < d.G) out[((int64_t)k * d.G + g) * d.P + p] = S[g];
}

// The material slots of one point, held in registers for the directions of a point-major loop.
struct PointSlots {
  float rho[kMaxSlots];
  int mat[kMaxSlots];
};

__device__ __forceinline__ void load_slots(const FcDev& d, int p, PointSlots& ps) {
#pragma unroll
  for (int s = 0; s < kMaxSlots; ++s) {
    ps.rho[s] = s < d.n_slots ? d.slot_density[(int64_t)s * d.P + p] : 0.0f;
    ps.mat[s] = s < d.n_slots ? d.slot_material[(int64_t)s * d.P + p] : 0;
  }
}

// The sweep's block order of first-collision cell means (source layout 3 of the persistent sweep): per group
// [blocks of 4^3 cells][8 octants][64 cells of the block][dp], blocks (row block, column block, slice block) in
// row-major order, the cells of octant o's run in o's task order (block_slot4: by diagonal plane from the octant's
// upstream corner, so the cells one plane of the task reads are adjacent), dp = directions per octant rounded up to 4
// (the padding zero), then CITRUS_FC_GROUP_PAD unused floats. One sweep task (block, octant, group) reads one
// contiguous run of 64 dp floats; the pad skews the groups' and views' regions across HBM channels (unpadded they
// all start on the same channel, and the persistent sweep runs the same block of 16 domains of one XCD together).
__host__ __device__ __forceinline__ int64_t fc_sample_group_floats(int R, int C, int S, int dp) {
  return (int64_t)((R + 3) / 4) * ((C + 3) / 4) * ((S + 3) / 4) * 8 * 64 * dp + CITRUS_FC_GROUP_PAD;
}

// offset of octant o's run of block (r/4, c/4, s/4), and the position of cell (r, c, s) in it
__device__ __forceinline__ int64_t fc_sample_run(int C, int S, int dp, int r, int c, int s, int o) {
  const int64_t blk = ((int64_t)(r >> 2) * ((C + 3) / 4) + (c >> 2)) * ((S + 3) / 4) + (s >> 2);
  return ((blk * 8 + o) * 64 + block_slot4(o, c & 3, r & 3, s & 3)) * dp;
}

// S of the four output groups 4 ch .. 4 ch + 3 at one eta, for a wave-uniform channel with the slots in registers:
// eval_source restricted to one channel (the same terms in the same order, so the values are equal bit for bit), the
// table rows of four source groups requested together before they are added.
template <int GMAX>
__device__ __forceinline__ f4 eval_source_channel_batched(const FcDev& d, const PointData<GMAX>& pd,
                                                          const PointSlots& ps, float eta, int ch) {
  int i0;
  float w1;
  table_position(d, eta, i0, w1);
  f4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0;
  const int64_t node = (int64_t)d.M * d.G * GMAX;
#pragma unroll
  for (int s = 0; s < kMaxSlots; ++s) {
    if (s >= d.n_slots) break;
    const float rho = ps.rho[s];
    if (rho == 0.0f) continue;  // as in eval_source
    const float* t0 = d.table + (int64_t)i0 * node + (int64_t)ps.mat[s] * d.G * GMAX;
#pragma unroll
    for (int gb = 0; gb < GMAX / 4; ++gb) {
      if (gb > ch) break;
      f4 x0[4], x1[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int gi = 4 * gb + j;
        const int row = gi < d.G ? gi : d.G - 1;
        x0[j] = reinterpret_cast<const f4*>(t0 + row * GMAX)[ch];
        x1[j] = reinterpret_cast<const f4*>(t0 + node + row * GMAX)[ch];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float cv = rho * pd.ph[4 * gb + j];
        acc0 = fma4(cv, x0[j], acc0);
        acc1 = fma4(cv, x1[j], acc1);
      }
    }
  }
  return interpolate4(acc0, acc1, w1);
}

// Sampling-correction factors of one (point, group) from its sums (first_collision_correction_factors,
// uncollided.py:611-640): in float64 with the host formula's operations in its order and none contracted, so that
// the factors equal the host's float64 result rounded to float32 bit for bit.
__device__ __forceinline__ void correction_factors(float i0, float i1, float i2, float e0, float e1, float& scale,
                                                   float& beta) {
#pragma clang fp contract(off)
  const double I0 = i0, I1 = i1, I2 = i2, E0 = e0, E1 = e1;
  const bool pos = E0 > 0.0;
  const double c = pos ? E1 / E0 : 0.0;
  const double den = I2 - c * I1;
  const bool ok = pos && fabs(den) > 1e-12 * fabs(I0);
  double b = ok ? (c * I0 - I1) / den : 0.0;
  b = b < -1.0 ? -1.0 : (b > 1.0 ? 1.0 : b);  // a clamp that keeps a NaN, as torch's
  const double d0 = I0 + b * I1;
  scale = (float)(d0 > 0.0 ? E0 / d0 : 1.0);
  beta = (float)b;
}

// One thread per (point, output channel of four groups): the channel is uniform per workgroup (the heaviest, which
// gathers the most table rows, first) and each thread runs all K directions for its four groups. Splitting the
// groups over threads keeps the accumulators of a thread at 4 groups, so 4x as many points' gathers are in flight.
// The directions (unit vector and weight) pass through LDS in tiles of one per thread; the point's material slots stay
// in registers. Per group the terms and their order are those of the all-group form.
// d8 > 0: the samples go out in the sweep's block order instead (fc_sample_run), the K = 8 d8 directions octant
// by octant in runs of four, and the threads take the points in that order too (one wave per 4^3 block), so that a
// wave's consecutive run stores fill the block's contiguous 64 x dp floats of one octant while it is in L2; in grid
// order a wave's stores fall into 16 blocks at the stride of a cell, and lines left L2 partly written.
template <int GMAX, bool BLOCKED>
__global__ __launch_bounds__(kBlock) void fc_sums_kernel(FcDev d, const float* __restrict__ weights,
                                                         const float* __restrict__ mu_s_l, int lp1,
                                                         float* __restrict__ out, float* __restrict__ samples,
                                                         int blocks_per_channel, int d8, float* __restrict__ fscale,
                                                         float* __restrict__ fbeta, float* __restrict__ fu) {
  __shared__ f4 sdir[kBlock];  // (omega x, y, z, weight) of the tile's directions
  constexpr int NCH = GMAX / 4;
  const int ch = NCH - 1 - (int)(blockIdx.x / blocks_per_channel);
  const int t = (int)(blockIdx.x % blocks_per_channel) * kBlock + threadIdx.x;
  if (4 * ch >= d.G) return;  // uniform over the workgroup
  int p = t;
  if (BLOCKED) {  // t = 64 block + cell of the block, both row-major
    const int bc = (d.C + 3) / 4, bs = (d.S + 3) / 4, blk = t / 64, lc = t % 64;
    const int r = (blk / (bc * bs)) * 4 + lc / 16, c = ((blk / bs) % bc) * 4 + (lc / 4) % 4,
              s = (blk % bs) * 4 + lc % 4;
    p = r < d.R && c < d.C && s < d.S ? (r * d.C + c) * d.S + s : d.P;
  }
  const bool active = p < d.P;
  const int pp = active ? p : d.P - 1;  // inactive threads only help stage the directions
  const int r = pp / (d.C * d.S), c = (pp / d.S) % d.C, s = pp % d.S;
  float ux, uy, uz;
  primary_direction(d, r, c, s, ux, uy, uz);
  PointData<GMAX> pd;
  load_point(d, pp, pd);
  PointSlots ps;
  load_slots(d, pp, ps);
  f4 I0 = {0.f, 0.f, 0.f, 0.f}, I1 = I0, I2 = I0;
  if (BLOCKED) {
    // block order: per octant the d8 samples in runs of eight (32 bytes, whole sectors when dp % 8 == 0; padding
    // zero), each run two 16-byte stores per group; the sums in direction order as below
    const int dp = (d8 + 3) & ~3;
    if ((int)threadIdx.x < d.K)
      sdir[threadIdx.x] = f4{d.omegas[3 * threadIdx.x], d.omegas[3 * threadIdx.x + 1], d.omegas[3 * threadIdx.x + 2],
                             weights[threadIdx.x]};
    __syncthreads();
    if (active) {
      const int64_t gstride = fc_sample_group_floats(d.R, d.C, d.S, dp);
      for (int o = 0; o < 8; ++o) {
        const int64_t base = fc_sample_run(d.C, d.S, dp, r, c, s, o);
        for (int d0 = 0; d0 < dp; d0 += 8) {
          f4 run[4][2];
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int dd = d0 + j;
            if (dd >= d8) {
#pragma unroll
              for (int gl = 0; gl < 4; ++gl) run[gl][j / 4][j % 4] = 0.0f;
              continue;
            }
            const f4 dir = sdir[o * d8 + dd];
            const float eta = direction_cosine(ux, uy, uz, dir.x, dir.y, dir.z);
            const f4 S = eval_source_channel_batched(d, pd, ps, eta, ch);
            const float w = dir.w;
            const float we = w * eta;
            const float wee = we * eta;
#pragma unroll
            for (int gl = 0; gl < 4; ++gl) {
              I0[gl] = fmaf(w, S[gl], I0[gl]);
              I1[gl] = fmaf(we, S[gl], I1[gl]);
              I2[gl] = fmaf(wee, S[gl], I2[gl]);
              run[gl][j / 4][j % 4] = S[gl];
            }
          }
#pragma unroll
          for (int gl = 0; gl < 4; ++gl) {
            if (4 * ch + gl >= d.G) continue;
            f4* dst = reinterpret_cast<f4*>(samples + (4 * ch + gl) * gstride + base + d0);
            dst[0] = run[gl][0];
            if (d0 + 4 < dp) dst[1] = run[gl][1];
          }
        }
      }
    }
  }
  // directions in runs of eight: with samples and K % 8 == 0 each group's eight samples leave as two adjacent 16-byte
  // stores, one whole 32-byte sector (runs of four leave half sectors, which reach memory as whole ones)
  constexpr int RUN = 8;
  const bool vec = samples && (d.K % RUN) == 0;
  for (int t0 = 0; t0 < (BLOCKED ? 0 : d.K); t0 += kBlock) {
    const int kt = min(kBlock, d.K - t0);
    __syncthreads();  // the previous tile is no longer read
    if ((int)threadIdx.x < kt) {
      const int k = t0 + (int)threadIdx.x;
      sdir[threadIdx.x] = f4{d.omegas[3 * k], d.omegas[3 * k + 1], d.omegas[3 * k + 2], weights[k]};
    }
    __syncthreads();
    if (!active) continue;
    for (int j0 = 0; j0 < kt; j0 += RUN) {
      const int k0 = t0 + j0;
      f4 run[4][RUN / 4];
#pragma unroll
      for (int j = 0; j < RUN; ++j) {
        const int k = k0 + j;
        if (k >= d.K) break;
        const f4 dir = sdir[j0 + j];
        const float eta = direction_cosine(ux, uy, uz, dir.x, dir.y, dir.z);
        const f4 S = eval_source_channel_batched(d, pd, ps, eta, ch);
        const float w = dir.w;
        const float we = w * eta;
        const float wee = we * eta;
#pragma unroll
        for (int gl = 0; gl < 4; ++gl) {
          I0[gl] = fmaf(w, S[gl], I0[gl]);
          I1[gl] = fmaf(we, S[gl], I1[gl]);
          I2[gl] = fmaf(wee, S[gl], I2[gl]);
          run[gl][j / 4][j % 4] = S[gl];
          const int g = 4 * ch + gl;
          if (samples && !vec && g < d.G) samples[((int64_t)g * d.P + p) * d.K + k] = S[gl];
        }
      }
      if (vec) {
#pragma unroll
        for (int gl = 0; gl < 4; ++gl) {
          if (4 * ch + gl >= d.G) continue;
          f4* dst = reinterpret_cast<f4*>(samples + ((int64_t)(4 * ch + gl) * d.P + p) * d.K + k0);
#pragma unroll
          for (int h = 0; h < RUN / 4; ++h) dst[h] = run[gl][h];
        }
      }
    }
  }
  // BLOCKED with factor outputs: per group [cells + CITRUS_FC_GROUP_PAD] and per cell u, in the block order of t
  // (cells past the grid zero); u from the lightest channel's threads
  const int64_t fcells = (int64_t)((d.R + 3) / 4) * ((d.C + 3) / 4) * ((d.S + 3) / 4) * 64;
  const bool factors = BLOCKED && fscale != nullptr && t < fcells;
  if (!active) {
    if (factors) {
#pragma unroll
      for (int gl = 0; gl < 4; ++gl)
        if (4 * ch + gl < d.G) {
          fscale[(4 * ch + gl) * (fcells + CITRUS_FC_GROUP_PAD) + t] = 0.0f;
          fbeta[(4 * ch + gl) * (fcells + CITRUS_FC_GROUP_PAD) + t] = 0.0f;
        }
      if (ch == 0) reinterpret_cast<f4*>(fu)[t] = f4{0.f, 0.f, 0.f, 0.f};
    }
    return;
  }
  // exact Legendre moments l = 0, 1 of the same source (no quadrature)
  f4 E0 = {0.f, 0.f, 0.f, 0.f}, E1 = E0;
  for (int sl = 0; sl < d.n_slots; ++sl) {
    const float rho = d.slot_density[(int64_t)sl * d.P + p];
    const float* ms = mu_s_l + (int64_t)d.slot_material[(int64_t)sl * d.P + p] * d.G * d.G * lp1;
#pragma unroll
    for (int gi = 0; gi < GMAX; ++gi) {
      if (gi / 4 > ch) break;
      const float cv = rho * pd.ph[gi];  // 0 for g' >= G
#pragma unroll
      for (int gl = 0; gl < 4; ++gl) {
        const int go = 4 * ch + gl;
        if (go < gi) continue;
        const int a = (gi < d.G ? gi : d.G - 1) * d.G + (go < d.G ? go : d.G - 1);
        E0[gl] = fmaf(cv, ms[a * lp1 + 0], E0[gl]);
        E1[gl] = fmaf(cv, ms[a * lp1 + 1], E1[gl]);
      }
    }
  }
  const int64_t plane = (int64_t)d.G * d.P;
#pragma unroll
  for (int gl = 0; gl < 4; ++gl) {
    const int g = 4 * ch + gl;
    if (g >= d.G) continue;
    const int64_t o = (int64_t)g * d.P + p;
    out[o] = I0[gl];
    out[plane + o] = I1[gl];
    out[2 * plane + o] = I2[gl];
    out[3 * plane + o] = E0[gl];
    out[4 * plane + o] = E1[gl];
    if (factors) {
      float sc, be;
      correction_factors(I0[gl], I1[gl], I2[gl], E0[gl], E1[gl], sc, be);
      fscale[g * (fcells + CITRUS_FC_GROUP_PAD) + t] = sc;
      fbeta[g * (fcells + CITRUS_FC_GROUP_PAD) + t] = be;
    }
  }
  if (factors && ch == 0) reinterpret_cast<f4*>(fu)[t] = f4{ux, uy, uz, 0.0f};
}

// In-place sampling correction of captured samples [G][P][K]: S scale (1 + beta eta), the operations and order of
// fc_moments_kernel at one point per cell, so the result equals its cell-mean layout bit for bit.
__global__ __launch_bounds__(kBlock) void fc_correct_kernel(FcDev d, const float* __restrict__ scale,
                                                           const float* __restrict__ beta,
                                                           float* __restrict__ samples) {
  const int64_t PK = (int64_t)d.P * d.K;  // (p, k), k fastest, over a stride loop; the groups in a loop
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < PK; idx += (int64_t)gridDim.x * kBlock) {
    const int k = (int)(idx % d.K);
    const int p = (int)(idx / d.K);
    const int r = p / (d.C * d.S), c = (p / d.S) % d.C, s = p % d.S;
    float ux, uy, uz;
    primary_direction(d, r, c, s, ux, uy, uz);
    const float eta = direction_cosine(ux, uy, uz, d.omegas[3 * k], d.omegas[3 * k + 1], d.omegas[3 * k + 2]);
    for (int g = 0; g < d.G; ++g) {
      const int64_t gp = (int64_t)g * d.P + p;
      const float v = samples[g * PK + idx];
      samples[g * PK + idx] = v * scale[gp] * fmaf(beta[gp], eta, 1.0f);
    }
  }
}

// Omega_UC of every point as the kernels above form it (FP64, rounded to float32), [P][4] with w = 0: the table the
// sweep reads when it applies the sampling correction to stored samples.
__global__ __launch_bounds__(kBlock) void fc_directions_kernel(FcDev d, float* __restrict__ out) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= d.P) return;
  const int r = p / (d.C * d.S), c = (p / d.S) % d.C, s = p % d.S;
  float ux, uy, uz;
  primary_direction(d, r, c, s, ux, uy, uz);
  reinterpret_cast<f4*>(out)[p] = f4{ux, uy, uz, 0.0f};
}

// fc_correct_kernel on cell means stored [V][G][N][M] with the directions given as tables: u [V][N][4] (from
// fc_directions_kernel) and omega [M][4]; the same operations, so the values are equal bit for bit. Thread per
// (view, cell, direction), the groups in a loop.
__global__ __launch_bounds__(kBlock) void fc_correct_means_kernel(float* __restrict__ samples,
                                                                 const float* __restrict__ scale,
                                                                 const float* __restrict__ beta,
                                                                 const float* __restrict__ u4,
                                                                 const float* __restrict__ omega4, int G, int64_t N,
                                                                 int M, int64_t total) {
  // (v, n, m), m fastest, over a stride loop
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * kBlock) {
    const int m = (int)(idx % M);
    const int64_t vn = idx / M;
    const int64_t n = vn % N, v = vn / N;
    const f4 u = reinterpret_cast<const f4*>(u4)[vn];
    const f4 o = reinterpret_cast<const f4*>(omega4)[m];
    const float eta = direction_cosine(u.x, u.y, u.z, o.x, o.y, o.z);
    for (int g = 0; g < G; ++g) {
      const int64_t gp = (v * G + g) * N + n;
      const float x = samples[gp * M + m];
      samples[gp * M + m] = x * scale[gp] * fmaf(beta[gp], eta, 1.0f);
    }
  }
}

template <int GMAX>
__global__ __launch_bounds__(kBlock) void fc_moments_kernel(FcDev d, const float* __restrict__ scale,
                                                            const float* __restrict__ beta, int f, int layout,
                                                            float* __restrict__ out) {
  const int Rc = d.R / f, Cc = d.C / f, Sc = d.S / f;
  const int Nc = Rc * Cc * Sc;
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (int64_t)Nc * d.K) return;
  // group-major output runs directions fastest, the other layouts cells fastest (coalesced stores either way)
  const bool dir_fast = layout >= 2;
  const int cell = dir_fast ? (int)(idx / d.K) : (int)(idx % Nc);
  const int k = dir_fast ? (int)(idx % d.K) : (int)(idx / Nc);
  const int rc = cell / (Cc * Sc), cc = (cell / Sc) % Cc, sc = cell % Sc;
  const float ox = d.omegas[3 * k], oy = d.omegas[3 * k + 1], oz = d.omegas[3 * k + 2];
  f4 acc[GMAX];  // per group: sum of v, v u_x, v u_y, v u_z
#pragma unroll
  for (int g = 0; g < GMAX; ++g) acc[g] = f4{0.f, 0.f, 0.f, 0.f};
  for (int jr = 0; jr < f; ++jr)
    for (int jc = 0; jc < f; ++jc)
      for (int js = 0; js < f; ++js) {
        const int r = rc * f + jr, c = cc * f + jc, s = sc * f + js;
        const int p = (r * d.C + c) * d.S + s;
        float ux, uy, uz;
        primary_direction(d, r, c, s, ux, uy, uz);
        PointData<GMAX> pd;
        load_point(d, p, pd);
        const float eta = direction_cosine(ux, uy, uz, ox, oy, oz);
        float S[GMAX];
        eval_source(d, pd, eta, S);
        // local coordinates u_j = (2j + 1)/f - 1 (x along cols, y along rows, z along slices)
        const float u_x = (2.0f * jc + 1.0f) / f - 1.0f;
        const float u_y = (2.0f * jr + 1.0f) / f - 1.0f;
        const float u_z = (2.0f * js + 1.0f) / f - 1.0f;
        const bool corrected = scale != nullptr;
#pragma unroll
        for (int g = 0; g < GMAX; ++g) {
          const int gg = g < d.G ? g : d.G - 1;  // groups past G carry S = 0; their loads are clamped
          float v = S[g];
          const float sc = corrected ? scale[(int64_t)gg * d.P + p] : 1.0f;
          const float be = corrected ? beta[(int64_t)gg * d.P + p] : 0.0f;
          v = v * sc * fmaf(be, eta, 1.0f);
          acc[g] += v * f4{1.0f, u_x, u_y, u_z};
        }
      }
  const float n3 = (float)(f * f * f);
  float slope = 0.0f;  // 1 / (n3 mean(u^2)) along one axis, 0 for f = 1
  if (f > 1) {
    float u2 = 0.0f;
    for (int j = 0; j < f; ++j) {
      const float u = (2.0f * j + 1.0f) / f - 1.0f;
      u2 += u * u;
    }
    slope = 1.0f / (n3 * (u2 / f));
  }
  // element (g, moment j) sits at base + g gstride + j jstride
  int64_t base, gstride, jstride;
  if (layout == 0) {
    base = (int64_t)k * d.G * 4 * Nc + cell; gstride = 4 * (int64_t)Nc; jstride = Nc;
  } else if (layout == 1) {
    base = ((int64_t)k * Nc + cell) * d.G * 4; gstride = 4; jstride = 1;
  } else if (layout == 2) {
    base = ((int64_t)cell * d.K + k) * 4; gstride = (int64_t)Nc * d.K * 4; jstride = 1;
  } else {  // cell means only (f = 1: the slopes are zero)
    base = (int64_t)cell * d.K + k; gstride = (int64_t)Nc * d.K; jstride = 0;
  }
#pragma unroll
  for (int g = 0; g < GMAX; ++g) {
    if (g >= d.G) continue;
    const f4 m = f4{acc[g].x / n3, acc[g].y * slope, acc[g].z * slope, acc[g].w * slope};
    float* o = out + base + g * gstride;
    if (jstride == 0) {
      *o = m.x;
    } else if (jstride == 1) {
      *reinterpret_cast<f4*>(o) = m;
    } else {
      o[0] = m.x; o[jstride] = m.y; o[2 * jstride] = m.z; o[3 * jstride] = m.w;
    }
  }
}

// The source of the output channel ch (groups 4 ch .. 4 ch + 3) at eta node j, contracted once per point: node j of
// eval_source's loop on its own, restricted to one channel (per accumulator the slots and source groups in the same
// order, the same fused multiply-adds), so that interpolating two contracted nodes gives eval_source's value bit for
// bit. ph: the point's phi_UC of every group (0 past G), a row of LDS; ps: its material slots. The table rows of four
// source groups are requested together before they are added.
template <int GMAX>
__device__ __forceinline__ f4 contract_node_channel(const FcDev& d, const PointSlots& ps, const float* ph, int j,
                                                    int ch) {
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  const int64_t node = (int64_t)d.M * d.G * GMAX;
#pragma unroll
  for (int s = 0; s < kMaxSlots; ++s) {
    if (s >= d.n_slots) break;
    const float rho = ps.rho[s];
    if (rho == 0.0f) continue;  // as in eval_source
    const float* t0 = d.table + (int64_t)j * node + (int64_t)ps.mat[s] * d.G * GMAX;
#pragma unroll
    for (int gb = 0; gb < GMAX / 4; ++gb) {
      if (gb > ch) break;
      f4 x[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gi = 4 * gb + q;
        const int row = gi < d.G ? gi : d.G - 1;
        x[q] = reinterpret_cast<const f4*>(t0 + row * GMAX)[ch];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) acc = fma4(rho * ph[4 * gb + q], x[q], acc);
    }
  }
  return acc;
}

// contract_node_channel for G == GMAX (the emission kernel's straight-line form): the table through a buffer
// resource, one 32-bit offset per (node, slot), the rows of a batch in the loads' immediate offsets, so no address is
// formed per load; the same terms in the same order, so the value is the same bit for bit.
constexpr int kTableRsrcWord3 = 0x00020000;  // raw buffer, 32-bit data format

template <int GMAX>
__device__ __forceinline__ f4 contract_node_channel_exact(__amdgpu_buffer_rsrc_t table, const FcDev& d,
                                                          const PointSlots& ps, const float* ph, int j, int ch) {
  f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < kMaxSlots; ++s) {
    if (s >= d.n_slots) break;
    const float rho = ps.rho[s];
    if (rho == 0.0f) continue;  // as in eval_source
    const unsigned base = (unsigned)((j * d.M + ps.mat[s]) * GMAX * GMAX * 4 + ch * 16);
#pragma unroll
    for (int gb = 0; gb < GMAX / 4; ++gb) {
      if (gb > ch) break;
      f4 x[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        x[q] = __builtin_bit_cast(
            f4, __builtin_amdgcn_raw_buffer_load_b128(table, (int)(base + (4 * gb + q) * GMAX * 4), 0, 0));
#pragma unroll
      for (int q = 0; q < 4; ++q) acc = fma4(rho * ph[4 * gb + q], x[q], acc);
    }
  }
  return acc;
}

// Workgroup shape of the emission kernel: a run of kEmPoints consecutive points (coalesced outputs) times the
// directions of a tile of kEmTile, each point's directions spread over kEmLanes threads (kEmPerLane each).
constexpr int kEmThreads = 512;
constexpr int kEmPoints = CITRUS_EMISSION_BLOCK;
constexpr int kEmLanes = kEmThreads / kEmPoints;
constexpr int kEmTile = 64;
constexpr int kEmPerLane = kEmTile / kEmLanes;

template <int GMAX>
struct EmissionShape {
  static constexpr int kNodes = (GMAX <= 16 ? 384 : 352) / GMAX;  // eta nodes of a slice in LDS (24 at 16 groups)
  static constexpr int kStride = kNodes * GMAX + 4;  // floats per point: 16-byte rows of neighbouring points 4 apart
};
constexpr int kEmSumStride = kEmTile + 1;  // floats per point of the per-pair sums (rows in different banks)

// Attenuated total emission of the collapsed-cone image at the quadrature points (uncollided_kernels.py:713-830,
// scatter_image.py:829-875): e_g = (S_FC,g(point, w) + S_MC,g(cell, w)) exp(-tau_j,g) with the first-order in-cell
// optical depth tau_j = max(0, tau_cell - mu_cell (w . (r_j - r_cell))) (tau_cell exactly for f = 1).
//
// A workgroup takes kEmPoints consecutive points and every direction, in tiles of kEmTile. The directions of a tile
// leave a point within a narrow cone, so their eta = Omega_UC . w fall in a short run of table nodes: the workgroup
// finds that window, contracts each of its points' sources at every node of the window once into LDS
// (contract_node_channel, one (node, channel) per thread and step) and interpolates per (point, direction) from LDS --
// eval_source's value bit for bit, with about 7x fewer table gathers than two nodes per (point, direction) at 64
// panel-facing directions and 257 nodes. A window wider than LDS holds is taken in slices of kNodes nodes overlapping
// by one; each (point, direction) is evaluated in the slice that holds both of its nodes.
// Each (direction, workgroup) also sums sum_g e and h sum_g e over the workgroup's points into
// partial [K][blocks][2] (through LDS, in point order; no atomics): the footprint's emission-weighted mean height.
// COLLAPSED (one point per cell, signal weights): `multiple` is [K][2][cells], the multiple-collision part already
// attenuated and summed over the groups, weighted (0) and unweighted (1), by citrus_multiple_emission.
template <int GMAX, bool COLLAPSED, bool FAST>
__global__ __launch_bounds__(kEmThreads) __attribute__((amdgpu_waves_per_eu(4))) void emission_kernel(
    FcDev d, const float* __restrict__ tau, int tau_channels, const float* __restrict__ tau_mix,
    const float* __restrict__ mu_c, const float* __restrict__ multiple, int f, const float* __restrict__ heights,
    const float* __restrict__ signal_weight, float* __restrict__ out, float* __restrict__ partial) {
  using Shape = EmissionShape<GMAX>;
  constexpr int NCH = GMAX / 4;
  __shared__ f4 s_nodes[kEmPoints * Shape::kStride / 4];  // [point][node of the slice][GMAX] contracted sources
  __shared__ f4 s_dir[kEmTile];                            // the tile's directions
  __shared__ float s_weight[kEmTile * GMAX];               // their signal weights per group
  __shared__ float s_mix[32 * kMaxTauChannels];
  __shared__ float s_sum[2][kEmPoints * kEmSumStride];    // per (point, direction): sum_g e and h sum_g e
  __shared__ float s_ph[kEmPoints][GMAX + 1];              // phi_UC of the workgroup's points
  __shared__ int s_window[2];
  // the table as a buffer (the straight-line form's node contraction: 32-bit offsets)
  const __amdgpu_buffer_rsrc_t table_r = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(d.table), 0, FAST ? d.n_eta * d.M * d.G * GMAX * 4 : 0, kTableRsrcWord3);
  float* const s_node = reinterpret_cast<float*>(s_nodes);

  const int tid = threadIdx.x;
  const int pl = tid % kEmPoints, dl = tid / kEmPoints;
  const int p = blockIdx.x * kEmPoints + pl;
  const bool active = p < d.P;
  const int pp = active ? p : d.P - 1;  // inactive threads evaluate nothing but take part in the shuffles
  const int r = pp / (d.C * d.S), c = (pp / d.S) % d.C, s = pp % d.S;
  const int Rc = d.R / f, Cc = d.C / f, Sc = d.S / f;
  const int Nc = Rc * Cc * Sc;
  const int cell = ((r / f) * Cc + (c / f)) * Sc + (s / f);
  float ux, uy, uz;
  primary_direction(d, r, c, s, ux, uy, uz);
  for (int g = dl; g < GMAX; g += kEmLanes) s_ph[pl][g] = g < d.G ? d.phi_uc[(int64_t)g * d.P + pp] : 0.0f;
  PointSlots ps;
  load_slots(d, pp, ps);
  const float h = heights ? heights[pp] : 0.0f;
  // offset of the point from its cell centre (centred sub-point index times the pitch), dotted with w per direction
  const float half = 0.5f * (float)(f - 1);
  const float offx = ((float)(c % f) - half) * (float)d.dx, offy = ((float)(r % f) - half) * (float)d.dy,
              offz = ((float)(s % f) - half) * (float)d.dz;
  if (tau_mix && tid < d.G * tau_channels) s_mix[tid] = tau_mix[tid];
  const bool two = tau_mix && tau_channels <= 2;  // the depths of a direction in two registers

  // e of (point, direction j of the tile, cell means interpolated from the slice's LDS row a0) into out, returning
  // sum_g e (with the collapsed unweighted part)
  auto emit = [&](const int k, const int j, const f4 w, const f4* a0, const float w1) __attribute__((always_inline)) {
    if constexpr (FAST) {
      // the view step's form: collapsed, two traced depth channels, signal weights, one point per cell, G == GMAX
      const float mcw = multiple[(int64_t)k * 2 * Nc + cell];
      const float mcs = multiple[((int64_t)k * 2 + 1) * Nc + cell];
      const float tc0 = tau[(int64_t)k * 2 * Nc + cell];
      const float tc1 = tau[((int64_t)k * 2 + 1) * Nc + cell];
      const bool unreachable = tc0 == INFINITY;
      float sig = 0.0f, es = 0.0f;
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        const f4 v = interpolate4(a0[ch], a0[NCH + ch], w1);
        const f4 wv = reinterpret_cast<const f4*>(s_weight + j * GMAX)[ch];
#pragma unroll
        for (int gl = 0; gl < 4; ++gl) {
          const int g = 4 * ch + gl;
          float t = fmaf(tau_mix[2 * g + 1], tc1, tau_mix[2 * g] * tc0);
          t = unreachable ? INFINITY : t;
          const float e = v[gl] * attenuation(t);
          es += e;
          sig = fmaf(wv[gl], e, sig);
        }
      }
      out[(int64_t)k * d.P + p] = sig + mcw;
      return es + mcs;
    }
    float mcw = 0.0f, mcs = 0.0f;
    if constexpr (COLLAPSED) {
      mcw = multiple[(int64_t)k * 2 * Nc + cell];
      mcs = multiple[((int64_t)k * 2 + 1) * Nc + cell];
    }
    float tc0 = 0.0f, tc1 = 0.0f;
    if (two) {
      tc0 = tau[(int64_t)k * tau_channels * Nc + cell];
      tc1 = tau_channels > 1 ? tau[((int64_t)k * tau_channels + 1) * Nc + cell] : 0.0f;
    }
    const bool unreachable =
        tau_mix && (two ? tc0 : tau[(int64_t)k * tau_channels * Nc + cell]) == INFINITY;
    const float off = f > 1 ? w.x * offx + w.y * offy + w.z * offz : 0.0f;
    float sig = 0.0f, es = 0.0f;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      const f4 v = interpolate4(a0[ch], a0[NCH + ch], w1);
      const f4 wv = signal_weight ? reinterpret_cast<const f4*>(s_weight + j * GMAX)[ch] : f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int gl = 0; gl < 4; ++gl) {
        const int g = 4 * ch + gl;
        if (g >= d.G) continue;
        const int64_t kg = (int64_t)k * d.G + g;
        float t;
        if (two) {  // the mix coefficients are uniform: scalar loads
          t = fmaf(tau_mix[g * tau_channels], tc0, 0.0f);
          if (tau_channels > 1) t = fmaf(tau_mix[g * tau_channels + 1], tc1, t);
          if (unreachable) t = INFINITY;
        } else if (tau_mix) {
          t = 0.0f;
          for (int ci = 0; ci < tau_channels; ++ci)
            t = fmaf(s_mix[g * tau_channels + ci], tau[((int64_t)k * tau_channels + ci) * Nc + cell], t);
          if (unreachable) t = INFINITY;
        } else {
          t = tau[kg * Nc + cell];
        }
        if (f > 1) t = fmaxf(t - mu_c[(int64_t)g * Nc + cell] * off, 0.0f);
        const float src = COLLAPSED ? v[gl] : v[gl] + multiple[kg * Nc + cell];
        const float e = src * attenuation(t);
        es += e;
        if (signal_weight) {
          sig = fmaf(wv[gl], e, sig);
        } else {
          out[kg * d.P + p] = e;
        }
      }
    }
    if constexpr (COLLAPSED) {
      sig += mcw;
      es += mcs;
    }
    if (signal_weight) out[(int64_t)k * d.P + p] = sig;
    return es;
  };

  for (int k0 = 0; k0 < d.K; k0 += kEmTile) {
    const int kt = min(kEmTile, d.K - k0);
    __syncthreads();  // the previous tile's directions, weights, nodes and sums are no longer read
    if (tid < kt)
      s_dir[tid] = f4{d.omegas[3 * (k0 + tid)], d.omegas[3 * (k0 + tid) + 1], d.omegas[3 * (k0 + tid) + 2], 0.f};
    for (int i = tid; i < kEmPoints * kEmSumStride; i += kEmThreads) s_sum[0][i] = s_sum[1][i] = 0.0f;
    if (signal_weight)
      for (int i = tid; i < kt * d.G; i += kEmThreads)
        s_weight[(i / d.G) * GMAX + i % d.G] = signal_weight[(int64_t)k0 * d.G + i];
    if (tid == 0) {
      s_window[0] = 0x7fffffff;
      s_window[1] = -1;
    }
    __syncthreads();

    // ---- the window of eta nodes the tile's directions reach from the workgroup's points -------------------------
    int lo = 0x7fffffff, hi = -1;
    if (active) {
      for (int j = dl; j < kt; j += kEmLanes) {
        const f4 w = s_dir[j];
        int i0;
        float w1;
        table_position(d, direction_cosine(ux, uy, uz, w.x, w.y, w.z), i0, w1);
        lo = min(lo, i0);
        hi = max(hi, i0);
      }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      lo = min(lo, __shfl_xor(lo, m));
      hi = max(hi, __shfl_xor(hi, m));
    }
    if ((tid & 63) == 0) {
      atomicMin(&s_window[0], lo);
      atomicMax(&s_window[1], hi);
    }
    __syncthreads();
    const int jlo = s_window[0], jhi = s_window[1];  // lower nodes of the tile (jhi < jlo: no active point)

    // slices of lower nodes [a, a + ns - 2] with their upper nodes: nodes a .. a + ns - 1 in LDS
    for (int a = jlo; a <= jhi; a += Shape::kNodes - 1) {
      const int ns = min(Shape::kNodes, jhi + 2 - a);
      if (a > jlo) __syncthreads();  // the previous slice is no longer read
      if (active) {
        for (int q = dl; q < ns * NCH; q += kEmLanes) {
          const int jj = q / NCH, ch = q % NCH;
          reinterpret_cast<f4*>(s_node + pl * Shape::kStride + jj * GMAX)[ch] =
              FAST ? contract_node_channel_exact<GMAX>(table_r, d, ps, s_ph[pl], a + jj, ch)
                   : contract_node_channel<GMAX>(d, ps, s_ph[pl], a + jj, ch);
        }
      }
      __syncthreads();

      // ---- emission per (point, direction) whose nodes the slice holds ----------------------------------------------
#pragma unroll 2
      for (int i = 0; i < kEmPerLane; ++i) {
        const int j = dl + kEmLanes * i;
        if (!active || j >= kt) continue;
        const f4 w = s_dir[j];
        int i0;
        float w1;
        table_position(d, direction_cosine(ux, uy, uz, w.x, w.y, w.z), i0, w1);
        if (i0 < a || i0 > a + ns - 2) continue;
        const float es =
            emit(k0 + j, j, w, reinterpret_cast<const f4*>(s_node + pl * Shape::kStride + (i0 - a) * GMAX), w1);
        s_sum[0][pl * kEmSumStride + j] = es;
        s_sum[1][pl * kEmSumStride + j] = es * h;
      }
    }
    __syncthreads();
    // the direction's sums over the workgroup's points, in point order (zeros where a point has no pair)
    if (tid < kt) {
      float se = 0.0f, sh = 0.0f;
      for (int q = 0; q < kEmPoints; ++q) {
        se += s_sum[0][q * kEmSumStride + tid];
        sh += s_sum[1][q * kEmSumStride + tid];
      }
      float* o = partial + ((int64_t)(k0 + tid) * gridDim.x + blockIdx.x) * 2;
      o[0] = se;
      o[1] = sh;
    }
  }
}

int check_args(const CitrusFcArgs* a) {
  if (!a) return fail(CITRUS_ERR_ARG, "args is NULL");
  if (a->G < 1 || a->G > 32) return fail(CITRUS_ERR_ARG, "1..32 groups supported");
  if (a->M < 1 || a->n_eta < 2 || a->K < 0) return fail(CITRUS_ERR_ARG, "M >= 1, n_eta >= 2 and K >= 0 required");
  if (a->n_slots < 1 || a->n_slots > kMaxSlots) return fail(CITRUS_ERR_ARG, "1..4 material slots per point supported");
  if (a->eta_map && a->n_map < 2) return fail(CITRUS_ERR_ARG, "an adaptive map needs at least 2 knots");
  if (a->R < 1 || a->C < 1 || a->S < 1 || (int64_t)a->R * a->C * a->S > 0x7fffffff)
    return fail(CITRUS_ERR_ARG, "point grid must be non-empty and below 2^31 points");
  if (!a->table || !a->slot_material || !a->slot_density || !a->phi_uc || (a->K > 0 && !a->omegas))
    return fail(CITRUS_ERR_ARG, "NULL table or buffer");
  return CITRUS_OK;
}

template <typename Launch>
int dispatch_groups(int G, Launch&& launch) {
  if (G <= 4) return launch(std::integral_constant<int, 4>{});
  if (G <= 8) return launch(std::integral_constant<int, 8>{});
  if (G <= 16) return launch(std::integral_constant<int, 16>{});
  return launch(std::integral_constant<int, 32>{});
}

int finish(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return CITRUS_OK;
}

}  // namespace
}  // namespace citrus

extern "C" int citrus_fc_source(const CitrusFcArgs* args, float* out, void* stream) {
  using namespace citrus;
  if (int rc = check_args(args)) return rc;
  if (!out) return fail(CITRUS_ERR_ARG, "NULL output");
  const FcDev d = make_dev(*args);
  const int64_t n = (int64_t)d.P * d.K;
  if (n == 0) return CITRUS_OK;
  if (!grid_fits((n + kBlock - 1) / kBlock, kBlock))
    return fail(CITRUS_ERR_ARG, "points x directions exceed 2^32 work-items per launch: pass fewer directions");
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock)), block(kBlock);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  dispatch_groups(d.G, [&](auto gm) {
    hipLaunchKernelGGL((fc_source_kernel<decltype(gm)::value>), grid, block, 0, s, d, out);
    return 0;
  });
  return finish("first-collision source");
}

extern "C" int citrus_fc_sums(const CitrusFcArgs* args, const float* weights, const float* mu_s_l, int lp1,
                              float* out, float* samples, void* stream) {
  using namespace citrus;
  if (int rc = check_args(args)) return rc;
  if (!weights || !mu_s_l || !out) return fail(CITRUS_ERR_ARG, "NULL weights, moments or output");
  if (lp1 < 2) return fail(CITRUS_ERR_ARG, "the exact moments need Legendre order >= 1");
  const FcDev d = make_dev(*args);
  const int blocks = (d.P + kBlock - 1) / kBlock;
  if (!grid_fits((int64_t)blocks * (group_pad(d.G) / 4), kBlock))
    return fail(CITRUS_ERR_ARG, "points x group channels exceed 2^32 work-items per launch");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  dispatch_groups(d.G, [&](auto gm) {
    constexpr int NCH = decltype(gm)::value / 4;
    hipLaunchKernelGGL((fc_sums_kernel<decltype(gm)::value, false>), dim3((unsigned)(blocks * NCH)), dim3(kBlock), 0,
                       s, d, weights, mu_s_l, lp1, out, samples, blocks, 0, nullptr, nullptr, nullptr);
    return 0;
  });
  return finish("first-collision sums");
}

extern "C" int citrus_fc_sums_blocked(const CitrusFcArgs* args, const float* weights, const float* mu_s_l, int lp1,
                                      float* out, float* samples, int D8, float* scale, float* beta, float* u,
                                      void* stream) {
  using namespace citrus;
  if (int rc = check_args(args)) return rc;
  if (!weights || !mu_s_l || !out || !samples) return fail(CITRUS_ERR_ARG, "NULL weights, moments, output or samples");
  if ((scale == nullptr) != (beta == nullptr) || (scale == nullptr) != (u == nullptr))
    return fail(CITRUS_ERR_ARG, "scale, beta and u are given together or not at all");
  if ((reinterpret_cast<uintptr_t>(samples) & 15) != 0 || (reinterpret_cast<uintptr_t>(u) & 15) != 0)
    return fail(CITRUS_ERR_ARG, "samples and u must be 16-byte aligned (they are stored four floats at a time)");
  if (lp1 < 2) return fail(CITRUS_ERR_ARG, "the exact moments need Legendre order >= 1");
  if (D8 < 1 || args->K != 8 * D8 || args->K > kBlock)
    return fail(CITRUS_ERR_ARG, "the block order needs K = 8 D8 sweep directions (octant-major), at most 256");
  const FcDev d = make_dev(*args);
  // threads over the padded block order: 64 per 4^3 block
  const int64_t cells = (int64_t)((d.R + 3) / 4) * ((d.C + 3) / 4) * ((d.S + 3) / 4) * 64;
  const int blocks = (int)((cells + kBlock - 1) / kBlock);
  if (!grid_fits((int64_t)blocks * (group_pad(d.G) / 4), kBlock))
    return fail(CITRUS_ERR_ARG, "points x group channels exceed 2^32 work-items per launch");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  dispatch_groups(d.G, [&](auto gm) {
    constexpr int NCH = decltype(gm)::value / 4;
    hipLaunchKernelGGL((fc_sums_kernel<decltype(gm)::value, true>), dim3((unsigned)(blocks * NCH)), dim3(kBlock), 0,
                       s, d, weights, mu_s_l, lp1, out, samples, blocks, D8, scale, beta, u);
    return 0;
  });
  return finish("first-collision sums (block order)");
}

extern "C" int citrus_fc_correct(const CitrusFcArgs* args, const float* scale, const float* beta, float* samples,
                                 void* stream) {
  using namespace citrus;
  if (int rc = check_args(args)) return rc;
  if (!scale || !beta || !samples) return fail(CITRUS_ERR_ARG, "NULL scale, beta or samples");
  const FcDev d = make_dev(*args);
  const int64_t n = (int64_t)d.P * d.K;
  if (n == 0) return CITRUS_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(fc_correct_kernel, dim3(stride_grid(n, kBlock)), dim3(kBlock), 0, s, d, scale, beta, samples);
  return finish("first-collision correction");
}

extern "C" int citrus_fc_primary_directions(const CitrusFcArgs* args, float* out, void* stream) {
  using namespace citrus;
  if (int rc = check_args(args)) return rc;
  if (!out) return fail(CITRUS_ERR_ARG, "NULL output");
  const FcDev d = make_dev(*args);
  hipLaunchKernelGGL(fc_directions_kernel, dim3((unsigned)((d.P + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), d, out);
  return finish("first-collision primary directions");
}

extern "C" int citrus_fc_correct_means(float* samples, const float* scale, const float* beta, const float* u4,
                                       const float* omega4, int V, int G, int64_t N, int M, void* stream) {
  using namespace citrus;
  if (!samples || !scale || !beta || !u4 || !omega4) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (V < 1 || G < 1 || N < 1 || M < 1) return fail(CITRUS_ERR_ARG, "V, G, N and M must be >= 1");
  const int64_t total = (int64_t)V * N * M;
  hipLaunchKernelGGL(fc_correct_means_kernel, dim3(stride_grid(total, kBlock)), dim3(kBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), samples, scale, beta, u4, omega4, G, N, M, total);
  return finish("first-collision correction of cell means");
}

extern "C" int citrus_fc_moments(const CitrusFcArgs* args, const float* scale, const float* beta, int factor,
                                 int layout, float* out, void* stream) {
  using namespace citrus;
  if (int rc = check_args(args)) return rc;
  if (!out) return fail(CITRUS_ERR_ARG, "NULL output");
  if ((scale == nullptr) != (beta == nullptr)) return fail(CITRUS_ERR_ARG, "scale and beta must be given together");
  if (factor < 1 || args->R % factor || args->C % factor || args->S % factor)
    return fail(CITRUS_ERR_ARG, "the point grid must be divisible by the factor");
  if (layout < 0 || layout > 3) return fail(CITRUS_ERR_ARG, "layout must be 0, 1, 2 or 3");
  if (layout == 3 && factor != 1) return fail(CITRUS_ERR_ARG, "layout 3 (cell means) needs factor 1");
  const FcDev d = make_dev(*args);
  const int64_t n = (int64_t)(d.P / (factor * factor * factor)) * d.K;
  if (n == 0) return CITRUS_OK;
  if (!grid_fits((n + kBlock - 1) / kBlock, kBlock))
    return fail(CITRUS_ERR_ARG, "cells x directions exceed 2^32 work-items per launch: pass fewer directions");
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock)), block(kBlock);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  dispatch_groups(d.G, [&](auto gm) {
    hipLaunchKernelGGL((fc_moments_kernel<decltype(gm)::value>), grid, block, 0, s, d, scale, beta, factor, layout,
                       out);
    return 0;
  });
  return finish("first-collision moments");
}

extern "C" int citrus_attenuated_emission(const CitrusFcArgs* args, const float* tau, int tau_channels,
                                          const float* tau_mix, const float* mu_t_coarse, const float* multiple,
                                          int factor, const float* heights, const float* signal_weight, float* out,
                                          float* partial, void* stream) {
  using namespace citrus;
  if (int rc = check_args(args)) return rc;
  if (!tau || !mu_t_coarse || !multiple || !out || !partial) return fail(CITRUS_ERR_ARG, "NULL input or output");
  if (tau_mix && (tau_channels < 1 || tau_channels > kMaxTauChannels))
    return fail(CITRUS_ERR_ARG, "1..8 optical-depth channels with a mix table");
  if (factor < 1 || args->R % factor || args->C % factor || args->S % factor)
    return fail(CITRUS_ERR_ARG, "the point grid must be divisible by the factor");
  if (args->K > 65535) return fail(CITRUS_ERR_ARG, "at most 65535 directions per launch");
  const FcDev d = make_dev(*args);
  if (d.K == 0) return CITRUS_OK;
  const int64_t blocks = ((int64_t)d.P + kEmPoints - 1) / kEmPoints;
  if (!grid_fits(blocks, kEmThreads)) return fail(CITRUS_ERR_ARG, "points exceed 2^32 work-items per launch");
  const dim3 grid((unsigned)blocks), block(kEmThreads);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  dispatch_groups(d.G, [&](auto gm) {
    hipLaunchKernelGGL((emission_kernel<decltype(gm)::value, false, false>), grid, block, 0, s, d, tau, tau_channels,
                       tau_mix, mu_t_coarse, multiple, factor, heights, signal_weight, out, partial);
    return 0;
  });
  return finish("attenuated emission");
}

extern "C" int citrus_attenuated_emission_collapsed(const CitrusFcArgs* args, const float* tau, int tau_channels,
                                                    const float* tau_mix, const float* collapsed,
                                                    const float* heights, const float* signal_weight, float* out,
                                                    float* partial, void* stream) {
  using namespace citrus;
  if (int rc = check_args(args)) return rc;
  if (!tau || !collapsed || !signal_weight || !out || !partial)
    return fail(CITRUS_ERR_ARG, "NULL input or output (the collapsed form needs the signal weights)");
  if (tau_mix && (tau_channels < 1 || tau_channels > kMaxTauChannels))
    return fail(CITRUS_ERR_ARG, "1..8 optical-depth channels with a mix table");
  if (args->K > 65535) return fail(CITRUS_ERR_ARG, "at most 65535 directions per launch");
  const FcDev d = make_dev(*args);
  if (d.K == 0) return CITRUS_OK;
  const int64_t blocks = ((int64_t)d.P + kEmPoints - 1) / kEmPoints;
  if (!grid_fits(blocks, kEmThreads)) return fail(CITRUS_ERR_ARG, "points exceed 2^32 work-items per launch");
  const dim3 grid((unsigned)blocks), block(kEmThreads);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // the view step's form (two traced depth channels, G filling the kernel's group width) runs without per-group
  // tests
  const bool fast = tau_mix && tau_channels == 2;
  dispatch_groups(d.G, [&](auto gm) {
    constexpr int GM = decltype(gm)::value;
    if (fast && d.G == GM)
      hipLaunchKernelGGL((emission_kernel<GM, true, true>), grid, block, 0, s, d, tau, tau_channels, tau_mix, nullptr,
                         collapsed, 1, heights, signal_weight, out, partial);
    else
      hipLaunchKernelGGL((emission_kernel<GM, true, false>), grid, block, 0, s, d, tau, tau_channels, tau_mix,
                         nullptr, collapsed, 1, heights, signal_weight, out, partial);
    return 0;
  });
  return finish("attenuated emission (collapsed)");
}
