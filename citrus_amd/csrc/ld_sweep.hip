// LD discrete-ordinates wavefront sweep for gfx950 (MI355X).
//
// Work decomposition. The grid is cut into cubic blocks of B^3 cells. For one octant the blocks form diagonal
// "block planes" bi' + bj' + bk' = d (primes: coordinates mirrored into the octant's sweep order); all blocks of a
// block plane are independent. One workgroup sweeps one (view, octant, group, block) task for all dc directions of
// the chunk: it walks the block's 3B-2 cell planes, keeps the three face-flux arrays of every direction in LDS
// (updated in place, since the cells of one plane touch disjoint face entries), and exchanges the block's outgoing
// faces with its downstream neighbours through a global face buffer between launches.
//
// Octants o and 7-o mirror every axis, so block b sits on block plane d for o and K-d for 7-o (K = sum(nb_a-1)):
// one launch runs both octants of such a pair on block plane d without ever touching the same block twice, except
// on the middle plane when K is even, which is split into two launches. Launch order therefore fixes the order in
// which the eight octants add into each cell's moments, and the sweep is bitwise reproducible.
//
// Per cell plane the workgroup (1) stages the plane's scattering-source moments T [NSH][4] in LDS, (2) solves every
// (cell, direction) pair: source S_k = sum_n Y[d][n] T[n][k] (+ an explicit source), closed-form LD solve
// (citrus/scatter_solver/linear_discontinuous_cell.py:59-87, docs/source/method.rst:254-279), outgoing faces back to
// LDS, and (3) reduces the plane's angular fluence to moments phi[n][k] = sum_d Yw[d][n] psi[d][k] and adds them to
// HBM. The angular fluence never leaves the chip.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "citrus_hip.h"
#include "common.h"
#include "sweep_device.h"

namespace citrus {
namespace {

constexpr int kThreads = 128;


struct LaunchArgs {
  int R, C, S, N;
  int nbr, nbc, nbs, nblocks;  // blocks along rows (y), cols (x), slices (z)
  int G, D8, d0, dc, V;
  int first;      // this launch belongs to the first pass and the first chunk: first visits store phi
  int write_psi;  // psi overwrites the explicit source in place
  int src_layout;  // 0: cell-major chunk [8 dc][N][G][4]; 1: group-major [G][N][M][4], M = 8 D8; 2: [G][N][M] means
  int M;
  int ntask;
  const int4* tasks;  // (octant | first_in_pair << 3 | slot << 4, block row, block col, block slice)
  const float* mu;
  const float* T;
  float* src;
  float* phi;
  const float* Y;
  const float* Yw;
  const float* cdir;
  float* faces;
  int64_t tview, sview, fview;
  const uint16_t* T16;     // mixed operand: bfloat16 degrees l >= 2 [V][G][N][4][KH] (or NULL)
  const float* Tlo;        // mixed operand: float32 degrees l <= 1 [V][G][N][4][4]
  int64_t t16view;         // elements per view of T16
};

template <int NSH>
struct ReduceShape {
  static constexpr int kPer = NSH >= 16 ? 2 : 1;             // moments per thread in the reduction
  static constexpr int kChunks = (NSH + kPer - 1) / kPer;
};

template <int NSH, int B, bool HAS_T, bool HAS_SRC>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) void ld_sweep_kernel(LaunchArgs a) {
  constexpr int BB = B * B;
  constexpr PlaneTable<B> ptab{};
  constexpr int PMAX = ptab.max_plane();
  constexpr int NPL = PlaneTable<B>::kPlanes;
  constexpr int KPER = ReduceShape<NSH>::kPer;
  constexpr int KCH = ReduceShape<NSH>::kChunks;
  constexpr int TPT = (PMAX * NSH + kThreads - 1) / kThreads;  // float4 of T staged per thread and plane
  constexpr int RPT = (PMAX * KCH + kThreads - 1) / kThreads;  // reduction items per thread and plane
  const PlaneTable<B>& tab = planes<B>();

  extern __shared__ float lds[];
  const int dc = a.dc;
  float* face = lds;                            // [3 axes][3 components][BB positions][dc]
  float* tpl = face + 9 * dc * BB;              // [PMAX][NSH + 1][4]
  float* psl = tpl + PMAX * (NSH + 1) * 4;      // [PMAX][dc][4]
  float* ywl = psl + PMAX * dc * 4;             // [dc][NSH]
  int* slot_cell = reinterpret_cast<int*>(ywl + dc * NSH);      // [B^3] flat cell of each slot, or -1
  int* slot_pos = slot_cell + B * B * B;                         // [B^3] local (x, y, z) of each slot, packed
  float* tmu = reinterpret_cast<float*>(slot_pos + B * B * B);  // [2][PMAX] mu_t of the staged planes

  const int tid = threadIdx.x;
  const int task = blockIdx.x % a.ntask;
  const int rest = blockIdx.x / a.ntask;
  const int g = rest % a.G;
  const int v = rest / a.G;
  const int4 e = a.tasks[task];
  const int o = e.x & 7;
  const bool first_visit = a.first && ((e.x >> 3) & 1);
  const int slot = (e.x >> 4) & 1;
  const int bi = e.y, bj = e.z, bk = e.w;
  const int sx = (o & 4) ? -1 : 1, sy = (o & 2) ? -1 : 1, sz = (o & 1) ? -1 : 1;
  const float fsx = (float)sx, fsy = (float)sy, fsz = (float)sz;
  const int blk = (bi * a.nbc + bj) * a.nbs + bk;
  const int x0 = bj * B, y0 = bi * B, z0 = bk * B;

  const int64_t face_stride = (int64_t)dc * BB * 3;  // floats per (block, axis)
  float* gface_v = a.faces + (int64_t)v * a.fview + ((int64_t)slot * a.G + g) * a.nblocks * 3 * face_stride;
  const float* mu_g = a.mu + (int64_t)g * a.N;
  const float* T_g = HAS_T ? a.T + (int64_t)v * a.tview + (int64_t)g * a.N * NSH * 4 : nullptr;
  float* phi_g = a.phi + (int64_t)v * a.tview + (int64_t)g * a.N * NSH * 4;
  float* src_v = HAS_SRC ? a.src + (int64_t)v * a.sview : nullptr;

  // ---- per-task tables in LDS: slots, direction weights, incoming faces ------------------------------------------
  for (int sl = tid; sl < B * B * B; sl += kThreads) {
    const int packed = tab.cell[sl];
    const int pa = packed & 0xff, pb = (packed >> 8) & 0xff, pc = (packed >> 16) & 0xff;
    const int lx = sx > 0 ? pa : B - 1 - pa;
    const int ly = sy > 0 ? pb : B - 1 - pb;
    const int lz = sz > 0 ? pc : B - 1 - pc;
    const int x = x0 + lx, yy = y0 + ly, z = z0 + lz;
    slot_cell[sl] = (x >= a.C || yy >= a.R || z >= a.S) ? -1 : (yy * a.C + x) * a.S + z;
    slot_pos[sl] = lx | (ly << 8) | (lz << 16);
  }
  const int mbase = o * a.D8 + a.d0;
  for (int i = tid; i < dc * NSH; i += kThreads) ywl[i] = a.Yw[(int64_t)(mbase + i / NSH) * NSH + i % NSH];
  const int cpr = kThreads / dc;  // cells per round of the solve phase
  const int my_d = tid % dc;
  const bool solver = tid < cpr * dc;
  float y[NSH];
#pragma unroll
  for (int n = 0; n < NSH; ++n) y[n] = HAS_T ? a.Y[(int64_t)(mbase + my_d) * NSH + n] : 0.0f;
  const float cx = a.cdir[(mbase + my_d) * 3 + 0];
  const float cy = a.cdir[(mbase + my_d) * 3 + 1];
  const float cz = a.cdir[(mbase + my_d) * 3 + 2];
  {
    const int ub[3][3] = {{bi, bj - sx, bk}, {bi - sy, bj, bk}, {bi, bj, bk - sz}};
    for (int ax = 0; ax < 3; ++ax) {
      const int ui = ub[ax][0], uj = ub[ax][1], uk = ub[ax][2];
      const bool inside = ui >= 0 && ui < a.nbr && uj >= 0 && uj < a.nbc && uk >= 0 && uk < a.nbs;
      float* dst = face + ax * face_stride;
      if (inside) {
        const float* srcf = gface_v + ((int64_t)((ui * a.nbc + uj) * a.nbs + uk) * 3 + ax) * face_stride;
        for (int i = tid; i < face_stride; i += kThreads) dst[i] = srcf[i];
      } else {
        for (int i = tid; i < face_stride; i += kThreads) dst[i] = 0.0f;
      }
    }
  }
  __syncthreads();

  // ---- cell planes --------------------------------------------------------------------------------------------
  // Every global load below is unconditional (clamped to a valid address, the value discarded when unused), issued
  // in a fixed order, and the plane loop is fully unrolled, so the compiler counts outstanding loads exactly: the
  // moments T and mu_t of plane p+2 and the old fluence of plane p are in flight during the solve of plane p, and
  // the barriers (lds_barrier) order LDS only.
  const float4* T4 = reinterpret_cast<const float4*>(T_g);
  float4* tpl4 = reinterpret_cast<float4*>(tpl);
  struct Staged {  // one cell plane's prefetched inputs, held in registers
    f4 t[TPT];
    float mu;
  };
  f4 ppre[RPT][KPER];
  const f4* T4v = reinterpret_cast<const f4*>(T_g);

  auto load_plane = [&](const int pl) __attribute__((always_inline)) -> Staged {
    Staged st;
    const int pb = ptab.offset[pl];
    const int P = ptab.offset[pl + 1] - pb;
    {
      const int c = slot_cell[pb + (tid < P ? tid : P - 1)];
      st.mu = mu_g[c >= 0 ? c : 0];
    }
#pragma unroll
    for (int r = 0; r < TPT; ++r) {
      if (HAS_T) {
        int i = tid + r * kThreads;
        i = i < P * NSH ? i : P * NSH - 1;
        const int ci = i / NSH, n = i - ci * NSH;
        const int c = slot_cell[pb + ci];
        st.t[r] = T4v[(int64_t)(c >= 0 ? c : 0) * NSH + n];
      } else {
        st.t[r] = f4{0.f, 0.f, 0.f, 0.f};
      }
    }
    return st;
  };

  // stages plane p from `cur`, runs it, and returns the prefetched inputs of plane p + 2
  auto plane = [&](const int p, const Staged cur) __attribute__((always_inline)) -> Staged {
    const int which = p & 1;
    const int pbeg = ptab.offset[p];
    const int P = ptab.offset[p + 1] - pbeg;
    // (1) stage this plane's mu_t and T in LDS
    if (tid < P) tmu[which * PMAX + tid] = cur.mu;
    if (HAS_T) {
#pragma unroll
      for (int r = 0; r < TPT; ++r) {
        const int i = tid + r * kThreads;
        if (i < P * NSH) reinterpret_cast<f4*>(tpl4)[i + i / NSH] = cur.t[r];  // cell stride NSH + 1: no conflicts
      }
    }
    // (2) request the fluence moments this plane adds to, then the next-but-one plane
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      int i = tid + r * kThreads;
      i = i < P * KCH ? i : P * KCH - 1;
      const int ci = i / KCH, n0 = (i - ci * KCH) * KPER;
      const int c = slot_cell[pbeg + ci];
      const f4* src = reinterpret_cast<const f4*>(phi_g + (int64_t)(c >= 0 ? c : 0) * NSH * 4);
#pragma unroll
      for (int j = 0; j < KPER; ++j) ppre[r][j] = src[(n0 + j < NSH) ? n0 + j : NSH - 1];
    }
    // (the last two planes reload the final plane: a few wasted loads instead of a select between aggregates)
    const Staged next = load_plane(p + 2 < NPL ? p + 2 : NPL - 1);
    lds_barrier();
    // (3) solve every (cell, direction) pair of the plane
    if (solver) {
      for (int ci = tid / dc; ci < P; ci += cpr) {
        const int cell = slot_cell[pbeg + ci];
        const int packed = slot_pos[pbeg + ci];
        const int lx = packed & 0xff, ly = (packed >> 8) & 0xff, lz = (packed >> 16) & 0xff;
        float4 pout = make_float4(0.f, 0.f, 0.f, 0.f);
        if (cell >= 0) {
          float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
          if (HAS_T) {
            const float4* t4 = tpl4 + ci * (NSH + 1);
#pragma unroll
            for (int n = 0; n < NSH; ++n) {
              const float4 t = t4[n];
              s0 = fmaf(y[n], t.x, s0);
              s1 = fmaf(y[n], t.y, s1);
              s2 = fmaf(y[n], t.z, s2);
              s3 = fmaf(y[n], t.w, s3);
            }
          }
          float4* sp = nullptr;
          if (HAS_SRC && a.src_layout == 2) {  // cell means only: the slopes are zero
            s0 += src_v[((int64_t)g * a.N + cell) * a.M + (o * a.D8 + a.d0 + my_d)];
          } else if (HAS_SRC) {
            sp = reinterpret_cast<float4*>(
                src_v + (a.src_layout == 0 ? (((int64_t)(o * dc + my_d) * a.N + cell) * a.G + g) * 4
                                           : (((int64_t)g * a.N + cell) * a.M + (o * a.D8 + a.d0 + my_d)) * 4));
            const float4 q = *sp;
            s0 += q.x; s1 += q.y; s2 += q.z; s3 += q.w;
          }
          // face arrays are [component][position][direction]: the lanes of one cell read consecutive words;
          // component 0 is the face average psi_0 + s_a psi_a, 1 and 2 the transverse moments
          const int cs = BB * dc;
          float* fx = face + 0 * face_stride + (ly * B + lz) * dc + my_d;
          float* fy = face + 1 * face_stride + (lx * B + lz) * dc + my_d;
          float* fz = face + 2 * face_stride + (lx * B + ly) * dc + my_d;
          const float pbx = fx[0], xy = fx[cs], xz = fx[2 * cs];
          const float pby = fy[0], yx = fy[cs], yz = fy[2 * cs];
          const float pbz = fz[0], zx = fz[cs], zy = fz[2 * cs];
          const float mu = tmu[which * PMAX + ci];
          const float r0 = s0 + (cx * pbx + cy * pby + cz * pbz);
          const float rx = s1 - 3.0f * fsx * cx * pbx + (cy * yx + cz * zx);
          const float ry = s2 - 3.0f * fsy * cy * pby + (cx * xy + cz * zy);
          const float rz = s3 - 3.0f * fsz * cz * pbz + (cx * xz + cy * yz);
          const float csum = cx + cy + cz;
          const float dx = 3.0f * cx + (csum - cx) + mu;
          const float dy = 3.0f * cy + (csum - cy) + mu;
          const float dz = 3.0f * cz + (csum - cz) + mu;
          const float qx = cx / dx, qy = cy / dy, qz = cz / dz;
          const float den = csum + mu + 3.0f * (cx * qx + cy * qy + cz * qz);
          const float p0 = (r0 - (fsx * qx * rx + fsy * qy * ry + fsz * qz * rz)) / den;
          const float px = (rx + 3.0f * fsx * cx * p0) / dx;
          const float py = (ry + 3.0f * fsy * cy * p0) / dy;
          const float pz = (rz + 3.0f * fsz * cz * p0) / dz;
          fx[0] = p0 + fsx * px; fx[cs] = py; fx[2 * cs] = pz;
          fy[0] = p0 + fsy * py; fy[cs] = px; fy[2 * cs] = pz;
          fz[0] = p0 + fsz * pz; fz[cs] = px; fz[2 * cs] = py;
          pout = make_float4(p0, px, py, pz);
          if (HAS_SRC && a.write_psi) *sp = pout;
        }
        reinterpret_cast<float4*>(psl)[ci * dc + my_d] = pout;
      }
    }
    lds_barrier();
    // (4) reduce the plane's angular fluence to moments and add them into phi
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int i = tid + r * kThreads;
      if (i >= P * KCH) continue;
      const int ci = i / KCH, n0 = (i - ci * KCH) * KPER;
      const int cell = slot_cell[pbeg + ci];
      float4 acc[KPER];
#pragma unroll
      for (int j = 0; j < KPER; ++j) acc[j] = make_float4(0.f, 0.f, 0.f, 0.f);
      const float4* ps = reinterpret_cast<const float4*>(psl) + ci * dc;
      for (int d = 0; d < dc; ++d) {
        const float4 q = ps[d];
#pragma unroll
        for (int j = 0; j < KPER; ++j) {
          const float w = (n0 + j < NSH) ? ywl[d * NSH + n0 + j] : 0.0f;
          acc[j].x = fmaf(w, q.x, acc[j].x);
          acc[j].y = fmaf(w, q.y, acc[j].y);
          acc[j].z = fmaf(w, q.z, acc[j].z);
          acc[j].w = fmaf(w, q.w, acc[j].w);
        }
      }
      if (cell >= 0) {
        float4* dst = reinterpret_cast<float4*>(phi_g + (int64_t)cell * NSH * 4);
#pragma unroll
        for (int j = 0; j < KPER; ++j) {
          if (n0 + j < NSH) {
            float4 val = acc[j];
            if (!first_visit) {
              val.x += ppre[r][j].x; val.y += ppre[r][j].y; val.z += ppre[r][j].z; val.w += ppre[r][j].w;
            }
            dst[n0 + j] = val;
          }
        }
      }
    }
    // the next plane writes tpl and tmu (of the other parity) before its first barrier: tpl was last read before the
    // barrier above, and psl, which the reduction reads, is next written after the next barrier
    return next;
  };

  Staged sa = load_plane(0);
  Staged sb = load_plane(NPL > 1 ? 1 : 0);
#pragma unroll
  for (int p = 0; p < NPL; p += 2) {
    sa = plane(p, sa);
    if (p + 1 < NPL) sb = plane(p + 1, sb);
  }
  __syncthreads();

  // ---- outgoing faces for the downstream blocks ---------------------------------------------------------------
  {
    const int db[3][3] = {{bi, bj + sx, bk}, {bi + sy, bj, bk}, {bi, bj, bk + sz}};
    for (int ax = 0; ax < 3; ++ax) {
      const int di = db[ax][0], dj = db[ax][1], dk = db[ax][2];
      if (!(di >= 0 && di < a.nbr && dj >= 0 && dj < a.nbc && dk >= 0 && dk < a.nbs)) continue;
      float* dst = gface_v + ((int64_t)blk * 3 + ax) * face_stride;
      const float* srcf = face + ax * face_stride;
      for (int i = tid; i < face_stride; i += kThreads) dst[i] = srcf[i];
    }
  }
}

// Launch-sequence form of the matrix-core sweep: one workgroup per (task of the launch's block diagonal, group,
// view), decoded from the block index.
template <int NSH, int DC4, int SRC, bool BF16>
__global__ __launch_bounds__(64) void ld_sweep_mfma_kernel(LaunchArgs a) {
  extern __shared__ float lds[];
  constexpr int KH = 16 * ((NSH - 4 + 15) / 16);
  const int task = blockIdx.x % a.ntask;
  const int rest = blockIdx.x / a.ntask;
  const int g = rest % a.G;
  const int v = rest / a.G;
  const int4 e = a.tasks[task];
  const int slot = (e.x >> 4) & 1;
  MfmaTask t;
  t.R = a.R; t.C = a.C; t.S = a.S; t.N = a.N;
  t.nbr = a.nbr; t.nbc = a.nbc; t.nbs = a.nbs;
  t.o = e.x & 7;
  t.bi = e.y; t.bj = e.z; t.bk = e.w;
  t.first_visit = a.first && ((e.x >> 3) & 1);
  t.gface = a.faces + (int64_t)v * a.fview + ((int64_t)slot * a.G + g) * a.nblocks * 3 * (48 * a.dc);
  t.mu = a.mu + (int64_t)g * a.N;
  if (BF16) {
    t.T = a.T16 + (int64_t)v * a.t16view + (int64_t)g * a.N * 4 * KH;
    t.Tlo = a.Tlo + ((int64_t)v * a.G + g) * a.N * 16;
  } else {
    t.T = a.T + (int64_t)v * a.tview + (int64_t)g * a.N * NSH * 4;
    t.Tlo = nullptr;
  }
  t.phi = a.phi + (int64_t)v * a.tview + (int64_t)g * a.N * NSH * 4;
  t.src = SRC != 0 ? a.src + (int64_t)v * a.sview + (int64_t)g * a.N * a.M * (SRC == 2 ? 1 : 4) : nullptr;
  t.cdir = a.cdir;
  t.fc_scale = t.fc_beta = t.fc_u = nullptr;  // source layout 3 runs on the persistent sweep only
  mfma_task<NSH, DC4, SRC, BF16>(t, DirTables{a.D8, a.d0, a.dc, a.M, a.Y, a.Yw}, lds, threadIdx.x);
}


template <int NSH, int HAS_SRC, bool BF16>
hipError_t launch_mfma_src(const LaunchArgs& la, int nwg, hipStream_t stream) {
  const size_t lds = mfma_lds_bytes(la.dc);
  const dim3 grid(nwg), block(64);
  switch ((la.dc + 3) / 4) {
    case 1: hipLaunchKernelGGL((ld_sweep_mfma_kernel<NSH, 1, HAS_SRC, BF16>), grid, block, lds, stream, la); break;
    case 2: hipLaunchKernelGGL((ld_sweep_mfma_kernel<NSH, 2, HAS_SRC, BF16>), grid, block, lds, stream, la); break;
    case 3: hipLaunchKernelGGL((ld_sweep_mfma_kernel<NSH, 3, HAS_SRC, BF16>), grid, block, lds, stream, la); break;
    case 4: hipLaunchKernelGGL((ld_sweep_mfma_kernel<NSH, 4, HAS_SRC, BF16>), grid, block, lds, stream, la); break;
    case 5: hipLaunchKernelGGL((ld_sweep_mfma_kernel<NSH, 5, HAS_SRC, BF16>), grid, block, lds, stream, la); break;
    case 6: hipLaunchKernelGGL((ld_sweep_mfma_kernel<NSH, 6, HAS_SRC, BF16>), grid, block, lds, stream, la); break;
    case 7: hipLaunchKernelGGL((ld_sweep_mfma_kernel<NSH, 7, HAS_SRC, BF16>), grid, block, lds, stream, la); break;
    case 8: hipLaunchKernelGGL((ld_sweep_mfma_kernel<NSH, 8, HAS_SRC, BF16>), grid, block, lds, stream, la); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int NSH, bool BF16>
hipError_t launch_mfma_layout(const LaunchArgs& la, int nwg, hipStream_t stream) {
  if (!la.src) return launch_mfma_src<NSH, 0, BF16>(la, nwg, stream);
  return la.src_layout == 2 ? launch_mfma_src<NSH, 2, BF16>(la, nwg, stream)
                            : launch_mfma_src<NSH, 1, BF16>(la, nwg, stream);
}

template <int NSH>
hipError_t launch_mfma_nsh(const LaunchArgs& la, int nwg, hipStream_t stream) {
  return launch_mfma_layout<NSH, false>(la, nwg, stream);
}

// the bfloat16 expansion is built for P3 and P5 (16 and 36 moments)
template <int NSH>
hipError_t launch_mfma_bf16(const LaunchArgs& la, int nwg, hipStream_t stream) {
  return launch_mfma_layout<NSH, true>(la, nwg, stream);
}

hipError_t launch_mfma(int nsh, const LaunchArgs& la, int nwg, hipStream_t stream) {
  if (la.T16) {
    if (nsh == 16) return launch_mfma_bf16<16>(la, nwg, stream);
    if (nsh == 36) return launch_mfma_bf16<36>(la, nwg, stream);
    return hipErrorInvalidValue;
  }
  switch (nsh) {
    case 1: return launch_mfma_nsh<1>(la, nwg, stream);
    case 4: return launch_mfma_nsh<4>(la, nwg, stream);
    case 9: return launch_mfma_nsh<9>(la, nwg, stream);
    case 16: return launch_mfma_nsh<16>(la, nwg, stream);
    case 25: return launch_mfma_nsh<25>(la, nwg, stream);
    case 36: return launch_mfma_nsh<36>(la, nwg, stream);
    default: return hipErrorInvalidValue;
  }
}

// The matrix-core kernel serves the moment-source sweep unless CITRUS_SWEEP_KERNEL=vector selects the vector kernel
// (both give the same results to FP32 rounding; the switch exists to compare them). Its face-buffer layout differs
// from the vector kernel's, so the choice is made once per sweep.
bool use_mfma(const LaunchArgs& la, int nsh) {
  const char* sel = std::getenv("CITRUS_SWEEP_KERNEL");
  const bool vector_only = sel && std::string(sel) == "vector";
  // moments of one (view, group), and its group-major source, addressable by a buffer resource
  return !vector_only && (la.T || la.T16) && (!la.src || la.src_layout != 0) && nsh <= 36 && la.dc <= 32 &&
         !la.write_psi &&
         (int64_t)la.N * nsh * 16 < (int64_t)kDrop && (!la.src || (int64_t)la.N * la.M * 16 < (int64_t)kDrop);
}

template <int B>
constexpr int plane_max() {
  return PlaneTable<B>().max_plane();
}

template <int NSH, int B>
size_t lds_bytes(int dc) {
  return sizeof(float) * (size_t)(9 * dc * B * B + plane_max<B>() * (NSH + 1) * 4 + plane_max<B>() * dc * 4 +
                                  dc * NSH + 2 * B * B * B + 2 * plane_max<B>());
}

constexpr size_t kMaxLds = 64 * 1024;

template <int NSH, int B>
hipError_t launch_one(const LaunchArgs& la, int nwg, hipStream_t stream) {
  const size_t lds = lds_bytes<NSH, B>(la.dc);
  if (lds > kMaxLds) return hipErrorInvalidConfiguration;
  if (la.T && !la.src)
    hipLaunchKernelGGL((ld_sweep_kernel<NSH, B, true, false>), dim3(nwg), dim3(kThreads), lds, stream, la);
  else if (la.T)
    hipLaunchKernelGGL((ld_sweep_kernel<NSH, B, true, true>), dim3(nwg), dim3(kThreads), lds, stream, la);
  else
    hipLaunchKernelGGL((ld_sweep_kernel<NSH, B, false, true>), dim3(nwg), dim3(kThreads), lds, stream, la);
  return hipGetLastError();
}

template <int B>
hipError_t launch_nsh(int nsh, const LaunchArgs& la, int nwg, hipStream_t stream) {
  switch (nsh) {
    case 1: return launch_one<1, B>(la, nwg, stream);
    case 4: return launch_one<4, B>(la, nwg, stream);
    case 9: return launch_one<9, B>(la, nwg, stream);
    case 16: return launch_one<16, B>(la, nwg, stream);
    case 25: return launch_one<25, B>(la, nwg, stream);
    case 36: return launch_one<36, B>(la, nwg, stream);
    case 49: return launch_one<49, B>(la, nwg, stream);
    case 64: return launch_one<64, B>(la, nwg, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace
}  // namespace citrus

// ======================================================================================================================
// schedule
// ======================================================================================================================

struct CitrusSweepPlan {
  int R, C, S, B;
  int nbr, nbc, nbs, nblocks;
  int K;  // sum of (blocks - 1) over the axes: block planes run 0..K
  // launches in order: (offset into the task table, count)
  std::vector<std::pair<int, int>> launches;
  std::vector<int> launch_pass;  // pass (octant pair 0..3) of each launch
  int4* d_tasks = nullptr;
  int ntasks = 0;
};

namespace {

int block_plane(int o, int bi, int bj, int bk, int nbr, int nbc, int nbs) {
  const int a = (o & 4) ? nbc - 1 - bj : bj;
  const int b = (o & 2) ? nbr - 1 - bi : bi;
  const int c = (o & 1) ? nbs - 1 - bk : bk;
  return a + b + c;
}

}  // namespace

extern "C" int citrus_sweep_plan_create(int R, int C, int S, int block, CitrusSweepPlan** out) {
  using namespace citrus;
  if (!out) return fail(CITRUS_ERR_ARG, "plan pointer is NULL");
  *out = nullptr;
  if (R < 1 || C < 1 || S < 1) return fail(CITRUS_ERR_ARG, "grid shape must be positive");
  if (block != 4) return fail(CITRUS_ERR_UNSUPPORTED, "block edge must be 4");
  auto* plan = new CitrusSweepPlan();
  plan->R = R; plan->C = C; plan->S = S; plan->B = block;
  plan->nbr = (R + block - 1) / block;
  plan->nbc = (C + block - 1) / block;
  plan->nbs = (S + block - 1) / block;
  plan->nblocks = plan->nbr * plan->nbc * plan->nbs;
  plan->K = (plan->nbr - 1) + (plan->nbc - 1) + (plan->nbs - 1);
  const int K = plan->K;
  std::vector<int4> table;
  for (int pass = 0; pass < 4; ++pass) {
    const int oa = pass, ob = 7 - pass;
    for (int d = 0; d <= K; ++d) {
      // octant oa then octant ob; on the middle plane of an even K each gets its own launch (oa first)
      const bool split = (2 * d == K);
      for (int part = 0; part < (split ? 2 : 1); ++part) {
        const int start = (int)table.size();
        for (int which = 0; which < 2; ++which) {
          if (split && which != part) continue;
          const int o = which == 0 ? oa : ob;
          for (int bi = 0; bi < plan->nbr; ++bi)
            for (int bj = 0; bj < plan->nbc; ++bj)
              for (int bk = 0; bk < plan->nbs; ++bk) {
                const int dd = block_plane(o, bi, bj, bk, plan->nbr, plan->nbc, plan->nbs);
                if (dd != d) continue;
                // the pair's first visit of this block: the octant whose block plane comes first (oa on a tie)
                const bool first = which == 0 ? (2 * d <= K) : (2 * d < K);
                table.push_back(make_int4(o | ((first ? 1 : 0) << 3) | (which << 4), bi, bj, bk));
              }
        }
        const int count = (int)table.size() - start;
        if (count > 0) {
          plan->launches.emplace_back(start, count);
          plan->launch_pass.push_back(pass);
        }
      }
    }
  }
  plan->ntasks = (int)table.size();
  hipError_t e = hipMalloc(&plan->d_tasks, sizeof(int4) * table.size());
  if (e == hipSuccess) e = hipMemcpy(plan->d_tasks, table.data(), sizeof(int4) * table.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (plan->d_tasks) (void)hipFree(plan->d_tasks);
    delete plan;
    return fail(CITRUS_ERR_HIP, std::string("sweep plan upload: ") + hipGetErrorString(e));
  }
  *out = plan;
  return CITRUS_OK;
}

extern "C" void citrus_sweep_plan_destroy(CitrusSweepPlan* plan) {
  if (!plan) return;
  if (plan->d_tasks) (void)hipFree(plan->d_tasks);
  delete plan;
}

extern "C" int64_t citrus_sweep_face_floats(const CitrusSweepPlan* plan, int G, int dc, int V) {
  if (!plan || G < 1 || dc < 1 || V < 1) return -1;
  return (int64_t)V * 2 * G * plan->nblocks * 3 * (int64_t)dc * plan->B * plan->B * 3;
}

extern "C" int citrus_sweep_run(const CitrusSweepPlan* plan, const CitrusSweepArgs* args, void* stream) {
  using namespace citrus;
  if (!plan || !args) return fail(CITRUS_ERR_ARG, "plan or args is NULL");
  const CitrusSweepArgs& A = *args;
  if (A.G < 1 || A.V < 1) return fail(CITRUS_ERR_ARG, "G and V must be >= 1");
  if (A.dc < 1 || A.dc > 32) return fail(CITRUS_ERR_ARG, "dc must be in 1..32");
  if (A.d0 < 0 || A.d0 + A.dc > A.D8) return fail(CITRUS_ERR_ARG, "chunk [d0, d0+dc) outside the octant");
  if (!A.mu_t || !A.phi || !A.Y || !A.Yw || !A.cdir || !A.faces) return fail(CITRUS_ERR_ARG, "NULL table or buffer");
  if (!A.T && !A.T_bf16 && !A.source) return fail(CITRUS_ERR_ARG, "need T moments and/or an explicit source");
  if (A.write_psi && !A.source) return fail(CITRUS_ERR_ARG, "write_psi needs the explicit source buffer");
  if (A.source_layout < 0 || A.source_layout > 2) return fail(CITRUS_ERR_ARG, "source_layout must be 0, 1 or 2");
  if (A.T_bf16 && (!A.T_low || A.T)) return fail(CITRUS_ERR_ARG, "T_bf16 needs T_low and no float32 T");
  if (A.T_bf16 && A.nsh != 16 && A.nsh != 36)
    return fail(CITRUS_ERR_UNSUPPORTED, "the bfloat16 contraction is built for P3 and P5 (nsh 16 and 36)");
  if (A.source_layout != 0 && A.write_psi) return fail(CITRUS_ERR_ARG, "write_psi needs the cell-major layout");
  const int N = plan->R * plan->C * plan->S;
  LaunchArgs la{};
  la.R = plan->R; la.C = plan->C; la.S = plan->S; la.N = N;
  la.nbr = plan->nbr; la.nbc = plan->nbc; la.nbs = plan->nbs; la.nblocks = plan->nblocks;
  la.G = A.G; la.D8 = A.D8; la.d0 = A.d0; la.dc = A.dc; la.V = A.V;
  la.write_psi = A.write_psi;
  la.mu = A.mu_t; la.T = A.T; la.src = A.source; la.phi = A.phi;
  la.Y = A.Y; la.Yw = A.Yw; la.cdir = A.cdir; la.faces = A.faces;
  la.tview = (int64_t)A.G * N * A.nsh * 4;
  la.src_layout = A.source_layout;
  la.M = 8 * A.D8;
  la.sview = A.source_layout == 0 ? (int64_t)8 * A.dc * N * A.G * 4
                                   : (int64_t)A.G * N * la.M * (A.source_layout == 1 ? 4 : 1);
  la.fview = (int64_t)2 * A.G * plan->nblocks * 3 * (int64_t)A.dc * plan->B * plan->B * 3;
  la.T16 = reinterpret_cast<const uint16_t*>(A.T_bf16);
  la.Tlo = A.T_low;
  la.t16view = (int64_t)A.G * N * 4 * ((A.nsh - 4 + 15) / 16 * 16);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  la.ntask = plan->launches.empty() ? 0 : plan->launches[0].second;
  const bool mfma = use_mfma(la, A.nsh);
  if (la.T16 && !mfma) return fail(CITRUS_ERR_UNSUPPORTED, "the bfloat16 contraction runs on the matrix-core kernel only");
  for (size_t l = 0; l < plan->launches.size(); ++l) {
    la.tasks = plan->d_tasks + plan->launches[l].first;
    la.ntask = plan->launches[l].second;
    la.first = (!A.accumulate && plan->launch_pass[l] == 0) ? 1 : 0;
    const int64_t nwg = (int64_t)la.ntask * A.G * A.V;
    if (!grid_fits(nwg, kThreads)) return fail(CITRUS_ERR_ARG, "too many workgroups in one launch");
    hipError_t e = mfma ? launch_mfma(A.nsh, la, (int)nwg, s) : launch_nsh<4>(A.nsh, la, (int)nwg, s);
    if (e == hipErrorInvalidValue) return fail(CITRUS_ERR_UNSUPPORTED, "nsh must be (L+1)^2 with L <= 7");
    if (e == hipErrorInvalidConfiguration) return fail(CITRUS_ERR_UNSUPPORTED, "chunk needs more than 64 KiB of LDS");
    if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("sweep launch: ") + hipGetErrorString(e));
  }
  return CITRUS_OK;
}
