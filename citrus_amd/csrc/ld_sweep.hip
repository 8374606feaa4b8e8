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

namespace citrus {
namespace {

constexpr int kThreads = 128;

// Native 4-wide float vector for register-resident prefetch buffers (HIP's float4 wrapper is copied through memory,
// which keeps arrays of it out of registers).
typedef float f4 __attribute__((ext_vector_type(4)));

// Workgroup barrier that orders LDS accesses only: outstanding global loads stay in flight across it (a full
// __syncthreads also waits for them). Every cross-thread exchange in the sweep kernel goes through LDS.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }


// Cells of a B^3 block grouped by diagonal plane a + b + c = p, in sweep-local coordinates.
template <int B>
struct PlaneTable {
  static constexpr int kPlanes = 3 * B - 2;
  int offset[3 * B - 1];
  int cell[B * B * B];  // a | b << 8 | c << 16  (a along x, b along y, c along z)
  constexpr PlaneTable() : offset(), cell() {
    int n = 0;
    for (int p = 0; p < kPlanes; ++p) {
      offset[p] = n;
      for (int a = 0; a < B; ++a)
        for (int b = 0; b < B; ++b) {
          int c = p - a - b;
          if (c >= 0 && c < B) cell[n++] = a | (b << 8) | (c << 16);
        }
    }
    offset[kPlanes] = n;
  }
  constexpr int max_plane() const {
    int m = 0;
    for (int p = 0; p < kPlanes; ++p) m = (offset[p + 1] - offset[p]) > m ? (offset[p + 1] - offset[p]) : m;
    return m;
  }
};

__constant__ PlaneTable<4> kPlanes4 = PlaneTable<4>();

template <int B>
__device__ __forceinline__ const PlaneTable<B>& planes();
template <>
__device__ __forceinline__ const PlaneTable<4>& planes<4>() { return kPlanes4; }

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
  const uint16_t* T16;     // bfloat16 operand of the source expansion [V][G][N][4][KP] (or NULL)
  const float* tinv;       // [V][G] scale undoing the operand's per-group power of two
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

// ---------------------------------------------------------------------------------------------------------------------
// Matrix-core variant for the moment-source sweep (T given, plus an optional group-major explicit source), B = 4,
// NSH <= 36, dc <= 32.
//
// One wave sweeps one (view, octant, group, block) task. Both contractions of a cell plane run on
// v_mfma_f32_16x16x4_f32 in full FP32, four cells (16 rows: cell x LD component) per M-tile:
//   source expansion  S[(cell,k)][d] = sum_n T[(cell,k)][n] Y[d][n]        (K = SH moments, N = directions)
//   moment reduction  phi[(cell,k)][n] += sum_d psi[(cell,k)][d] Yw[d][n]  (K = directions, N = SH moments)
// Operand layouts of the 16x16x4 f32 instruction (lane l): A[i = l%16][k = l/16], B[k = l/16][j = l%16],
// D[i = 4(l/16) + r][j = l%16] in result register r. So an A fragment of T is one scalar load from the cell-major
// moments, the expansion leaves all four LD components of one (cell, direction) in one lane for the closed-form solve,
// and the reduction leaves phi[cell][n][0..3], one float4 of the moment array, in one lane. Only the angular fluence
// changes layout between the two (through a 2.2 KB LDS tile). The whole task is one wave, so cell planes need no
// barrier: LDS accesses of one wave execute in program order.
//
// The kernel body has no branches: moments are read and written through buffer resources, whose out-of-range
// accesses return zero or are dropped, and lanes without a cell or direction send their face updates to a private
// scratch row of LDS. The fully unrolled plane loop is then a single block the compiler can schedule across, so the
// matrix-core chains of one M-tile overlap the solve of another and the loads of the next plane.
// Face arrays (LDS and the inter-block buffer) are [axis][position][component][direction].
// ---------------------------------------------------------------------------------------------------------------------

constexpr int kMfmaPsiStride = 136;  // floats per cell row of the psi tile ([d][k], 32 directions + 8): == 8 mod 32
// Buffer-resource word 3 of gfx9-family raw buffers; accesses at kDrop or beyond num_records (< kDrop) read zero and
// write nothing.
constexpr int kBufferWord3 = 0x00020000;
constexpr unsigned kDrop = 0x80000000u;
typedef unsigned u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

typedef short s4 __attribute__((ext_vector_type(4)));  // four bfloat16 bit patterns
typedef unsigned u2 __attribute__((ext_vector_type(2)));

// D += A B with A [16][16] and B [16][16] bfloat16 (lane l: A[l%16][4(l/16)..+3], B[4(l/16)..+3][l%16]), fp32 D
__device__ __forceinline__ f4 mfma_bf16(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ short bf16_bits(float f) {  // round to nearest even
  const unsigned b = __float_as_uint(f);
  return (short)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

__device__ __forceinline__ f4 buf_load4(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}

__device__ __forceinline__ s4 buf_load_s4(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(s4, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

__device__ __forceinline__ void buf_store4(__amdgpu_buffer_rsrc_t r, unsigned off, f4 v) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, (int)off, 0, 0);
}

// SRC: 0 no explicit source; 1 group-major LD moments (layout 1, read per plane); 2 group-major cell means (layout 2,
// prefetched one plane ahead with T, the slopes zero)
template <int NSH, int DC4, int SRC, bool BF16>
__global__ __launch_bounds__(64) void ld_sweep_mfma_kernel(LaunchArgs a) {
  constexpr bool HAS_SRC = SRC != 0;
  constexpr int B = 4;
  constexpr PlaneTable<B> ptab{};
  constexpr int NPL = PlaneTable<B>::kPlanes;
  constexpr int NT = (DC4 + 3) / 4;     // direction tiles of the expansion
  constexpr int KS = BF16 ? 1 : (NSH + 3) / 4;  // k-steps of the float32 expansion
  constexpr int KB = BF16 ? (NSH + 15) / 16 : 1;  // k-steps of the bfloat16 expansion (16 moments each)
  constexpr int KP = 16 * ((NSH + 15) / 16);      // moments per row of the bfloat16 operand
  constexpr int NT2 = (NSH + 15) / 16;  // moment tiles of the reduction
  constexpr int PS = kMfmaPsiStride;
  constexpr unsigned CB = NSH * 16;     // bytes of one cell's moments
  const PlaneTable<B>& tab = planes<B>();

  extern __shared__ float lds[];
  const int dc = a.dc;
  const int face_stride = 48 * dc;                           // floats per axis: [16 positions][3 components][dc]
  float* face = lds;                                         // [3 axes] of the above
  float* psl = face + 3 * face_stride;                       // [4 cells][PS]
  int* slot_cell = reinterpret_cast<int*>(psl + 4 * PS);     // [64] flat cell of each slot, or -1
  int* slot_pos = slot_cell + 64;                            // [64] local (x, y, z) of each slot, packed
  float* scratch_row = reinterpret_cast<float*>(slot_pos + 64);  // [2 dc + 64] face updates of padding lanes

  const int lane = threadIdx.x;
  const int q = lane >> 4, r16 = lane & 15;
  const int acell = r16 >> 2, acomp = r16 & 3;  // A-fragment row: (cell of the M-tile, LD component)
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

  float* gface_v = a.faces + (int64_t)v * a.fview + ((int64_t)slot * a.G + g) * a.nblocks * 3 * face_stride;
  const __amdgpu_buffer_rsrc_t mu_r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(a.mu + (int64_t)g * a.N), 0, a.N * 4, kBufferWord3);
  const __amdgpu_buffer_rsrc_t T_r =
      BF16 ? __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(a.T16 + (int64_t)v * a.t16view +
                                                                      (int64_t)g * a.N * 4 * KP),
                                               0, a.N * 8 * KP, kBufferWord3)
           : __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(a.T + (int64_t)v * a.tview +
                                                                   (int64_t)g * a.N * NSH * 4),
                                               0, a.N * (int)CB, kBufferWord3);
  const float tinv = BF16 ? a.tinv[v * a.G + g] : 1.0f;
  const __amdgpu_buffer_rsrc_t phi_r = __builtin_amdgcn_make_buffer_rsrc(
      a.phi + (int64_t)v * a.tview + (int64_t)g * a.N * NSH * 4, 0, a.N * (int)CB, kBufferWord3);
  // explicit source, group-major [G][N][M][4] (layout 1: the 16 directions of one cell in a D tile are 256 contiguous
  // bytes) or cell means [G][N][M] (layout 2, zero slopes: 64 contiguous bytes)
  constexpr int src_words = SRC == 2 ? 1 : 4;
  const __amdgpu_buffer_rsrc_t src_r = __builtin_amdgcn_make_buffer_rsrc(
      HAS_SRC ? a.src + (int64_t)v * a.sview + (int64_t)g * a.N * a.M * src_words : a.phi, 0,
      HAS_SRC ? a.N * a.M * 4 * src_words : 0, kBufferWord3);

  // ---- slot tables, incoming faces (all requested before the first is used), direction fragments ------------------
  {
    const int packed = tab.cell[lane];
    const int pa = packed & 0xff, pb = (packed >> 8) & 0xff, pc = (packed >> 16) & 0xff;
    const int lx = sx > 0 ? pa : B - 1 - pa;
    const int ly = sy > 0 ? pb : B - 1 - pb;
    const int lz = sz > 0 ? pc : B - 1 - pc;
    const int x = x0 + lx, yy = y0 + ly, z = z0 + lz;
    slot_cell[lane] = (x >= a.C || yy >= a.R || z >= a.S) ? -1 : (yy * a.C + x) * a.S + z;
    slot_pos[lane] = lx | (ly << 8) | (lz << 16);
  }
  constexpr int FPL = 3 * NT;  // float4 per lane and axis: face_stride / 4 = 12 dc <= 192 NT
  const int n4 = face_stride / 4;
  {
    // a block without an upstream neighbour along an axis reads its own (valid) entries and stores zeros instead
    const int ub[3][3] = {{bi, bj - sx, bk}, {bi - sy, bj, bk}, {bi, bj, bk - sz}};
    f4 fin[3][FPL];
    bool inside[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const int ui = ub[ax][0], uj = ub[ax][1], uk = ub[ax][2];
      inside[ax] = ui >= 0 && ui < a.nbr && uj >= 0 && uj < a.nbc && uk >= 0 && uk < a.nbs;
      const int ublk = inside[ax] ? (ui * a.nbc + uj) * a.nbs + uk : blk;
      const f4* srcf = reinterpret_cast<const f4*>(gface_v + ((int64_t)ublk * 3 + ax) * face_stride);
#pragma unroll
      for (int r = 0; r < FPL; ++r) {
        const int i = lane + 64 * r;
        fin[ax][r] = srcf[i < n4 ? i : n4 - 1];
      }
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      f4* dst = reinterpret_cast<f4*>(face + ax * face_stride);
#pragma unroll
      for (int r = 0; r < FPL; ++r) {
        const int i = lane + 64 * r;
        if (i < n4) dst[i] = inside[ax] ? fin[ax][r] : f4{0.f, 0.f, 0.f, 0.f};
      }
    }
  }
  const int mbase = o * a.D8 + a.d0;
  float yf[NT][KS];  // expansion B fragments: Y[d = 16 nt + l%16][n = 4 s + l/16]
  s4 yb[NT][KB];     // bfloat16 expansion B fragments: Y[d = 16 nt + l%16][n = 16 s + 4 (l/16) + 0..3]
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int d = nt * 16 + r16;
    const int dd = d < dc ? d : dc - 1;
    if (!BF16) {
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int n = 4 * s + q;
        const float val = a.Y[(int64_t)(mbase + dd) * NSH + (n < NSH ? n : NSH - 1)];
        yf[nt][s] = (d < dc && n < NSH) ? val : 0.0f;
      }
    } else {
#pragma unroll
      for (int s = 0; s < KB; ++s) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = 16 * s + 4 * q + j;
          const float val = a.Y[(int64_t)(mbase + dd) * NSH + (n < NSH ? n : NSH - 1)];
          yb[nt][s][j] = bf16_bits((d < dc && n < NSH) ? val : 0.0f);
        }
      }
    }
  }
  float ywf[NT2][DC4];  // reduction B fragments: Yw[d = 4 s + l/16][n = 16 t + l%16]
#pragma unroll
  for (int t = 0; t < NT2; ++t) {
    const int n = t * 16 + r16;
#pragma unroll
    for (int s = 0; s < DC4; ++s) {
      const int d = 4 * s + q;
      const float val = a.Yw[(int64_t)(mbase + (d < dc ? d : dc - 1)) * NSH + (n < NSH ? n : NSH - 1)];
      ywf[t][s] = (d < dc && n < NSH) ? val : 0.0f;
    }
  }
  float cxs[NT], cys[NT], czs[NT];  // direction cosines of the D-layout direction d = 16 nt + l%16
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int d = nt * 16 + r16;
    const int m = mbase + (d < dc ? d : dc - 1);
    cxs[nt] = a.cdir[m * 3 + 0];
    cys[nt] = a.cdir[m * 3 + 1];
    czs[nt] = a.cdir[m * 3 + 2];
  }

  // ---- cell planes --------------------------------------------------------------------------------------------
  // The T fragments and mu_t of plane p+1 and the old moments of plane p are in flight while plane p is computed.
  struct Staged {
    float t[3][KS];
    s4 tb[3][KB];
    float mu[3];
    float sm[3][SRC == 2 ? NT : 1];  // cell-mean source of (cell 4 mt + l/16, direction 16 nt + l%16)
  };
  auto load_plane = [&](const int pl) __attribute__((always_inline)) -> Staged {
    Staged st;
    const int pb = ptab.offset[pl];
    const int P = ptab.offset[pl + 1] - pb;
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) {
      if (mt * 4 >= P) break;
      const int sa = 4 * mt + acell;
      const int ca = sa < P ? slot_cell[pb + (sa < P ? sa : P - 1)] : -1;
      if (!BF16) {
        const unsigned ta = ca >= 0 ? (unsigned)ca * CB + 4u * acomp : kDrop;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          const int n = 4 * s + q;
          st.t[mt][s] = buf_load(T_r, n < NSH ? ta + 16u * n : kDrop);
        }
      } else {  // row (cell, component): KP contiguous bfloat16 moments, four per lane and k-step
        const unsigned tb = ca >= 0 ? ((unsigned)ca * 4u + (unsigned)acomp) * (2u * KP) + 8u * q : kDrop;
#pragma unroll
        for (int s = 0; s < KB; ++s) st.tb[mt][s] = buf_load_s4(T_r, ca >= 0 ? tb + 32u * s : kDrop);
      }
      const int sd = 4 * mt + q;
      const int cd = sd < P ? slot_cell[pb + (sd < P ? sd : P - 1)] : -1;
      st.mu[mt] = buf_load(mu_r, cd >= 0 ? 4u * cd : kDrop);
      if (SRC == 2) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const int d = nt * 16 + r16;
          const unsigned e = (unsigned)cd * a.M + (unsigned)(mbase + d);
          st.sm[mt][nt] = buf_load(src_r, (cd >= 0 && d < dc) ? e * 4u : kDrop);
        }
      }
    }
    return st;
  };

  auto plane = [&](const int p, const Staged cur) __attribute__((always_inline)) -> Staged {
    const int pb = ptab.offset[p];
    const int P = ptab.offset[p + 1] - pb;
    Staged next = cur;
    if (p + 1 < NPL) next = load_plane(p + 1);
    f4 pold[3][NT2];  // moments this plane adds to (zero on first visits: those loads are dropped)
    int cells[3];
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) {
      if (mt * 4 >= P) break;
      const int sd = 4 * mt + q;
      cells[mt] = sd < P ? slot_cell[pb + (sd < P ? sd : P - 1)] : -1;
#pragma unroll
      for (int t = 0; t < NT2; ++t) {
        const int n = t * 16 + r16;
        const bool use = !first_visit && cells[mt] >= 0 && n < NSH;
        pold[mt][t] = buf_load4(phi_r, use ? (unsigned)cells[mt] * CB + 16u * n : kDrop);
      }
    }
    f4 srcv[3][NT];  // explicit source of (cell 4 mt + l/16, direction 16 nt + l%16)
    if (SRC == 2) {
#pragma unroll
      for (int mt = 0; mt < 3; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) srcv[mt][nt] = f4{cur.sm[mt][nt], 0.f, 0.f, 0.f};
    } else if (SRC == 1) {
#pragma unroll
      for (int mt = 0; mt < 3; ++mt) {
        if (mt * 4 >= P) break;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const int d = nt * 16 + r16;
          const bool use = cells[mt] >= 0 && d < dc;
          const unsigned e = (unsigned)cells[mt] * a.M + (unsigned)(mbase + d);
          srcv[mt][nt] = buf_load4(src_r, use ? e * 16u : kDrop);
        }
      }
    }
#pragma unroll
    for (int mt = 0; mt < 3; ++mt) {
      if (mt * 4 >= P) break;
      // (1) source expansion
      f4 sacc[NT];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        sacc[nt] = f4{0.f, 0.f, 0.f, 0.f};
        if (!BF16) {
#pragma unroll
          for (int s = 0; s < KS; ++s) sacc[nt] = mfma4(cur.t[mt][s], yf[nt][s], sacc[nt]);
        } else {
#pragma unroll
          for (int s = 0; s < KB; ++s) sacc[nt] = mfma_bf16(cur.tb[mt][s], yb[nt][s], sacc[nt]);
          sacc[nt] *= tinv;
        }
      }
      // (2) closed-form LD solve of (cell 4 mt + l/16, direction 16 nt + l%16)
      const int sd = 4 * mt + q;
      const bool cvalid = cells[mt] >= 0;
      const int packed = slot_pos[pb + (sd < P ? sd : P - 1)];
      const int lx = packed & 0xff, ly = (packed >> 8) & 0xff, lz = (packed >> 16) & 0xff;
      const float mu = cur.mu[mt];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const int d = nt * 16 + r16;
        const bool valid = cvalid && d < dc;
        // component c of a face entry sits at +c*dc; component 0 is the face average psi_0 + s_a psi_a, 1 and 2 the
        // transverse moments
        float* fx = valid ? face + 0 * face_stride + (ly * B + lz) * 3 * dc + d : scratch_row + lane;
        float* fy = valid ? face + 1 * face_stride + (lx * B + lz) * 3 * dc + d : scratch_row + lane;
        float* fz = valid ? face + 2 * face_stride + (lx * B + ly) * 3 * dc + d : scratch_row + lane;
        const float pbx = fx[0], xy = fx[dc], xz = fx[2 * dc];
        const float pby = fy[0], yx = fy[dc], yz = fy[2 * dc];
        const float pbz = fz[0], zx = fz[dc], zy = fz[2 * dc];
        const float cx = cxs[nt], cy = cys[nt], cz = czs[nt];
        const f4 sv = HAS_SRC ? sacc[nt] + srcv[mt][nt] : sacc[nt];
        const float r0 = sv.x + (cx * pbx + cy * pby + cz * pbz);
        const float rx = sv.y - 3.0f * fsx * cx * pbx + (cy * yx + cz * zx);
        const float ry = sv.z - 3.0f * fsy * cy * pby + (cx * xy + cz * zy);
        const float rz = sv.w - 3.0f * fsz * cz * pbz + (cx * xz + cy * yz);
        const float csum = cx + cy + cz;
        // hardware reciprocals (1 ulp) in place of IEEE divisions: four per solve instead of seven divisions
        const float idx = __builtin_amdgcn_rcpf(3.0f * cx + (csum - cx) + mu);
        const float idy = __builtin_amdgcn_rcpf(3.0f * cy + (csum - cy) + mu);
        const float idz = __builtin_amdgcn_rcpf(3.0f * cz + (csum - cz) + mu);
        const float qx = cx * idx, qy = cy * idy, qz = cz * idz;
        const float iden = __builtin_amdgcn_rcpf(csum + mu + 3.0f * (cx * qx + cy * qy + cz * qz));
        const float p0 = (r0 - (fsx * qx * rx + fsy * qy * ry + fsz * qz * rz)) * iden;
        const float px = (rx + 3.0f * fsx * cx * p0) * idx;
        const float py = (ry + 3.0f * fsy * cy * p0) * idy;
        const float pz = (rz + 3.0f * fsz * cz * p0) * idz;
        fx[0] = p0 + fsx * px; fx[dc] = py; fx[2 * dc] = pz;
        fy[0] = p0 + fsy * py; fy[dc] = px; fy[2 * dc] = pz;
        fz[0] = p0 + fsz * pz; fz[dc] = px; fz[2 * dc] = py;
        // padding rows and directions contribute exact zeros to the reduction
        const f4 psi = valid ? f4{p0, px, py, pz} : f4{0.f, 0.f, 0.f, 0.f};
        *reinterpret_cast<f4*>(psl + q * PS + d * 4) = psi;
      }
      // (3) moment reduction over the chunk's directions, added into phi
      f4 facc[NT2];
#pragma unroll
      for (int t = 0; t < NT2; ++t) facc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < DC4; ++s) {
        const float av = psl[acell * PS + (4 * s + q) * 4 + acomp];
#pragma unroll
        for (int t = 0; t < NT2; ++t) facc[t] = mfma4(av, ywf[t][s], facc[t]);
      }
#pragma unroll
      for (int t = 0; t < NT2; ++t) {
        const int n = t * 16 + r16;
        buf_store4(phi_r, (cvalid && n < NSH) ? (unsigned)cells[mt] * CB + 16u * n : kDrop, facc[t] + pold[mt][t]);
      }
    }
    return next;
  };

  Staged cur = load_plane(0);
#pragma unroll
  for (int p = 0; p < NPL; ++p) cur = plane(p, cur);

  // ---- outgoing faces for the downstream blocks ---------------------------------------------------------------
  {
    const int db[3][3] = {{bi, bj + sx, bk}, {bi + sy, bj, bk}, {bi, bj, bk + sz}};
    for (int ax = 0; ax < 3; ++ax) {
      const int di = db[ax][0], dj = db[ax][1], dk = db[ax][2];
      if (!(di >= 0 && di < a.nbr && dj >= 0 && dj < a.nbc && dk >= 0 && dk < a.nbs)) continue;
      f4* dst = reinterpret_cast<f4*>(gface_v + ((int64_t)blk * 3 + ax) * face_stride);
      const f4* srcf = reinterpret_cast<const f4*>(face + ax * face_stride);
      for (int i = lane; i < n4; i += 64) dst[i] = srcf[i];
    }
  }
}

size_t mfma_lds_bytes(int dc) { return sizeof(float) * (size_t)(144 * dc + 4 * kMfmaPsiStride + 128 + 2 * dc + 64); }

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
  if (A.T_bf16 && (!A.T_inv_scale || A.T)) return fail(CITRUS_ERR_ARG, "T_bf16 needs T_inv_scale and no float32 T");
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
  la.tinv = A.T_inv_scale;
  la.t16view = (int64_t)A.G * N * 4 * ((A.nsh + 15) / 16 * 16);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  la.ntask = plan->launches.empty() ? 0 : plan->launches[0].second;
  const bool mfma = use_mfma(la, A.nsh);
  if (la.T16 && !mfma) return fail(CITRUS_ERR_UNSUPPORTED, "the bfloat16 contraction runs on the matrix-core kernel only");
  for (size_t l = 0; l < plan->launches.size(); ++l) {
    la.tasks = plan->d_tasks + plan->launches[l].first;
    la.ntask = plan->launches[l].second;
    la.first = (!A.accumulate && plan->launch_pass[l] == 0) ? 1 : 0;
    const int64_t nwg = (int64_t)la.ntask * A.G * A.V;
    if (nwg > 0x7fffffff) return fail(CITRUS_ERR_ARG, "too many workgroups in one launch");
    hipError_t e = mfma ? launch_mfma(A.nsh, la, (int)nwg, s) : launch_nsh<4>(A.nsh, la, (int)nwg, s);
    if (e == hipErrorInvalidValue) return fail(CITRUS_ERR_UNSUPPORTED, "nsh must be (L+1)^2 with L <= 7");
    if (e == hipErrorInvalidConfiguration) return fail(CITRUS_ERR_UNSUPPORTED, "chunk needs more than 64 KiB of LDS");
    if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("sweep launch: ") + hipGetErrorString(e));
  }
  return CITRUS_OK;
}
