// Device code shared by the LD sweep kernels (ld_sweep.hip: one launch per block diagonal; ld_sweep_persistent.hip:
// one persistent launch): vector types, buffer-resource helpers, the cell-plane table of a 4^3 block and the
// matrix-core task body that sweeps one (view, octant, group, block) task.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "block_planes.h"
#include "common.h"

namespace citrus {
namespace {

// Native 4-wide float vector for register-resident prefetch buffers (HIP's float4 wrapper is copied through memory,
// which keeps arrays of it out of registers).
typedef float f4 __attribute__((ext_vector_type(4)));

// Workgroup barrier that orders LDS accesses only: outstanding global loads stay in flight across it (a full
// __syncthreads also waits for them). Every cross-thread exchange in the sweep kernel goes through LDS.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }


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

// Device-scope (SC1) buffer accesses: the cache policy of agent-scope atomics. Loads miss the CU's L1 and stores
// write through, so data one wave stores before raising a flag is what another wave, on any CU or XCD, loads after
// seeing it -- with no L2 writeback or invalidate.
constexpr int kScopeDevice = 16;  // CPol::SC1

template <bool COH>
__device__ __forceinline__ f4 buf_load4_c(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, COH ? kScopeDevice : 0));
}

template <bool COH>
__device__ __forceinline__ void buf_store4_c(__amdgpu_buffer_rsrc_t r, unsigned off, f4 v) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, (int)off, 0, COH ? kScopeDevice : 0);
}

// One (view, octant, group, block) task of the matrix-core sweep, resolved to its view's buffers.
struct MfmaTask {
  int R, C, S, N;         // grid of the task's view (cells along rows y, cols x, slices z)
  int nbr, nbc, nbs;      // blocks along rows, cols and slices
  int o, bi, bj, bk;      // octant and block
  int first_visit;        // the block's first visit of the sweep: phi is stored instead of added to
  float* gface;           // inter-block faces of the (view, slot, group): [blocks][3 axes][16 positions][3][dc]
  const float* mu;        // total attenuation of the group [N]
  const void* T;          // moments of the (view, group): float32 [N][NSH][4], or (mixed operand) the bfloat16
                          // degrees l >= 2 [N][4][KH]
  const float* Tlo;       // mixed operand: the float32 degrees l <= 1, [N][4][4] (component-major), else unused
  float* phi;             // fluence moments of the (view, group) [N][NSH][4]
  const float* src;       // explicit source of the (view, group): [N][M][4] (SRC 1), [N][M] (SRC 2), the block-order
                          // cell means [blocks][8][64][dp] (SRC 3), or NULL
  const float* cdir;      // [8 D8][3] c_a = |Omega_a| / Delta_a on the view's grid
  const float* fc_scale;  // SRC 3: [blocks][64] scale of the first-collision sampling correction of the (view, group)
  const float* fc_beta;   // SRC 3: [blocks][64] its beta
  const float* fc_u;      // SRC 3: [blocks][64][4] Omega_UC of each cell of the view
};

// Direction tables and chunk shared by every task of a sweep.
struct DirTables {
  int D8, d0, dc, M;      // directions per octant, chunk [d0, d0 + dc), M = 8 D8
  const float* Y;         // [M][NSH] real spherical harmonics of the reordered directions
  const float* Yw;        // [M][NSH] the same times the quadrature weights
  const float* fc_omega = nullptr;  // SRC 3: [M][4] the reordered directions (x, y, z, 0)
};


// SRC: 0 no explicit source; 1 group-major LD moments (layout 1, read per plane); 2 group-major cell means (layout 2,
// prefetched one plane ahead with T, the slopes zero); 3 the cell means before the first-collision sampling
// correction, stored in block order (per block and octant the 64 cells' dp = D8 rounded up to 4 directions, one
// contiguous run per task with the cells in the task's plane order; the correction factors and Omega_UC per block of
// 64 cells, row-major), the correction applied here as
// they are read: mean scale (1 + beta eta) with eta = u[cell] . omega[m] (direction_cosine), the operations of the
// correction kernels. COH: the faces and moments other tasks of the same launch
// produce or consume are accessed at device scope (the persistent sweep); the launch-sequence sweep exchanges them
// across kernel boundaries only.
template <int NSH, int DC4, int SRC, bool BF16, bool COH = false>
__device__ __forceinline__ void mfma_task(const MfmaTask& a, const DirTables& dt, float* lds, const int lane) {
  constexpr bool HAS_SRC = SRC != 0;
  constexpr int B = 4;
  constexpr PlaneTable<B> ptab{};
  constexpr int NPL = PlaneTable<B>::kPlanes;
  constexpr int NT = (DC4 + 3) / 4;     // direction tiles of the expansion
  // float32 k-steps (4 moments each): every moment, or with the mixed operand the degrees l <= 1 (moments 0..3)
  constexpr int KS = BF16 ? 1 : (NSH + 3) / 4;
  constexpr int KB = BF16 ? (NSH - 4 + 15) / 16 : 1;  // bfloat16 k-steps (16 moments each) of the degrees l >= 2
  constexpr int KH = 16 * KB;                         // moments per row of the bfloat16 part (n = 4 .. 4 + KH)
  constexpr int NT2 = (NSH + 15) / 16;  // moment tiles of the reduction
  constexpr int PS = kMfmaPsiStride;
  constexpr unsigned CB = NSH * 16;     // bytes of one cell's moments
  const PlaneTable<B>& tab = planes<B>();

  const int dc = dt.dc;
  const int face_stride = 48 * dc;                           // floats per axis: [16 positions][3 components][dc]
  float* face = lds;                                         // [3 axes] of the above
  float* psl = face + 3 * face_stride;                       // [4 cells][PS]
  int* slot_cell = reinterpret_cast<int*>(psl + 4 * PS);     // [64] flat cell of each slot, or -1
  int* slot_pos = slot_cell + 64;                            // [64] local (x, y, z) of each slot, packed
  float* scratch_row = reinterpret_cast<float*>(slot_pos + 64);  // [2 dc + 64] face updates of padding lanes

  const int q = lane >> 4, r16 = lane & 15;
  const int acell = r16 >> 2, acomp = r16 & 3;  // A-fragment row: (cell of the M-tile, LD component)
  const int o = a.o;
  const bool first_visit = a.first_visit != 0;
  const int bi = a.bi, bj = a.bj, bk = a.bk;
  const int sx = (o & 4) ? -1 : 1, sy = (o & 2) ? -1 : 1, sz = (o & 1) ? -1 : 1;
  const float fsx = (float)sx, fsy = (float)sy, fsz = (float)sz;
  const int blk = (bi * a.nbc + bj) * a.nbs + bk;
  const int x0 = bj * B, y0 = bi * B, z0 = bk * B;

  float* gface_v = a.gface;
  // faces of every block of the (view, slot, group): [blocks][3 axes][face_stride]
  const __amdgpu_buffer_rsrc_t face_r = __builtin_amdgcn_make_buffer_rsrc(
      gface_v, 0, a.nbr * a.nbc * a.nbs * 3 * face_stride * 4, kBufferWord3);
  const __amdgpu_buffer_rsrc_t mu_r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(a.mu), 0, a.N * 4, kBufferWord3);
  const __amdgpu_buffer_rsrc_t T_r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(a.T), 0, BF16 ? a.N * 8 * KH : a.N * (int)CB, kBufferWord3);
  const __amdgpu_buffer_rsrc_t Tlo_r = __builtin_amdgcn_make_buffer_rsrc(
      BF16 ? const_cast<float*>(a.Tlo) : const_cast<void*>(a.T), 0, BF16 ? a.N * 64 : 0, kBufferWord3);
  const __amdgpu_buffer_rsrc_t phi_r = __builtin_amdgcn_make_buffer_rsrc(a.phi, 0, a.N * (int)CB, kBufferWord3);
  // explicit source, group-major [N][M][4] (layout 1: the 16 directions of one cell in a D tile are 256 contiguous
  // bytes) or cell means [N][M] (layout 2, zero slopes: 64 contiguous bytes)
  constexpr int src_words = SRC == 2 ? 1 : 4;
  const int nblocks = a.nbr * a.nbc * a.nbs;
  const int dp = (dt.D8 + 3) & ~3;  // SRC 3: stored directions per octant
  const unsigned run64 = ((unsigned)blk * 8u + (unsigned)o) * 64u;  // SRC 3: the task's run, in cells
  const __amdgpu_buffer_rsrc_t src_r = __builtin_amdgcn_make_buffer_rsrc(
      HAS_SRC ? const_cast<float*>(a.src) : a.phi, 0,
      SRC == 3 ? nblocks * 8 * 64 * dp * 4 : (HAS_SRC ? a.N * dt.M * 4 * src_words : 0), kBufferWord3);
  // SRC 3: the per-cell factors of the first-collision sampling correction, in block order
  const __amdgpu_buffer_rsrc_t fs_r = __builtin_amdgcn_make_buffer_rsrc(
      SRC == 3 ? const_cast<float*>(a.fc_scale) : a.phi, 0, SRC == 3 ? nblocks * 64 * 4 : 0, kBufferWord3);
  const __amdgpu_buffer_rsrc_t fb_r = __builtin_amdgcn_make_buffer_rsrc(
      SRC == 3 ? const_cast<float*>(a.fc_beta) : a.phi, 0, SRC == 3 ? nblocks * 64 * 4 : 0, kBufferWord3);
  const __amdgpu_buffer_rsrc_t fu_r = __builtin_amdgcn_make_buffer_rsrc(
      SRC == 3 ? const_cast<float*>(a.fc_u) : a.phi, 0, SRC == 3 ? nblocks * 64 * 16 : 0, kBufferWord3);

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
      const unsigned fbase = (unsigned)((ublk * 3 + ax) * face_stride) * 4u;
#pragma unroll
      for (int r = 0; r < FPL; ++r) {
        const int i = lane + 64 * r;
        if (COH)
          fin[ax][r] = buf_load4_c<true>(face_r, fbase + 16u * (unsigned)(i < n4 ? i : n4 - 1));
        else
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
  const int mbase = o * dt.D8 + dt.d0;
  float yf[NT][KS];  // float32 expansion B fragments: Y[d = 16 nt + l%16][n = 4 s + l/16]
  s4 yb[NT][KB];     // bfloat16 expansion B fragments: Y[d = 16 nt + l%16][n = 4 + 16 s + 4 (l/16) + 0..3]
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int d = nt * 16 + r16;
    const int dd = d < dc ? d : dc - 1;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int n = 4 * s + q;
      const float val = dt.Y[(int64_t)(mbase + dd) * NSH + (n < NSH ? n : NSH - 1)];
      yf[nt][s] = (d < dc && n < NSH) ? val : 0.0f;
    }
    if (BF16) {
#pragma unroll
      for (int s = 0; s < KB; ++s) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = 4 + 16 * s + 4 * q + j;
          const float val = dt.Y[(int64_t)(mbase + dd) * NSH + (n < NSH ? n : NSH - 1)];
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
      const float val = dt.Yw[(int64_t)(mbase + (d < dc ? d : dc - 1)) * NSH + (n < NSH ? n : NSH - 1)];
      ywf[t][s] = (d < dc && n < NSH) ? val : 0.0f;
    }
  }
  float cxs[NT], cys[NT], czs[NT];  // direction cosines of the D-layout direction d = 16 nt + l%16
  f4 fomg[SRC == 3 ? NT : 1];       // SRC 3: that direction's unit vector
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int d = nt * 16 + r16;
    const int m = mbase + (d < dc ? d : dc - 1);
    cxs[nt] = a.cdir[m * 3 + 0];
    cys[nt] = a.cdir[m * 3 + 1];
    czs[nt] = a.cdir[m * 3 + 2];
    if (SRC == 3) fomg[nt] = reinterpret_cast<const f4*>(dt.fc_omega)[m];
  }

  // ---- cell planes --------------------------------------------------------------------------------------------
  // The T fragments and mu_t of plane p+1 and the old moments of plane p are in flight while plane p is computed.
  constexpr bool MEANS = SRC == 2 || SRC == 3;  // the explicit source is a cell mean per (cell, direction)
  struct Staged {
    float t[3][KS];
    s4 tb[3][KB];
    float mu[3];
    float sm[3][MEANS ? NT : 1];  // cell-mean source of (cell 4 mt + l/16, direction 16 nt + l%16)
  };
  // SRC 3: the cell of (4 mt + l/16) in block order, (row, column, slice) row-major within the block
  auto block_cell = [&](const int pb, const int P, const int sd) __attribute__((always_inline)) -> unsigned {
    const int pk = slot_pos[pb + (sd < P ? sd : P - 1)];
    return (unsigned)blk * 64u + (unsigned)(((((pk >> 8) & 0xff) * 4) + (pk & 0xff)) * 4 + ((pk >> 16) & 0xff));
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
      } else {
        // row (cell, component): its 4 float32 moments of degree <= 1 (one per lane) and its KH bfloat16 moments of
        // degree >= 2 (four per lane and k-step)
        const unsigned row = (unsigned)ca * 4u + (unsigned)acomp;
        st.t[mt][0] = buf_load(Tlo_r, ca >= 0 ? row * 16u + 4u * q : kDrop);
        const unsigned tb = ca >= 0 ? row * (2u * KH) + 8u * q : kDrop;
#pragma unroll
        for (int s = 0; s < KB; ++s) st.tb[mt][s] = buf_load_s4(T_r, ca >= 0 ? tb + 32u * s : kDrop);
      }
      const int sd = 4 * mt + q;
      const int cd = sd < P ? slot_cell[pb + (sd < P ? sd : P - 1)] : -1;
      st.mu[mt] = buf_load(mu_r, cd >= 0 ? 4u * cd : kDrop);
      // SRC 3: the task's run holds the cells in this task order (slot pb + sd), so a plane's cells are adjacent
      unsigned slot = (unsigned)(pb + (sd < P ? sd : P - 1));
      if (SRC == 3) asm volatile("" : "+v"(slot));  // formed here, not hoisted for every plane up front
      if (MEANS) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const int d = nt * 16 + r16;
          const unsigned e = SRC == 3 ? (run64 + slot) * (unsigned)dp + (unsigned)(dt.d0 + d)
                                      : (unsigned)cd * dt.M + (unsigned)(mbase + d);
          st.sm[mt][nt] = buf_load(src_r, (cd >= 0 && d < dc) ? e * 4u : kDrop);
        }
      }
    }
    return st;
  };

  auto plane = [&](const int p, const Staged cur) __attribute__((always_inline)) -> Staged {
    const int pb = ptab.offset[p];
    const int P = ptab.offset[p + 1] - pb;
    // Loads retire in issue order, so a wait for one load also waits for every load issued before it. Below 8
    // direction tiles (at 8 the kernel then spills) the SRC 3 correction factors, used right after the expansion, are
    // requested before the next plane's prefetch, which the wait for them then leaves in flight.
    constexpr bool kFactorsFirst = SRC == 3 && DC4 < 8;
    // SRC 3: this plane's correction factors (scale, beta, Omega_UC) of the cell of (4 mt + l/16); held for one plane
    // only: prefetching them a plane ahead with the means doubles their registers
    float fs[3], fb[3];
    f4 fu[3];
    int cells[3];  // the flat cell of (4 mt + l/16), or -1
    auto load_factors = [&]() __attribute__((always_inline)) {
#pragma unroll
      for (int mt = 0; mt < 3; ++mt) {
        if (mt * 4 >= P) break;
        const unsigned bcell = block_cell(pb, P, 4 * mt + q);
        const int sd = 4 * mt + q;
        const bool on = kFactorsFirst ? sd < P && slot_cell[pb + (sd < P ? sd : P - 1)] >= 0 : cells[mt] >= 0;
        fs[mt] = buf_load(fs_r, on ? 4u * bcell : kDrop);
        fb[mt] = buf_load(fb_r, on ? 4u * bcell : kDrop);
        fu[mt] = buf_load4(fu_r, on ? 16u * bcell : kDrop);
      }
    };
    if (kFactorsFirst) load_factors();
    // the moments this plane adds to (zero on first visits: those loads are dropped)
    f4 pold[3][NT2];
    auto load_pold = [&]() __attribute__((always_inline)) {
#pragma unroll
      for (int mt = 0; mt < 3; ++mt) {
        if (mt * 4 >= P) break;
        const int sd = 4 * mt + q;
        cells[mt] = sd < P ? slot_cell[pb + (sd < P ? sd : P - 1)] : -1;
#pragma unroll
        for (int t = 0; t < NT2; ++t) {
          const int n = t * 16 + r16;
          const bool use = !first_visit && cells[mt] >= 0 && n < NSH;
          pold[mt][t] = buf_load4_c<COH>(phi_r, use ? (unsigned)cells[mt] * CB + 16u * n : kDrop);
        }
      }
    };
    Staged next = cur;
    if (p + 1 < NPL) next = load_plane(p + 1);
    load_pold();
    if (SRC == 3 && !kFactorsFirst) load_factors();
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
          const unsigned e = (unsigned)cells[mt] * dt.M + (unsigned)(mbase + d);
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
#pragma unroll
        for (int s = 0; s < KS; ++s) sacc[nt] = mfma4(cur.t[mt][s], yf[nt][s], sacc[nt]);
        if (BF16) {
#pragma unroll
          for (int s = 0; s < KB; ++s) sacc[nt] = mfma_bf16(cur.tb[mt][s], yb[nt][s], sacc[nt]);
        }
      }
      if (SRC == 3) {
        // the corrected cell means; the products stay rounded as the correction kernels round them, not fused into
        // the source's later addition
#pragma clang fp contract(off)
        const f4 u = fu[mt];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const f4 om = fomg[nt];
          const float eta = direction_cosine(u.x, u.y, u.z, om.x, om.y, om.z);
          srcv[mt][nt] = f4{cur.sm[mt][nt] * fs[mt] * fmaf(fb[mt], eta, 1.0f), 0.f, 0.f, 0.f};
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
        buf_store4_c<COH>(phi_r, (cvalid && n < NSH) ? (unsigned)cells[mt] * CB + 16u * n : kDrop,
                          facc[t] + pold[mt][t]);
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
      if (COH) {
        const unsigned obase = (unsigned)((blk * 3 + ax) * face_stride) * 4u;
        for (int i = lane; i < n4; i += 64) buf_store4_c<true>(face_r, obase + 16u * (unsigned)i, srcf[i]);
      } else {
        for (int i = lane; i < n4; i += 64) dst[i] = srcf[i];
      }
    }
  }
}


// LDS bytes of one matrix-core task: faces [3][48 dc], the psi tile, the slot tables and the padding lanes' row
inline size_t mfma_lds_bytes(int dc) {
  return sizeof(float) * (size_t)(144 * dc + 4 * kMfmaPsiStride + 128 + 2 * dc + 64);
}

}  // namespace
}  // namespace citrus
