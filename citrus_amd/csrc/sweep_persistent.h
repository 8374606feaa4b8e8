// Persistent LD sweep for gfx950 (shared declarations and the kernel template): every (view, group, octant, block)
// task of a sweep in ONE launch, for views whose transport grids may differ in shape and cell size (cropped helical
// views, coarse and fine multigrid stages).
//
// Schedule. The tasks and their order are those of the launch-sequence sweep (ld_sweep.hip): octant pairs (o, 7-o)
// in passes 0..3, each pass walking the block diagonals of its two octants. Instead of a launch per diagonal, a
// fixed grid of resident single-wave workgroups takes tasks from per-XCD queues with an atomic ticket and waits, per
// task, only for the tasks it actually depends on:
//   - the upstream block along each axis in the same octant (its outgoing faces are this task's incoming faces);
//   - the previous visit of the same block (the order in which the eight octants add into a cell's moments is that
//     of the launch sequence, so the two schedules give bit-identical moments);
//   - the readers of the face slot this task overwrites (the downstream blocks of the octant that last used it).
// Every dependency has a smaller ticket in the same queue, so the task with the smallest unfinished ticket can always
// run: the waves that hold earlier tickets are resident, and no wave ever waits on a later ticket. Every wait is
// bounded: a wave that spins past the limit raises the abort status, and every wave leaves at its next check.
//
// Coherence. The faces and moments one task hands to another are loaded and stored at device scope (SC1, the cache
// policy of agent-scope atomics: loads miss the CU's L1, stores write through), the producer waits for its stores
// (s_waitcnt vmcnt(0)) before storing its done flag, and the consumer issues its loads only after seeing the flag.
// No L2 writeback or invalidate is needed on any task. The read-only inputs (moments T, mu_t, sources, direction
// tables) keep the ordinary cached path.
//
// XCD affinity. Each (view, group) domain is assigned to one of the 8 XCDs, and a wave takes tasks only from the
// queue of the XCD it runs on (the hardware register XCC_ID): a domain's mu_t and its flags stay in one XCD's L2,
// and the per-queue tickets are touched by one XCD only.

#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "citrus_hip.h"
#include "sweep_device.h"

namespace citrus {

inline constexpr int kQueues = 8;                 // XCDs of an MI355X
inline constexpr int kMaxViews = CITRUS_PERSISTENT_MAX_VIEWS;
inline constexpr int kCtrlInts = 64;              // [0, 8) tickets, [8] status
inline constexpr int kStatusWord = 8;
inline constexpr int kSpinLimit = 1 << 21;        // sleeps of one wait before the sweep is aborted (about a second)

struct PersistentView {
  int R, C, S, N;
  int nbr, nbc, nbs, K;
  int64_t face_off;   // floats into the face workspace: [2 slots][G][blocks][3][48 dc]
  int64_t flag_off;   // ints into the flag array: [G][8 octants][blocks]
  const float* mu;    // [G][N]
  const void* T;      // [G][N][NSH][4] float32, or the mixed operand's bfloat16 degrees l >= 2 [G][N][4][KH]
  const float* Tlo;   // the mixed operand's float32 degrees l <= 1 [G][N][4][4]
  const float* src;   // [G][N][M] or [G][N][M][4], source layout 3 [G][blocks * 8 * 64 * dp + pad] (block order), or NULL
  float* phi;         // [G][N][NSH][4]
  const float* cdir;  // [8 D8][3]
  const float* fc_scale;  // source layout 3: [G][blocks * 64 + pad] scale and beta of the first-collision correction
  const float* fc_beta;
  const float* fc_u;      // source layout 3: [blocks][64][4] Omega_UC of each cell
};

struct PersistentArgs {
  int G, D8, d0, dc, M, nviews, accumulate, spin_limit;
  int waves_per_cu;  // host side only: resident workgroups per CU of the launch (0: the occupancy)
  const float* Y;
  const float* Yw;
  const int4* tasks;  // (view | group << 8 | octant << 16, block row, block col, block slice), queue by queue
  const int* qoff;    // queue q holds tasks [qoff[q], qoff[q + 1])
  float* faces;
  int* flags;
  int* ctrl;
  int* sink;          // caller's sticky status word (OR of every abort code), or NULL
  const float* fc_omega;  // source layout 3: [M][4] the reordered directions
  PersistentView views[kMaxViews];
};


// Launch the persistent kernel of one moment count (P3: 16, P5: 36) on a grid of resident waves. layout: -1 no explicit
// source, 1 or 2 the group-major source layouts, 3 layout 2 with the first-collision correction applied in the sweep. Defined in ld_sweep_persistent_p3.hip / _p5.hip.
hipError_t persistent_launch_p3(const PersistentArgs& a, bool bf16, int layout, size_t lds, hipStream_t s);
hipError_t persistent_launch_p5(const PersistentArgs& a, bool bf16, int layout, size_t lds, hipStream_t s);

namespace {

// Block plane (block diagonal) of block (bi, bj, bk) for octant o: coordinates mirrored into the octant's sweep order.
__host__ __device__ __forceinline__ int block_plane_of(int o, int bi, int bj, int bk, int nbr, int nbc, int nbs) {
  const int a = (o & 4) ? nbc - 1 - bj : bj;
  const int b = (o & 2) ? nbr - 1 - bi : bi;
  const int c = (o & 1) ? nbs - 1 - bk : bk;
  return a + b + c;
}

// Whether octant o's visit of a block on its block plane d is the first of its octant pair (o, 7 - o): the octant
// whose plane comes first, the pass's first octant o = pass on the middle plane.
__host__ __device__ __forceinline__ bool first_in_pair(int o, int d, int K) {
  return o < 4 ? (2 * d <= K) : (2 * d < K);
}

// Calls f(octant, bi, bj, bk) for every task that task (o, bi, bj, bk) of one (view, group) must wait for:
//   - the upstream block along x (cols), y (rows) and z (slices) in octant o (read-after-write of the faces);
//   - the previous visit of the block (the other octant of the pair, or the pair before's second visitor), which
//     fixes the order of the additions into the block's moments;
//   - the downstream blocks of the octant that last wrote the face slot this task writes (the previous pass's octant
//     of the slot), which must have read it (write-after-read of the faces).
template <class F>
__host__ __device__ __forceinline__ void for_each_dependency(int nbr, int nbc, int nbs, int K, int o, int bi, int bj,
                                                             int bk, F&& f) {
  auto inside = [&](int i, int j, int k) { return i >= 0 && i < nbr && j >= 0 && j < nbc && k >= 0 && k < nbs; };
  const int pass = o < 4 ? o : 7 - o;
  const int sx = (o & 4) ? -1 : 1, sy = (o & 2) ? -1 : 1, sz = (o & 1) ? -1 : 1;
  if (inside(bi, bj - sx, bk)) f(o, bi, bj - sx, bk);
  if (inside(bi - sy, bj, bk)) f(o, bi - sy, bj, bk);
  if (inside(bi, bj, bk - sz)) f(o, bi, bj, bk - sz);
  if (!first_in_pair(o, block_plane_of(o, bi, bj, bk, nbr, nbc, nbs), K)) {
    f(7 - o, bi, bj, bk);
  } else if (pass > 0) {
    const int pa = pass - 1;
    f(first_in_pair(pa, block_plane_of(pa, bi, bj, bk, nbr, nbc, nbs), K) ? 7 - pa : pa, bi, bj, bk);
  }
  if (pass > 0) {
    const int ow = o < 4 ? pass - 1 : 8 - pass;
    const int wx = (ow & 4) ? -1 : 1, wy = (ow & 2) ? -1 : 1, wz = (ow & 1) ? -1 : 1;
    if (inside(bi, bj + wx, bk)) f(ow, bi, bj + wx, bk);
    if (inside(bi + wy, bj, bk)) f(ow, bi + wy, bj, bk);
    if (inside(bi, bj, bk + wz)) f(ow, bi, bj, bk + wz);
  }
}

__device__ __forceinline__ int load_flag(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Abandon the sweep: the workspace's status word (read by every wave's next check) and the caller's sticky word.
__device__ __forceinline__ void abort_sweep(int* status, int* sink, int code) {
  __hip_atomic_store(status, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (sink) __hip_atomic_fetch_or(sink, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until the flag at `p` is set; false when the sweep is (or gets) aborted. Every lane of the wave runs it on the
// same address and every branch is on a wave-uniform (scalar) value: control flow inside the task loop must never
// depend on the lane, or the compiler may split the wave across the loop's exits.
__device__ __forceinline__ bool wait_flag(const int* p, int* status, int* sink, int limit) {
  int spins = 0;
  while (__builtin_amdgcn_readfirstlane(load_flag(p)) == 0) {
    if (++spins > limit) {
      abort_sweep(status, sink, CITRUS_PERSISTENT_TIMEOUT);
      return false;
    }
    if ((spins & 31) == 0 && __builtin_amdgcn_readfirstlane(load_flag(status)) != 0) return false;
    __builtin_amdgcn_s_sleep(2);
  }
  return true;
}

template <int NSH, int DC4, int SRC, bool BF16>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void ld_sweep_persistent_kernel(
    PersistentArgs a) {
  extern __shared__ float lds[];
  constexpr int KH = 16 * ((NSH - 4 + 15) / 16);
  const int lane = threadIdx.x;
  int* status = a.ctrl + kStatusWord;
  const int xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15;  // HW_REG_XCC_ID, bits [3:0]
  if (xcc >= kQueues) {
    abort_sweep(status, a.sink, CITRUS_PERSISTENT_BAD_XCC);
    return;
  }
  const int qbeg = a.qoff[xcc];
  const int qlen = a.qoff[xcc + 1] - qbeg;
  const DirTables dt{a.D8, a.d0, a.dc, a.M, a.Y, a.Yw, a.fc_omega};
  const int face_stride = 48 * a.dc;
  for (;;) {
    // lane 0 adds 1 and the others 0, so the wave takes one ticket whether or not the adds are combined
    const int ticket = __builtin_amdgcn_readfirstlane(
        __hip_atomic_fetch_add(a.ctrl + xcc, lane == 0 ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (ticket >= qlen) break;
    const int4 e = a.tasks[qbeg + ticket];
    const int v = e.x & 0xff, g = (e.x >> 8) & 0xff, o = (e.x >> 16) & 7;
    const int bi = e.y, bj = e.z, bk = e.w;
    const PersistentView& pv = a.views[v];
    const int nb = pv.nbr * pv.nbc * pv.nbs;
    const int blk = (bi * pv.nbc + bj) * pv.nbs + bk;
    const int pass = o < 4 ? o : 7 - o;
    const int which = o < 4 ? 0 : 1;  // face slot: the pass's first octant (o = pass) or its mirror (7 - pass)
    const bool first = first_in_pair(o, block_plane_of(o, bi, bj, bk, pv.nbr, pv.nbc, pv.nbs), pv.K);
    int* fl = a.flags + pv.flag_off + (int64_t)g * 8 * nb;  // [8][nb] done flags of the (view, group)

    bool ok = true;
    for_each_dependency(pv.nbr, pv.nbc, pv.nbs, pv.K, o, bi, bj, bk, [&](int oo, int i, int j, int k) {
      if (ok) ok = wait_flag(fl + oo * nb + (i * pv.nbc + j) * pv.nbs + k, status, a.sink, a.spin_limit);
    });
    if (!ok) break;

    MfmaTask t;
    t.R = pv.R; t.C = pv.C; t.S = pv.S; t.N = pv.N;
    t.nbr = pv.nbr; t.nbc = pv.nbc; t.nbs = pv.nbs;
    t.o = o; t.bi = bi; t.bj = bj; t.bk = bk;
    t.first_visit = (!a.accumulate && pass == 0 && first) ? 1 : 0;
    t.gface = a.faces + pv.face_off + ((int64_t)which * a.G + g) * nb * 3 * face_stride;
    t.mu = pv.mu + (int64_t)g * pv.N;
    if (BF16) {
      t.T = static_cast<const uint16_t*>(pv.T) + (int64_t)g * pv.N * 4 * KH;
      t.Tlo = pv.Tlo + (int64_t)g * pv.N * 16;
    } else {
      t.T = static_cast<const float*>(pv.T) + (int64_t)g * pv.N * NSH * 4;
      t.Tlo = nullptr;
    }
    t.phi = pv.phi + (int64_t)g * pv.N * NSH * 4;
    // SRC 3: per group [blocks][8][64][dp] cell means and [blocks][64] factors (block order), each group's region
    // followed by CITRUS_FC_GROUP_PAD floats
    const int64_t fc_cells = (int64_t)nb * 64;
    t.src = SRC == 0   ? nullptr
            : SRC == 3 ? pv.src + (int64_t)g * (fc_cells * 8 * ((a.D8 + 3) & ~3) + CITRUS_FC_GROUP_PAD)
                       : pv.src + (int64_t)g * pv.N * a.M * (SRC == 1 ? 4 : 1);
    t.cdir = pv.cdir;
    t.fc_scale = SRC == 3 ? pv.fc_scale + (int64_t)g * (fc_cells + CITRUS_FC_GROUP_PAD) : nullptr;
    t.fc_beta = SRC == 3 ? pv.fc_beta + (int64_t)g * (fc_cells + CITRUS_FC_GROUP_PAD) : nullptr;
    t.fc_u = SRC == 3 ? pv.fc_u : nullptr;
    // the lane index is re-read per task so that nothing derived from it is hoisted out of the task loop and kept
    // live across tasks (the task body needs every register it has)
    int task_lane;
    asm volatile("v_mov_b32 %0, %1" : "=v"(task_lane) : "v"(lane));
    mfma_task<NSH, DC4, SRC, BF16, true>(t, dt, lds, task_lane);

    // every (device-scope) store of the task has completed before its done flag is raised
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __hip_atomic_store(fl + o * nb + blk, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // every lane, one value
  }
}

template <int NSH, int DC4, int SRC, bool BF16>
hipError_t launch_persistent_t(const PersistentArgs& a, size_t lds, hipStream_t s) {
  auto kern = ld_sweep_persistent_kernel<NSH, DC4, SRC, BF16>;
  int occ = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 64, lds);
  if (e != hipSuccess) return e;
  int dev = 0, cus = 0;
  e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  // resident waves only (at most the occupancy, fewer when asked), a multiple of the XCD count so that every XCD
  // (dispatched round robin) gets the same number
  const int per_cu = a.waves_per_cu > 0 ? std::min(a.waves_per_cu, std::max(1, occ)) : std::max(1, occ);
  int nwg = per_cu * cus;
  nwg = std::max(kQueues, nwg / kQueues * kQueues);
  hipLaunchKernelGGL(kern, dim3(nwg), dim3(64), lds, s, a);
  return hipGetLastError();
}

template <int NSH, int SRC, bool BF16>
hipError_t launch_persistent_dc(const PersistentArgs& a, size_t lds, hipStream_t s) {
  switch ((a.dc + 3) / 4) {
    case 1: return launch_persistent_t<NSH, 1, SRC, BF16>(a, lds, s);
    case 2: return launch_persistent_t<NSH, 2, SRC, BF16>(a, lds, s);
    case 3: return launch_persistent_t<NSH, 3, SRC, BF16>(a, lds, s);
    case 4: return launch_persistent_t<NSH, 4, SRC, BF16>(a, lds, s);
    case 5: return launch_persistent_t<NSH, 5, SRC, BF16>(a, lds, s);
    case 6: return launch_persistent_t<NSH, 6, SRC, BF16>(a, lds, s);
    case 7: return launch_persistent_t<NSH, 7, SRC, BF16>(a, lds, s);
    case 8: return launch_persistent_t<NSH, 8, SRC, BF16>(a, lds, s);
    default: return hipErrorInvalidValue;
  }
}

template <int NSH, bool BF16>
hipError_t launch_persistent_src(const PersistentArgs& a, int layout, size_t lds, hipStream_t s) {
  if (layout < 0) return launch_persistent_dc<NSH, 0, BF16>(a, lds, s);
  if (layout == 3) return launch_persistent_dc<NSH, 3, BF16>(a, lds, s);
  return layout == 2 ? launch_persistent_dc<NSH, 2, BF16>(a, lds, s) : launch_persistent_dc<NSH, 1, BF16>(a, lds, s);
}

}  // namespace
}  // namespace citrus
