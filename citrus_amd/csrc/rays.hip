// Exact line integrals of voxel fields (Siddon / Amanatides-Woo traversal) for gfx950.
//
// One thread traces one ray. The ray's end points are generated in the kernel (explicit segments, focal spot to
// lattice points, focal spot to pixel centres, pixel lines along a direction, voxel centres to the panel), so no
// per-ray geometry ever goes through HBM. The segment clipping, entry voxel and plane parameters are computed and
// compared in float64 with one IEEE operation per step (contraction is off in this file), which makes the voxel
// sequence identical to a float64 CPU traversal written with the same operations; the field values are summed in
// float32 with Kahan compensation (the shared-sequence downstream kernel adds its exact float32 products in float64),
// so long rays keep float32 relative accuracy. Fields are channel-last
// [rows][cols][slices][channels] so the channels of a voxel are one contiguous read.
//
// Traversal rules (citrus/scatter_solver/kernels/ray_tracing_kernels.py:178-327): slab clip to [a_lo, a_hi] within
// [0, 1]; entry index floor (ceil - 1 along a descending axis) clamped to the grid; step to the nearest plane, ties in
// the order x, y, z; the next plane parameter is recomputed from the voxel index at every step.

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "citrus_hip.h"
#include "common.h"

namespace citrus {
namespace {

constexpr int kRayThreads = 128;
constexpr double kBig = 1e30;

struct Box {
  double lo[3], hi[3], delta[3];
  int n[3];  // voxels along x (cols), y (rows), z (slices)
};

enum Mode : int {
  kSegments = 0,   // p0[i], p1[i] given (p0 may be one shared point)
  kLattice = 1,    // p0 = point, p1 = centre of lattice point i
  kPixels = 2,     // p0 = focal spot, p1 = (sub-)pixel centre
  kOblique = 3,    // line through pixel centre along direction k, clipped to the box
  kDownstream = 4  // voxel centre to the panel along direction k
};

struct Panel {
  double source[3], center[3], uh[3], vh[3], nh[3];
  double sdd;
  int curved, rows, cols;
  double drow, dch, off_row, off_ch;
};

struct RayJob {
  Box box;
  int nch;
  int64_t nrays;   // rays per direction (pixels, lattice points, voxels or segments)
  int K;           // directions (1 unless oblique / downstream)
  const float* vol;
  int64_t vol_k_stride;  // elements between the fields of consecutive directions (0: shared field)
  // segments
  const double* p0;
  const double* p1;
  int p0_shared;
  // lattice: rays from lat_point to origin + (j dx, i dy, k dz), shape (rows, cols, slices)
  double lat_point[3], lat_origin[3], lat_delta[3];
  int lat_shape[3];
  // pixels
  Panel panel;
  int samples;  // sub-rays per pixel axis
  // directions
  const double* omega;  // [K][3]
  double half_diag;     // oblique: half length of the segment
  double box_center[3];
  double diag;          // downstream: clamp of the panel distance
  // output: out[k * ok + c * oc + i] (i = ray within direction)
  float* out;
  int64_t ok, oc;
  // optional voxel-sequence capture (segments mode)
  int32_t* seq_idx;
  float* seq_len;
  int32_t* seq_count;
  int seq_max;
};

__device__ __forceinline__ double entry_index(double p0, double d, double a, double lo, double delta, int n) {
  const double f = (p0 + a * d - lo) / delta;
  double i = d < 0.0 ? ceil(f) - 1.0 : floor(f);
  i = fmax(i, 0.0);
  i = fmin(i, (double)(n - 1));
  return i;
}

__device__ __forceinline__ double plane_param(double lo, double delta, int idx, bool pos, double p0, double inv,
                                              bool nz) {
  const int k = idx + (pos ? 1 : 0);
  return nz ? (lo + (double)k * delta - p0) * inv : kBig;
}

// Integrate the channel-last field along p0 -> p1. Returns |p1 - p0| and accumulates chord * value per channel.
// INCR: the next plane's parameter along the crossed axis is the current one plus that axis's plane spacing
// (delta |1/d|) instead of being formed again from the plane's index: one addition instead of a conversion, a fused
// multiply-add and a product per step. The parameters then differ from the exact ones by rounding only, so the
// voxel sequence is the same except where two planes are crossed within that rounding of each other (a segment of
// that length); for the images, not for the captured sequences.
template <int CH, bool CAPTURE, bool INCR = false>
__device__ void trace(const RayJob& J, const double* p0, const double* p1, const float* vol, float* sum,
                      int32_t* sidx, float* slen, int32_t* scount) {
  const Box& B = J.box;
  double d[3], inv[3], amin[3], amax[3];
  bool nz[3], pos[3];
  for (int a = 0; a < 3; ++a) {
    d[a] = p1[a] - p0[a];
    nz[a] = d[a] != 0.0;
    inv[a] = 1.0 / (nz[a] ? d[a] : 1.0);
    const double a1 = (B.lo[a] - p0[a]) * inv[a];
    const double a2 = (B.hi[a] - p0[a]) * inv[a];
    const bool ins = (p0[a] >= B.lo[a]) && (p0[a] <= B.hi[a]);
    amin[a] = nz[a] ? fmin(a1, a2) : (ins ? -kBig : kBig);
    amax[a] = nz[a] ? fmax(a1, a2) : (ins ? kBig : -kBig);
    pos[a] = d[a] > 0.0;
  }
  double a_lo = fmax(fmax(fmax(amin[0], amin[1]), amin[2]), 0.0);
  double a_hi = fmin(fmin(fmin(amax[0], amax[1]), amax[2]), 1.0);
  const double length = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  int count = 0;
  if (a_hi > a_lo) {
    int idx[3];
    double an[3], spacing[3];
    int step[3];
    for (int a = 0; a < 3; ++a) {
      idx[a] = (int)entry_index(p0[a], d[a], a_lo, B.lo[a], B.delta[a], B.n[a]);
      step[a] = pos[a] ? 1 : (d[a] < 0.0 ? -1 : 0);
      an[a] = plane_param(B.lo[a], B.delta[a], idx[a], pos[a], p0[a], inv[a], nz[a]);
      spacing[a] = nz[a] ? B.delta[a] * fabs(inv[a]) : 0.0;
    }
    float comp[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) comp[c] = 0.0f;
    double a = a_lo;
    const int max_steps = B.n[0] + B.n[1] + B.n[2] + 3;
    for (int s = 0; s < max_steps; ++s) {
      // the axis crossed next, by the DDA's tie rule (x on a tie with y or z, y on a tie with z), from two
      // comparisons; its plane parameter is the least of the three
      const bool xy = an[0] <= an[1];
      const double a01 = xy ? an[0] : an[1];
      const bool first = a01 <= an[2];
      const double nxt = fmin(first ? a01 : an[2], a_hi);
      const double seg = fmax(nxt - a, 0.0);
      const bool inb = idx[0] >= 0 && idx[0] < B.n[0] && idx[1] >= 0 && idx[1] < B.n[1] && idx[2] >= 0 &&
                       idx[2] < B.n[2];
      if (inb && seg > 0.0) {
        const int64_t flat = ((int64_t)idx[1] * B.n[0] + idx[0]) * B.n[2] + idx[2];
        if (CAPTURE) {
          if (count < J.seq_max) {
            sidx[count] = (int32_t)flat;
            slen[count] = (float)seg;
          }
        } else {
          const float chord = (float)(seg * length);
          const float* v = vol + flat * J.nch;
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            if (c < J.nch) {  // Kahan-compensated sum
              const float y = v[c] * chord - comp[c];
              const float t = sum[c] + y;
              comp[c] = (t - sum[c]) - y;
              sum[c] = t;
            }
          }
        }
        ++count;
      }
      const bool sx = first && xy, sy = first && !xy, sz = !first;
      if (INCR) {
        if (sx) { idx[0] += step[0]; an[0] += spacing[0]; }
        if (sy) { idx[1] += step[1]; an[1] += spacing[1]; }
        if (sz) { idx[2] += step[2]; an[2] += spacing[2]; }
      } else {
        if (sx) { idx[0] += step[0]; an[0] = plane_param(B.lo[0], B.delta[0], idx[0], pos[0], p0[0], inv[0], nz[0]); }
        if (sy) { idx[1] += step[1]; an[1] = plane_param(B.lo[1], B.delta[1], idx[1], pos[1], p0[1], inv[1], nz[1]); }
        if (sz) { idx[2] += step[2]; an[2] = plane_param(B.lo[2], B.delta[2], idx[2], pos[2], p0[2], inv[2], nz[2]); }
      }
      if (!(nxt < a_hi)) break;
      a = nxt;
    }
  }
  if (CAPTURE) *scount = count;
}

__device__ void panel_point(const Panel& P, double u, double v, double* out) {
  if (!P.curved) {
    for (int a = 0; a < 3; ++a) out[a] = P.center[a] + u * P.uh[a] + v * P.vh[a];
    return;
  }
  const double g = u / P.sdd;
  const double su = P.sdd * sin(g);
  const double sh = sin(0.5 * g);
  const double sn = 2.0 * P.sdd * sh * sh;
  const double sv = v * cos(g);
  for (int a = 0; a < 3; ++a) out[a] = P.center[a] + su * P.uh[a] + sn * P.nh[a] + sv * P.vh[a];
}

__device__ __forceinline__ double pixel_u(const Panel& P, int col) {
  return ((double)col - (double)(P.cols - 1) / 2.0) * P.dch - P.off_ch;
}
__device__ __forceinline__ double pixel_v(const Panel& P, int row) {
  return ((double)row - (double)(P.rows - 1) / 2.0) * P.drow - P.off_row;
}

// panel distance of p along w (flat plane or the larger root on the cylinder), +inf when unreachable
__device__ double panel_distance(const Panel& P, const double* p, const double* w) {
  if (!P.curved) {
    const double den = w[0] * P.nh[0] + w[1] * P.nh[1] + w[2] * P.nh[2];
    const double num = (P.center[0] - p[0]) * P.nh[0] + (P.center[1] - p[1]) * P.nh[1] +
                       (P.center[2] - p[2]) * P.nh[2];
    return num / den;
  }
  const double dx = p[0] - P.source[0], dy = p[1] - P.source[1];
  const double a = w[0] * w[0] + w[1] * w[1];
  const double b = dx * w[0] + dy * w[1];
  const double c = dx * dx + dy * dy - P.sdd * P.sdd;
  const double disc = b * b - a * c;
  const double root = sqrt(fmax(disc, 0.0));
  const double den = b >= 0.0 ? b + root : (a > 0.0 ? a : 1.0);
  const double num = b >= 0.0 ? -c : root - b;
  if (!(a > 0.0 && disc >= 0.0 && den != 0.0)) return INFINITY;
  return num / den;
}

template <int MODE, int CH>
__global__ __launch_bounds__(kRayThreads) void ray_kernel(RayJob J) {
  const int64_t gid = (int64_t)blockIdx.x * kRayThreads + threadIdx.x;
  const int64_t total = J.nrays * J.K;
  if (gid >= total) return;
  const int k = (int)(gid / J.nrays);
  const int64_t i = gid - (int64_t)k * J.nrays;
  float sum[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) sum[c] = 0.0f;
  const float* vol = J.vol + (int64_t)k * J.vol_k_stride;
  double p0[3], p1[3];
  bool write_inf = false;
  if (MODE == kSegments) {
    const double* q0 = J.p0_shared ? J.p0 : J.p0 + 3 * i;
    for (int a = 0; a < 3; ++a) { p0[a] = q0[a]; p1[a] = J.p1[3 * i + a]; }
    if (J.seq_idx) {
      trace<CH, true>(J, p0, p1, vol, sum, J.seq_idx + i * J.seq_max, J.seq_len + i * J.seq_max,
                      J.seq_count + i);
      return;
    }
    trace<CH, false>(J, p0, p1, vol, sum, nullptr, nullptr, nullptr);
  } else if (MODE == kLattice) {
    const int S = J.lat_shape[2], C = J.lat_shape[1];
    const int kk = (int)(i % S), jj = (int)((i / S) % C), ii = (int)(i / ((int64_t)S * C));
    for (int a = 0; a < 3; ++a) p0[a] = J.lat_point[a];
    p1[0] = J.lat_origin[0] + (double)jj * J.lat_delta[0];
    p1[1] = J.lat_origin[1] + (double)ii * J.lat_delta[1];
    p1[2] = J.lat_origin[2] + (double)kk * J.lat_delta[2];
    trace<CH, false>(J, p0, p1, vol, sum, nullptr, nullptr, nullptr);
  } else if (MODE == kPixels) {
    const Panel& P = J.panel;
    const int row = (int)(i / P.cols), col = (int)(i % P.cols);
    const int n = J.samples;
    for (int a = 0; a < 3; ++a) p0[a] = P.source[a];
    if (n == 1) {
      panel_point(P, pixel_u(P, col), pixel_v(P, row), p1);
      trace<CH, false>(J, p0, p1, vol, sum, nullptr, nullptr, nullptr);
    } else {
      // -log(mean exp(-tau)) over the n x n midpoint lattice of the pixel, by a running log-sum-exp per channel
      float m[CH], acc[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) { m[c] = INFINITY; acc[c] = 0.0f; }
      for (int sv = 0; sv < n; ++sv)
        for (int su = 0; su < n; ++su) {
          const double ou = ((2.0 * su + 1.0) / (2.0 * n) - 0.5) * P.dch;
          const double ov = ((2.0 * sv + 1.0) / (2.0 * n) - 0.5) * P.drow;
          panel_point(P, pixel_u(P, col) + ou, pixel_v(P, row) + ov, p1);
          float t[CH];
#pragma unroll
          for (int c = 0; c < CH; ++c) t[c] = 0.0f;
          trace<CH, false>(J, p0, p1, vol, t, nullptr, nullptr, nullptr);
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            if (t[c] < m[c]) { acc[c] = acc[c] * expf(t[c] - m[c]) + 1.0f; m[c] = t[c]; }
            else acc[c] += expf(m[c] - t[c]);
          }
        }
#pragma unroll
      for (int c = 0; c < CH; ++c) sum[c] = m[c] - logf(acc[c] / (float)(n * n));
    }
  } else if (MODE == kOblique) {
    const Panel& P = J.panel;
    const int row = (int)(i / P.cols), col = (int)(i % P.cols);
    double p[3];
    panel_point(P, pixel_u(P, col), pixel_v(P, row), p);
    const double* w = J.omega + 3 * k;
    const double w2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double cp = ((J.box_center[0] - p[0]) * w[0] + (J.box_center[1] - p[1]) * w[1] +
                       (J.box_center[2] - p[2]) * w[2]) / w2;
    const double h = J.half_diag / sqrt(w2);
    for (int a = 0; a < 3; ++a) {
      p0[a] = p[a] + (cp - h) * w[a];
      p1[a] = p[a] + (cp + h) * w[a];
    }
    trace<CH, false, true>(J, p0, p1, vol, sum, nullptr, nullptr, nullptr);
  } else {  // kDownstream: voxel centre i of the traced grid to the panel along omega[k]
    const Box& B = J.box;
    const int S = B.n[2], C = B.n[0];
    const int kk = (int)(i % S), jj = (int)((i / S) % C), ii = (int)(i / ((int64_t)S * C));
    const double* w = J.omega + 3 * k;
    const Panel& P = J.panel;
    const double toward = -(w[0] * P.nh[0] + w[1] * P.nh[1] + w[2] * P.nh[2]);
    p0[0] = B.lo[0] + ((double)jj + 0.5) * B.delta[0];
    p0[1] = B.lo[1] + ((double)ii + 0.5) * B.delta[1];
    p0[2] = B.lo[2] + ((double)kk + 0.5) * B.delta[2];
    if (!(toward > 0.0)) {
      write_inf = true;
    } else {
      double t = panel_distance(P, p0, w);
      t = fmin(fmax(t, 0.0), J.diag);
      for (int a = 0; a < 3; ++a) p1[a] = p0[a] + t * w[a];
      trace<CH, false>(J, p0, p1, vol, sum, nullptr, nullptr, nullptr);
    }
  }
  float* out = J.out + (int64_t)k * J.ok + i;
#pragma unroll
  for (int c = 0; c < CH; ++c)
    if (c < J.nch) out[(int64_t)c * J.oc] = write_inf ? INFINITY : sum[c];
}

template <int MODE>
hipError_t launch_mode(const RayJob& J, hipStream_t s) {
  const int64_t total = J.nrays * J.K;
  if (total == 0) return hipSuccess;
  const int64_t blocks = (total + kRayThreads - 1) / kRayThreads;
  if (!grid_fits(blocks, kRayThreads)) return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks);
  if (J.nch <= 1) hipLaunchKernelGGL((ray_kernel<MODE, 1>), grid, dim3(kRayThreads), 0, s, J);
  else if (J.nch <= 2) hipLaunchKernelGGL((ray_kernel<MODE, 2>), grid, dim3(kRayThreads), 0, s, J);
  else if (J.nch <= 4) hipLaunchKernelGGL((ray_kernel<MODE, 4>), grid, dim3(kRayThreads), 0, s, J);
  else if (J.nch <= 8) hipLaunchKernelGGL((ray_kernel<MODE, 8>), grid, dim3(kRayThreads), 0, s, J);
  else if (J.nch <= 16) hipLaunchKernelGGL((ray_kernel<MODE, 16>), grid, dim3(kRayThreads), 0, s, J);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

int init_job(RayJob& J, const float* volume, int R, int C, int S, int nch, const double* box6) {
  if (R < 1 || C < 1 || S < 1) return fail(CITRUS_ERR_ARG, "grid shape must be positive");
  if (nch < 1 || nch > 16) return fail(CITRUS_ERR_ARG, "1..16 channels supported");
  if (!box6) return fail(CITRUS_ERR_ARG, "box is NULL");
  J = RayJob{};
  for (int a = 0; a < 3; ++a) {
    J.box.lo[a] = box6[a];
    J.box.delta[a] = box6[3 + a];
  }
  J.box.n[0] = C; J.box.n[1] = R; J.box.n[2] = S;
  for (int a = 0; a < 3; ++a) J.box.hi[a] = J.box.lo[a] + J.box.delta[a] * (double)J.box.n[a];
  J.nch = nch;
  J.vol = volume;
  J.K = 1;
  J.samples = 1;
  return CITRUS_OK;
}

int set_panel(RayJob& J, const double* panel20, const int* panel_ints) {
  if (!panel20 || !panel_ints) return fail(CITRUS_ERR_ARG, "panel description is NULL");
  Panel& P = J.panel;
  for (int a = 0; a < 3; ++a) {
    P.source[a] = panel20[a];
    P.center[a] = panel20[3 + a];
    P.uh[a] = panel20[6 + a];
    P.vh[a] = panel20[9 + a];
    P.nh[a] = panel20[12 + a];
  }
  P.sdd = panel20[15];
  P.drow = panel20[16];
  P.dch = panel20[17];
  P.off_row = panel20[18];
  P.off_ch = panel20[19];
  P.curved = panel_ints[0];
  P.rows = panel_ints[1];
  P.cols = panel_ints[2];
  if (P.rows < 1 || P.cols < 1) return fail(CITRUS_ERR_ARG, "panel must have rows and columns");
  return CITRUS_OK;
}

// ---- downstream optical depth from a shared crossing sequence (ray_tracing_kernels.py:518-639) ---------------------
// Rays that start at voxel centres along one direction are translates of each other by whole voxels, so they cross
// the same sequence of (voxel offset, chord) until they leave the grid. One thread per direction traverses it once
// from the centre of a voxel of an unbounded lattice (float64, the plane-stepping and tie rules of trace()) into a
// record; every voxel then sums its field over the first samples of the record that stay inside the grid, cut at its
// own panel distance.
//
// Record of one direction, in 32-bit words (kSeqHeader + 3 max_len + 1 + (C + 1) + (R + 1) + (S + 1), rounded up to
// a multiple of 4): {count, 0, 0, 0}, flat voxel offset [max_len] (int), chord [max_len] (cm, float), prefix
// [max_len + 1] (float, sum of the chords before each sample), then per axis x, y, z the table first[m] (int,
// m = 0..n_axis) of the first sample whose offset along that axis has magnitude >= m (count when none).
// A voxel at index j along a moving axis leaves the grid at sample first[n - j] (positive step) or first[j + 1]
// (negative step); its in-grid samples are [0, min over the axes).

constexpr int kSeqHeader = 4;

__host__ __device__ inline int seq_max_len(int R, int C, int S) { return R + C + S + 3; }

__host__ __device__ inline int64_t seq_record_words(int R, int C, int S) {
  const int64_t L = seq_max_len(R, C, S);
  const int64_t w = kSeqHeader + 3 * L + 1 + (int64_t)(C + 1) + (R + 1) + (S + 1);
  return (w + 3) & ~(int64_t)3;
}

__global__ void crossing_sequence_kernel(Box B, const double* __restrict__ omega, int K, int* __restrict__ records,
                                         int64_t words) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const int R = B.n[1], C = B.n[0], S = B.n[2];
  const int max_len = seq_max_len(R, C, S);
  int* rec = records + (int64_t)k * words;
  int* flat = rec + kSeqHeader;
  float* chord = reinterpret_cast<float*>(flat + max_len);
  float* prefix = chord + max_len;
  int* first[3];
  first[0] = reinterpret_cast<int*>(prefix + max_len + 1);  // x: C + 1 entries
  first[1] = first[0] + (C + 1);                            // y: R + 1 entries
  first[2] = first[1] + (R + 1);                            // z: S + 1 entries
  const double* w = omega + 3 * k;
  const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  int idx[3] = {0, 0, 0}, step[3];
  double an[3], inv[3];
  bool nz[3], pos[3];
  for (int a = 0; a < 3; ++a) {
    nz[a] = w[a] != 0.0;
    pos[a] = w[a] > 0.0;
    inv[a] = 1.0 / (nz[a] ? w[a] : 1.0);
    step[a] = pos[a] ? 1 : (w[a] < 0.0 ? -1 : 0);
    first[a][0] = 0;
  }
  // parameter along w of the next plane crossed on axis a, from the centre of voxel 0 of the lattice
  auto plane = [&](int a) {
    return nz[a] ? ((double)(idx[a] + (pos[a] ? 1 : 0)) * B.delta[a] - 0.5 * B.delta[a]) * inv[a] : kBig;
  };
  for (int a = 0; a < 3; ++a) an[a] = plane(a);
  double t = 0.0, cum = 0.0;
  int count = 0;
  for (int s = 0; s < max_len; ++s) {
    const double nxt = fmin(fmin(an[0], an[1]), an[2]);
    const double seg = nxt - t;
    if (seg > 0.0 && count < max_len) {
      flat[count] = (idx[1] * C + idx[0]) * S + idx[2];
      chord[count] = (float)(seg * wn);
      prefix[count] = (float)cum;
      cum += seg * wn;
      ++count;
    }
    const bool sx = (an[0] <= an[1]) && (an[0] <= an[2]);
    const bool sy = (an[1] < an[0]) && (an[1] <= an[2]);
    const bool sz = (an[2] < an[0]) && (an[2] < an[1]);
    const int a = sx ? 0 : (sy ? 1 : 2);
    if (!(sx || sy || sz) || !nz[a]) break;  // no moving axis (w = 0)
    idx[a] += step[a];
    an[a] = plane(a);
    t = nxt;
    const int m = abs(idx[a]);
    first[a][m] = count;  // the next stored sample is the first with this offset magnitude along a
    if (m >= B.n[a]) break;  // past every voxel of the grid along a: no start voxel reaches further
  }
  prefix[count] = (float)cum;
  for (int a = 0; a < 3; ++a)
    for (int m = abs(idx[a]) + 1; m <= B.n[a]; ++m) first[a][m] = count;
  rec[0] = count;
  rec[1] = rec[2] = rec[3] = 0;
}

// The channels of one voxel at byte offset `off` of the volume: one vector load for 2 or 4 channels.
template <int CH>
__device__ __forceinline__ void load_channels(const char* __restrict__ vb, unsigned off, float v[CH]) {
  if constexpr (CH == 2) {
    const float2 w = *reinterpret_cast<const float2*>(vb + off);
    v[0] = w.x; v[1] = w.y;
  } else if constexpr (CH == 4) {
    const float4 w = *reinterpret_cast<const float4*>(vb + off);
    v[0] = w.x; v[1] = w.y; v[2] = w.z; v[3] = w.w;
  } else {
#pragma unroll
    for (int c = 0; c < CH; ++c) v[c] = reinterpret_cast<const float*>(vb + off)[c];
  }
}

// EXACT: nch == CH and the volume is below 2^31 bytes. The samples of a ray that leaves the grid before the panel are
// then taken eight at a time: their offsets and chords from LDS and their channels (32-bit byte offsets, one vector
// load per voxel) are all requested before the first is added, and the additions run in sample order, so the sums are
// those of the one-sample loop bit for bit.
template <int CH, bool EXACT>
__global__ __launch_bounds__(256) void downstream_shared_kernel(Box B, Panel P, const float* __restrict__ vol,
                                                                int nch, const double* __restrict__ omega,
                                                                const int* __restrict__ records, int64_t words,
                                                                const float* __restrict__ mix, int nout,
                                                                float* __restrict__ out) {
  extern __shared__ int srec[];
  const int k = blockIdx.y;
  const int R = B.n[1], C = B.n[0], S = B.n[2];
  const int64_t N = (int64_t)R * C * S;
  const int max_len = seq_max_len(R, C, S);
  const int* rec = records + (int64_t)k * words;
  for (int s = threadIdx.x; s < (int)words; s += blockDim.x) srec[s] = rec[s];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int* flat = srec + kSeqHeader;
  const float* chord = reinterpret_cast<const float*>(flat + max_len);
  const float* prefix = chord + max_len;
  const int* first_x = reinterpret_cast<const int*>(prefix + max_len + 1);
  const int* first_y = first_x + (C + 1);
  const int* first_z = first_y + (R + 1);
  const int kk = (int)(i % S), jj = (int)((i / S) % C), ii = (int)(i / ((int64_t)S * C));
  const double* w = omega + 3 * k;
  const double toward = -(w[0] * P.nh[0] + w[1] * P.nh[1] + w[2] * P.nh[2]);
  float* o = out + (int64_t)k * nout * N + i;
  if (!(toward > 0.0)) {
    for (int g = 0; g < nout; ++g) o[(int64_t)g * N] = INFINITY;
    return;
  }
  const int lx = w[0] > 0.0 ? C - jj : (w[0] < 0.0 ? jj + 1 : C);
  const int ly = w[1] > 0.0 ? R - ii : (w[1] < 0.0 ? ii + 1 : R);
  const int lz = w[2] > 0.0 ? S - kk : (w[2] < 0.0 ? kk + 1 : S);
  const int n_in = min(first_x[lx], min(first_y[ly], first_z[lz]));
  double p0[3] = {B.lo[0] + ((double)jj + 0.5) * B.delta[0], B.lo[1] + ((double)ii + 0.5) * B.delta[1],
                  B.lo[2] + ((double)kk + 0.5) * B.delta[2]};
  const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double tp = fmax(panel_distance(P, p0, w), 0.0) * wn;  // cm to the panel
  const int nc = EXACT ? CH : nch;
  const float* base = vol + i * nc;
  // FP64 sums: a float32 value times a float32 length is exact in FP64, so each term is added with one rounding
  // (2 vector operations per channel and crossing against 5 for a compensated float32 sum, at the same FP64 rate)
  double sum[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) sum[c] = 0.0;
  auto add = [&](int s, float len) {
    const float* v = base + (int64_t)flat[s] * nc;
    const double l = (double)len;
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (c < nc) sum[c] = fma((double)v[c], l, sum[c]);
  };
  if ((double)prefix[n_in] * (1.0 + 1e-5) + 1e-6 < tp) {
    // the panel lies beyond the grid exit: every in-grid sample counts in full
    int s = 0;
    if constexpr (EXACT) {
      constexpr int U = 8;
      const char* vb = reinterpret_cast<const char*>(vol);
      const unsigned ib = (unsigned)i * (unsigned)(4 * CH);
      for (; s + U <= n_in; s += U) {
        float v[U][CH], len[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          len[u] = chord[s + u];
          load_channels<CH>(vb, ib + (unsigned)(flat[s + u] * (4 * CH)), v[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const double l = (double)len[u];
#pragma unroll
          for (int c = 0; c < CH; ++c) sum[c] = fma((double)v[u][c], l, sum[c]);
        }
      }
    }
    for (; s < n_in; ++s) add(s, chord[s]);
  } else {
    double cum = 0.0;
    for (int s = 0; s < n_in; ++s) {
      const double seg = (double)chord[s];
      if (cum + seg >= tp) {
        add(s, (float)(tp - cum));
        break;
      }
      add(s, chord[s]);
      cum += seg;
    }
  }
  if (mix) {  // one output per row of the mix table: sum over the traced channels
    for (int g = 0; g < nout; ++g) {
      float acc = 0.0f;
#pragma unroll
      for (int c = 0; c < CH; ++c)
        if (c < nc) acc = fmaf(mix[g * nc + c], (float)sum[c], acc);
      o[(int64_t)g * N] = acc;
    }
  } else {
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (c < nc) o[(int64_t)c * N] = (float)sum[c];
  }
}

int finish(hipError_t e, const char* what) {
  if (e == hipErrorInvalidValue) return fail(CITRUS_ERR_ARG, std::string(what) + ": too many rays or channels");
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return CITRUS_OK;
}

}  // namespace
}  // namespace citrus

using namespace citrus;

extern "C" int citrus_ray_segments(const float* volume, int R, int C, int S, int nch, const double* box,
                                   const double* p0, const double* p1, int p0_shared, int64_t n, float* out,
                                   void* stream) {
  RayJob J;
  if (int rc = init_job(J, volume, R, C, S, nch, box)) return rc;
  if (!volume || !p0 || !p1 || !out) return fail(CITRUS_ERR_ARG, "NULL buffer");
  J.p0 = p0; J.p1 = p1; J.p0_shared = p0_shared; J.nrays = n;
  J.out = out; J.ok = 0; J.oc = n;
  return finish(launch_mode<kSegments>(J, reinterpret_cast<hipStream_t>(stream)), "citrus_ray_segments");
}

extern "C" int citrus_ray_trace_indices(int R, int C, int S, const double* box, const double* p0, const double* p1,
                                        int64_t n, int max_len, int32_t* seq_idx, float* seq_len,
                                        int32_t* seq_count, void* stream) {
  RayJob J;
  if (int rc = init_job(J, nullptr, R, C, S, 1, box)) return rc;
  if (!p0 || !p1 || !seq_idx || !seq_len || !seq_count || max_len < 1) return fail(CITRUS_ERR_ARG, "NULL buffer");
  J.p0 = p0; J.p1 = p1; J.nrays = n;
  J.seq_idx = seq_idx; J.seq_len = seq_len; J.seq_count = seq_count; J.seq_max = max_len;
  return finish(launch_mode<kSegments>(J, reinterpret_cast<hipStream_t>(stream)), "citrus_ray_trace_indices");
}

extern "C" int citrus_ray_lattice(const float* volume, int R, int C, int S, int nch, const double* box,
                                  const double* point, const double* lattice_origin, const double* lattice_delta,
                                  const int* lattice_shape, float* out, void* stream) {
  RayJob J;
  if (int rc = init_job(J, volume, R, C, S, nch, box)) return rc;
  if (!volume || !point || !lattice_origin || !lattice_delta || !lattice_shape || !out)
    return fail(CITRUS_ERR_ARG, "NULL buffer");
  for (int a = 0; a < 3; ++a) {
    J.lat_point[a] = point[a];
    J.lat_origin[a] = lattice_origin[a];
    J.lat_delta[a] = lattice_delta[a];
    J.lat_shape[a] = lattice_shape[a];
  }
  J.nrays = (int64_t)lattice_shape[0] * lattice_shape[1] * lattice_shape[2];
  J.out = out; J.ok = 0; J.oc = J.nrays;
  return finish(launch_mode<kLattice>(J, reinterpret_cast<hipStream_t>(stream)), "citrus_ray_lattice");
}

extern "C" int citrus_ray_primary(const float* volume, int R, int C, int S, int nch, const double* box,
                                  const double* panel, const int* panel_ints, int samples, float* out,
                                  void* stream) {
  RayJob J;
  if (int rc = init_job(J, volume, R, C, S, nch, box)) return rc;
  if (int rc = set_panel(J, panel, panel_ints)) return rc;
  if (!volume || !out) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if (samples < 1 || samples > 16) return fail(CITRUS_ERR_ARG, "samples per pixel axis must be 1..16");
  J.samples = samples;
  J.nrays = (int64_t)J.panel.rows * J.panel.cols;
  J.out = out; J.ok = 0; J.oc = J.nrays;
  return finish(launch_mode<kPixels>(J, reinterpret_cast<hipStream_t>(stream)), "citrus_ray_primary");
}

extern "C" int citrus_ray_oblique(const float* field, int64_t field_k_stride, int R, int C, int S, int nch,
                                  const double* box, const double* panel, const int* panel_ints,
                                  const double* omega, int K, float* out, int64_t out_k_stride,
                                  int64_t out_c_stride, void* stream) {
  RayJob J;
  if (int rc = init_job(J, field, R, C, S, nch, box)) return rc;
  if (int rc = set_panel(J, panel, panel_ints)) return rc;
  if (!field || !omega || !out || K < 1) return fail(CITRUS_ERR_ARG, "NULL buffer or no directions");
  J.vol_k_stride = field_k_stride;
  J.omega = omega;
  J.K = K;
  double e2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    J.box_center[a] = J.box.lo[a] + 0.5 * (J.box.hi[a] - J.box.lo[a]);
    e2 += (J.box.hi[a] - J.box.lo[a]) * (J.box.hi[a] - J.box.lo[a]);
  }
  J.half_diag = 0.5 * sqrt(e2) * (1.0 + 1e-6) + 1e-6;
  J.nrays = (int64_t)J.panel.rows * J.panel.cols;
  J.out = out; J.ok = out_k_stride; J.oc = out_c_stride;
  return finish(launch_mode<kOblique>(J, reinterpret_cast<hipStream_t>(stream)), "citrus_ray_oblique");
}

extern "C" int citrus_ray_downstream(const float* volume, int R, int C, int S, int nch, const double* box,
                                     const double* panel, const int* panel_ints, const double* omega, int K,
                                     float* out, void* stream) {
  RayJob J;
  if (int rc = init_job(J, volume, R, C, S, nch, box)) return rc;
  if (int rc = set_panel(J, panel, panel_ints)) return rc;
  if (!volume || !omega || !out || K < 1) return fail(CITRUS_ERR_ARG, "NULL buffer or no directions");
  J.omega = omega;
  J.K = K;
  double e2 = 0.0;
  for (int a = 0; a < 3; ++a) e2 += (J.box.hi[a] - J.box.lo[a]) * (J.box.hi[a] - J.box.lo[a]);
  J.diag = sqrt(e2) * (1.0 + 1e-9) + 1e-9;
  J.nrays = (int64_t)R * C * S;
  J.out = out; J.oc = J.nrays; J.ok = (int64_t)nch * J.nrays;
  return finish(launch_mode<kDownstream>(J, reinterpret_cast<hipStream_t>(stream)), "citrus_ray_downstream");
}

namespace {
int check_sequence_grid(int R, int C, int S) {
  if (R < 1 || C < 1 || S < 1) return citrus::fail(CITRUS_ERR_ARG, "empty grid");
  if (R > 511 || C > 511 || S > 511) return citrus::fail(CITRUS_ERR_UNSUPPORTED, "at most 511 voxels per axis");
  return CITRUS_OK;
}
}  // namespace

extern "C" int64_t citrus_ray_crossing_sequence_bytes(int R, int C, int S) {
  if (R < 1 || C < 1 || S < 1 || R > 511 || C > 511 || S > 511) return -1;
  return 4 * citrus::seq_record_words(R, C, S);
}

extern "C" int citrus_ray_crossing_sequences(int R, int C, int S, const double* box, const double* omega, int K,
                                             void* sequences, void* stream) {
  using namespace citrus;
  if (int rc = check_sequence_grid(R, C, S)) return rc;
  RayJob J;
  if (int rc = init_job(J, nullptr, R, C, S, 1, box)) return rc;
  if (K < 0) return fail(CITRUS_ERR_ARG, "K must be >= 0");
  if (K == 0) return CITRUS_OK;
  if (!omega || !sequences) return fail(CITRUS_ERR_ARG, "NULL buffer");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(crossing_sequence_kernel, dim3((K + 63) / 64), dim3(64), 0, s, J.box, omega, K,
                     reinterpret_cast<int*>(sequences), seq_record_words(R, C, S));
  return finish(hipGetLastError(), "citrus_ray_crossing_sequences");
}

extern "C" int citrus_ray_downstream_shared(const float* volume, int R, int C, int S, int nch, const double* box,
                                            const double* panel, const int* panel_ints, const double* omega, int K,
                                            const void* sequences, const float* mix, int nout, float* out,
                                            void* stream) {
  using namespace citrus;
  if (int rc = check_sequence_grid(R, C, S)) return rc;
  RayJob J;
  if (int rc = init_job(J, volume, R, C, S, nch, box)) return rc;
  if (int rc = set_panel(J, panel, panel_ints)) return rc;
  if (K < 0 || K > 65535) return fail(CITRUS_ERR_ARG, "0..65535 directions per launch");
  if (mix ? (nout < 1 || nout > 64) : nout != nch) return fail(CITRUS_ERR_ARG, "nout must be nch without a mix table");
  if (K == 0) return CITRUS_OK;
  if (!volume || !omega || !out || !sequences) return fail(CITRUS_ERR_ARG, "NULL buffer");
  const int64_t words = seq_record_words(R, C, S);
  const int* rec = reinterpret_cast<const int*>(sequences);
  const int64_t N = (int64_t)R * C * S;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)((N + 255) / 256), (unsigned)K), block(256);
  const size_t lds = (size_t)words * 4;
  auto launch = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, block, lds, s, J.box, J.panel, volume, nch, omega, rec, words, mix, nout, out);
  };
  const bool small = N * nch * 4 < ((int64_t)1 << 31);  // 32-bit byte offsets of the batched samples
  if (small && nch == 1) launch(downstream_shared_kernel<1, true>);
  else if (small && nch == 2) launch(downstream_shared_kernel<2, true>);
  else if (small && nch == 3) launch(downstream_shared_kernel<3, true>);
  else if (small && nch == 4) launch(downstream_shared_kernel<4, true>);
  else if (nch <= 4) launch(downstream_shared_kernel<4, false>);
  else if (nch <= 8) launch(downstream_shared_kernel<8, false>);
  else launch(downstream_shared_kernel<16, false>);
  return finish(hipGetLastError(), "citrus_ray_downstream_shared");
}
