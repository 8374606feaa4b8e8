// Group transmission and the analytic uncollided fluence for gfx950.
//
// Group transmission T_g = sum_{e in g} w_e exp(-sum_m mu_m(e) P_m - mu_f(e) P_f) over the fine energies of the
// library, one thread per ray (citrus/scatter_solver/cross_sections.py:797-878,
// citrus/scatter_solver/kernels/transmission_kernels.py:190-275): the fine tables are staged in LDS, the ray's paths
// sit in registers, and every group sum is a register accumulator, so each ray costs one read of its paths and one
// write of its G transmissions.
//
// The uncollided fluence fuses the transmission with the source's angular weight and the inverse square law at the
// points of a lattice (citrus/scatter_solver/uncollided.py:160-197, source.py:289-446):
//   phi_UC,g(r) = b(Omega(r)) / |r - r0|^2 T_g(P(r), P_f(Omega(r))),
// b = forward * field weight (collimated field, off-focal fraction past its edge) * emission profile at the ray's panel
// point (bilinear, constant beyond the outermost pixel centres); P_f = filter density * thickness at that point.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "citrus_hip.h"
#include "common.h"

namespace citrus {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxGroups = 32;
constexpr int kMaxMaterials = 8;
constexpr int kMaxOffFocal = 16;

struct FineTables {
  const float* mu;       // [M][E]
  const int* group;      // [E]
  const float* weight;   // [E]
  const float* filter_mu;  // [E] or null
  int M, E, G;
};

struct SourceParams {
  double source[3], center[3], uh[3], vh[3], nh[3];
  double sdd;
  int curved, rows, cols;
  double drow, dch, off_row, off_ch;
  double margin_ch, margin_row;
  int n_off;
  double off_dist[kMaxOffFocal], off_frac[kMaxOffFocal];
  const float* profile;    // [rows][cols] or null
  const float* thickness;  // [rows][cols] or null
  double filter_density;
};

__device__ __forceinline__ void transmission(const FineTables& F, const float* smem_mu, const int* smem_group,
                                             const float* smem_w, const float* smem_fmu, const float* path,
                                             float path_f, float* T) {
  for (int g = 0; g < F.G; ++g) T[g] = 0.0f;
  for (int e = 0; e < F.E; ++e) {
    float tau = 0.0f;
    for (int m = 0; m < F.M; ++m) tau = fmaf(smem_mu[m * F.E + e], path[m], tau);
    if (smem_fmu) tau = fmaf(smem_fmu[e], path_f, tau);
    const float t = smem_w[e] * expf(-tau);
    const int g = smem_group[e];
#pragma unroll
    for (int gg = 0; gg < kMaxGroups; ++gg)
      if (gg == g) T[gg] += t;
  }
}

__device__ int stage_tables(const FineTables& F, float* lds) {
  // LDS layout: mu [M*E] | weight [E] | filter_mu [E] | group [E] (ints)
  for (int i = threadIdx.x; i < F.M * F.E; i += blockDim.x) lds[i] = F.mu[i];
  for (int i = threadIdx.x; i < F.E; i += blockDim.x) {
    lds[F.M * F.E + i] = F.weight[i];
    lds[(F.M + 1) * F.E + i] = F.filter_mu ? F.filter_mu[i] : 0.0f;
    reinterpret_cast<int*>(lds)[(F.M + 2) * F.E + i] = F.group[i];
  }
  __syncthreads();
  return 0;
}

__global__ __launch_bounds__(kThreads) void group_transmission_kernel(FineTables F, const float* paths,
                                                                      int64_t nrays, const float* filter_path,
                                                                      float* out) {
  extern __shared__ float lds[];
  stage_tables(F, lds);
  const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (r >= nrays) return;
  float path[kMaxMaterials];
  for (int m = 0; m < F.M; ++m) path[m] = paths[(int64_t)m * nrays + r];
  const float pf = filter_path ? filter_path[r] : 0.0f;
  float T[kMaxGroups];
  transmission(F, lds, reinterpret_cast<const int*>(lds) + (F.M + 2) * F.E, lds + F.M * F.E,
               F.filter_mu ? lds + (F.M + 1) * F.E : nullptr, path, pf, T);
#pragma unroll
  for (int g = 0; g < kMaxGroups; ++g)
    if (g < F.G) out[(int64_t)g * nrays + r] = T[g];
}

__device__ float bilinear(const float* img, int rows, int cols, double u, double v, const SourceParams& P) {
  double ch = (u + P.off_ch) / P.dch + (cols - 1) / 2.0;
  double rw = (v + P.off_row) / P.drow + (rows - 1) / 2.0;
  ch = fmin(fmax(ch, 0.0), cols - 1.0);
  rw = fmin(fmax(rw, 0.0), rows - 1.0);
  int c0 = (int)floor(ch), r0 = (int)floor(rw);
  c0 = c0 < 0 ? 0 : (c0 > cols - 2 ? (cols - 2 > 0 ? cols - 2 : 0) : c0);
  r0 = r0 < 0 ? 0 : (r0 > rows - 2 ? (rows - 2 > 0 ? rows - 2 : 0) : r0);
  const int c1 = c0 + 1 < cols ? c0 + 1 : cols - 1;
  const int r1 = r0 + 1 < rows ? r0 + 1 : rows - 1;
  const float wc = (float)fmin(fmax(ch - c0, 0.0), 1.0);
  const float wr = (float)fmin(fmax(rw - r0, 0.0), 1.0);
  const float top = img[r0 * cols + c0] + (img[r0 * cols + c1] - img[r0 * cols + c0]) * wc;
  const float bot = img[r1 * cols + c0] + (img[r1 * cols + c1] - img[r1 * cols + c0]) * wc;
  return top + (bot - top) * wr;
}

// Panel coordinates of the ray from the focal spot along d; false if it does not reach the panel forward.
__device__ bool panel_uv(const SourceParams& P, const double* d, double& u, double& v) {
  double t;
  if (!P.curved) {
    const double den = d[0] * P.nh[0] + d[1] * P.nh[1] + d[2] * P.nh[2];
    const double num = (P.center[0] - P.source[0]) * P.nh[0] + (P.center[1] - P.source[1]) * P.nh[1] +
                       (P.center[2] - P.source[2]) * P.nh[2];
    t = num / den;
  } else {  // from the focal spot the cylinder root is t = SDD / |d_xy|
    const double a = d[0] * d[0] + d[1] * d[1];
    t = a > 0.0 ? P.sdd / sqrt(a) : INFINITY;
  }
  if (!(isfinite(t) && t > 0.0)) return false;
  double hit[3];
  for (int k = 0; k < 3; ++k) hit[k] = P.source[k] + t * d[k];
  if (!P.curved) {
    double q[3];
    for (int k = 0; k < 3; ++k) q[k] = hit[k] - P.center[k];
    u = q[0] * P.uh[0] + q[1] * P.uh[1] + q[2] * P.uh[2];
    v = q[0] * P.vh[0] + q[1] * P.vh[1] + q[2] * P.vh[2];
  } else {
    double q[3];
    for (int k = 0; k < 3; ++k) q[k] = hit[k] - P.source[k];
    const double depth = -(q[0] * P.nh[0] + q[1] * P.nh[1] + q[2] * P.nh[2]);
    u = P.sdd * atan2(q[0] * P.uh[0] + q[1] * P.uh[1] + q[2] * P.uh[2], depth);
    v = P.sdd * (q[0] * P.vh[0] + q[1] * P.vh[1] + q[2] * P.vh[2]) / depth;
  }
  return true;
}

__device__ double field_weight(const SourceParams& P, double u, double v) {
  const double half_u = P.cols * P.dch / 2.0, half_v = P.rows * P.drow / 2.0;
  const double ou = fabs(u + P.off_ch), ov = fabs(v + P.off_row);
  const bool inside = ou <= half_u + P.margin_ch && ov <= half_v + P.margin_row;
  if (inside) return 1.0;
  if (P.n_off == 0) return 0.0;
  const double pu = fmax(ou - (half_u + P.margin_ch), 0.0), pv = fmax(ov - (half_v + P.margin_row), 0.0);
  const double dist = sqrt(pu * pu + pv * pv);
  const int n = P.n_off;
  if (dist > P.off_dist[n - 1]) return 0.0;
  int above = 0;  // number of nodes <= dist
  while (above < n && P.off_dist[above] <= dist) ++above;
  const int i0 = above - 1 < 0 ? 0 : (above - 1 > n - 1 ? n - 1 : above - 1);
  const int i1 = above > n - 1 ? n - 1 : above;
  const double span = P.off_dist[i1] - P.off_dist[i0];
  double w = span > 0.0 ? (dist - P.off_dist[i0]) / span : 0.0;
  w = fmin(fmax(w, 0.0), 1.0);
  return P.off_frac[i0] + (P.off_frac[i1] - P.off_frac[i0]) * w;
}

__global__ __launch_bounds__(kThreads) void uncollided_kernel(FineTables F, SourceParams P, const float* paths,
                                                              int64_t npts, double ox, double oy, double oz,
                                                              double dx, double dy, double dz, int Cq, int Sq,
                                                              float* out) {
  extern __shared__ float lds[];
  stage_tables(F, lds);
  const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (r >= npts) return;
  const int k = (int)(r % Sq), j = (int)((r / Sq) % Cq), i = (int)(r / ((int64_t)Sq * Cq));
  double d[3] = {ox + j * dx - P.source[0], oy + i * dy - P.source[1], oz + k * dz - P.source[2]};
  double u = 0.0, v = 0.0;
  const bool forward = panel_uv(P, d, u, v);
  double b = 0.0;
  float pf = 0.0f;
  if (forward) {
    b = field_weight(P, u, v);
    if (P.profile && b != 0.0) b *= (double)bilinear(P.profile, P.rows, P.cols, u, v, P);
  }
  if (P.thickness) pf = (float)((double)bilinear(P.thickness, P.rows, P.cols, u, v, P) * P.filter_density);
  const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float scale = (float)(b / r2);
  float path[kMaxMaterials];
  for (int m = 0; m < F.M; ++m) path[m] = paths[(int64_t)m * npts + r];
  float T[kMaxGroups];
  transmission(F, lds, reinterpret_cast<const int*>(lds) + (F.M + 2) * F.E, lds + F.M * F.E,
               F.filter_mu ? lds + (F.M + 1) * F.E : nullptr, path, pf, T);
#pragma unroll
  for (int g = 0; g < kMaxGroups; ++g)
    if (g < F.G) out[(int64_t)g * npts + r] = T[g] * scale;
}

int check_tables(const FineTables& F) {
  if (F.M < 1 || F.M > kMaxMaterials) return fail(CITRUS_ERR_ARG, "1..8 traced materials supported");
  if (F.G < 1 || F.G > kMaxGroups) return fail(CITRUS_ERR_ARG, "1..32 groups supported");
  if (F.E < 1) return fail(CITRUS_ERR_ARG, "no fine energies");
  if (!F.mu || !F.group || !F.weight) return fail(CITRUS_ERR_ARG, "NULL fine table");
  if (sizeof(float) * (size_t)(F.M + 3) * F.E > 64 * 1024) return fail(CITRUS_ERR_UNSUPPORTED, "fine tables > 64 KiB");
  return CITRUS_OK;
}

}  // namespace
}  // namespace citrus

using namespace citrus;

extern "C" int citrus_group_transmission(const float* paths, int M, int64_t nrays, const float* fine_mu,
                                         const int* fine_group, const float* fine_weight, int E, int G,
                                         const float* filter_path, const float* filter_mu, float* out,
                                         void* stream) {
  FineTables F{fine_mu, fine_group, fine_weight, filter_path ? filter_mu : nullptr, M, E, G};
  if (int rc = check_tables(F)) return rc;
  if (!paths || !out) return fail(CITRUS_ERR_ARG, "NULL buffer");
  if ((filter_path == nullptr) != (filter_mu == nullptr)) return fail(CITRUS_ERR_ARG, "filter path without table");
  if (nrays == 0) return CITRUS_OK;
  const size_t lds = sizeof(float) * (size_t)(M + 3) * E;
  const int64_t blocks = (nrays + kThreads - 1) / kThreads;
  if (!grid_fits(blocks, kThreads)) return fail(CITRUS_ERR_ARG, "too many rays for one launch (2^32)");
  hipLaunchKernelGGL(group_transmission_kernel, dim3((unsigned)blocks), dim3(kThreads), lds,
                     reinterpret_cast<hipStream_t>(stream), F, paths, nrays, filter_path, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("group transmission: ") + hipGetErrorString(e));
  return CITRUS_OK;
}

extern "C" int citrus_uncollided_fluence(const float* paths, int M, const double* lattice, const int* lattice_shape,
                                         const float* fine_mu, const int* fine_group, const float* fine_weight, int E,
                                         int G, const double* source, const int* source_ints, const double* off_focal,
                                         const float* profile, const float* thickness, const float* filter_mu,
                                         float* out, void* stream) {
  FineTables F{fine_mu, fine_group, fine_weight, (thickness && filter_mu) ? filter_mu : nullptr, M, E, G};
  if (int rc = check_tables(F)) return rc;
  if (!paths || !lattice || !lattice_shape || !source || !source_ints || !out) return fail(CITRUS_ERR_ARG, "NULL");
  if ((thickness == nullptr) != (filter_mu == nullptr)) return fail(CITRUS_ERR_ARG, "filter thickness without table");
  SourceParams P{};
  for (int k = 0; k < 3; ++k) {
    P.source[k] = source[k];
    P.center[k] = source[3 + k];
    P.uh[k] = source[6 + k];
    P.vh[k] = source[9 + k];
    P.nh[k] = source[12 + k];
  }
  P.sdd = source[15];
  P.drow = source[16];
  P.dch = source[17];
  P.off_row = source[18];
  P.off_ch = source[19];
  P.margin_row = source[20];
  P.margin_ch = source[21];
  P.filter_density = source[22];
  P.curved = source_ints[0];
  P.rows = source_ints[1];
  P.cols = source_ints[2];
  P.n_off = source_ints[3];
  if (P.n_off < 0 || P.n_off > kMaxOffFocal) return fail(CITRUS_ERR_ARG, "off-focal table of 0..16 nodes");
  if (P.n_off > 0 && !off_focal) return fail(CITRUS_ERR_ARG, "off-focal table is NULL");
  for (int k = 0; k < P.n_off; ++k) {
    P.off_dist[k] = off_focal[k];
    P.off_frac[k] = off_focal[P.n_off + k];
  }
  P.profile = profile;
  P.thickness = thickness;
  const int64_t npts = (int64_t)lattice_shape[0] * lattice_shape[1] * lattice_shape[2];
  if (npts == 0) return CITRUS_OK;
  const size_t lds = sizeof(float) * (size_t)(M + 3) * E;
  const int64_t blocks = (npts + kThreads - 1) / kThreads;
  if (!grid_fits(blocks, kThreads)) return fail(CITRUS_ERR_ARG, "too many points for one launch (2^32)");
  hipLaunchKernelGGL(uncollided_kernel, dim3((unsigned)blocks), dim3(kThreads), lds,
                     reinterpret_cast<hipStream_t>(stream), F, P, paths, npts, lattice[0], lattice[1], lattice[2],
                     lattice[3], lattice[4], lattice[5], lattice_shape[1], lattice_shape[2], out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CITRUS_ERR_HIP, std::string("uncollided fluence: ") + hipGetErrorString(e));
  return CITRUS_OK;
}
