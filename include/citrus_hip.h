/*
 * citrus_hip.h — C ABI of the MI355X (gfx950) transport hot path of the CITRUS scatter solver.
 *
 * Every entry point takes raw device pointers, sizes and a hipStream_t (passed as void*), enqueues its kernels on
 * that stream and returns 0 on success or a nonzero CITRUS_ERR_* code; citrus_last_error() returns a message for
 * the last failure on the calling thread. Entry points never allocate device memory except the *_plan_create
 * calls, which upload small index tables once, and never synchronize the stream.
 *
 * Layouts (all float32 unless noted; N = R*C*S cells, flat cell n = (i*C + j)*S + k with i rows (y), j cols (x),
 * k slices (z); NSH = (L+1)^2 spherical-harmonic moments; LD moment order (0, x, y, z)):
 *   mu_t         [G][N]                      total attenuation per group and cell, 1/cm
 *   moments      [V][G][N][NSH][4]           group-major cell moments (scattering-source moments T, fluence phi)
 *   chunk source [V][8*dc][N][G][4]          the reference's cell-major source chunk of 8*dc directions
 *   Y, Yw        [8*D8][NSH]                 real SH (and w_m * Y) at the octant-reordered directions m = o*D8 + d
 *   cdir         [8*D8][3]                   |Omega_a| / Delta_a of the reordered directions, 1/cm
 */
#ifndef CITRUS_HIP_H
#define CITRUS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CITRUS_OK 0
#define CITRUS_ERR_ARG 1
#define CITRUS_ERR_HIP 2
#define CITRUS_ERR_UNSUPPORTED 3

/* Message of the last failed call on the calling thread ("" if none). */
const char* citrus_last_error(void);

/* ABI version of this library; bumped whenever a signature or layout changes. */
int citrus_abi_version(void);

/* ---------------------------------------------------------------------------------------------------------------
 * LD discrete-ordinates sweep.
 *
 * Replaces the reference's fused plane sweep `sweep_planes(index, store, mu_cells, c)`
 * (citrus/scatter_solver/kernels/sweep_kernels.py:265-290), its PyTorch plane step `_plane_step`
 * (citrus/scatter_solver/transport_sweep.py:271-305), the moment product `moments.addmm_(weights, psi_chunk)`
 * (transport_sweep.py:524-528) and the scattering source `Y_selected @ moments_matrix`
 * (citrus/scatter_solver/source_iteration.py:94-103), fused into one pass so the angular fluence never reaches HBM.
 * --------------------------------------------------------------------------------------------------------------- */

/* Opaque per-grid schedule (block-diagonal task tables on the device). */
typedef struct CitrusSweepPlan CitrusSweepPlan;

/* Build the schedule of a grid of shape (R, C, S) swept in cubic blocks of edge `block` (4). Uploads the task
 * tables with hipMalloc/hipMemcpy (the only allocation of the sweep). Counterpart of make_sweep_context's
 * build_sweep_plan / build_flat_index (transport_sweep.py:214-231, sweep_kernels.py:61-78). */
int citrus_sweep_plan_create(int R, int C, int S, int block, CitrusSweepPlan** plan);
void citrus_sweep_plan_destroy(CitrusSweepPlan* plan);

/* Floats of the block-face scratch buffer that citrus_sweep_run needs for G groups, dc directions per octant and
 * V views. */
int64_t citrus_sweep_face_floats(const CitrusSweepPlan* plan, int G, int dc, int V);

typedef struct CitrusSweepArgs {
  int G;            /* energy groups */
  int nsh;          /* SH moments of T and phi: 1, 4, 9, 16, 25, 36, 49 or 64 */
  int D8;           /* directions per octant of the quadrature */
  int d0;           /* first per-octant direction of this chunk */
  int dc;           /* directions per octant in this chunk, 1..32 */
  int V;            /* views in the batch (they share mu_t and the direction tables) */
  int accumulate;   /* 0: phi is overwritten by this chunk; 1: phi += this chunk's moments */
  int write_psi;    /* 1: the angular-fluence moments overwrite `source` in place (needs source) */
  const float* mu_t;     /* [G][N] */
  const float* T;        /* [V][G][N][nsh][4] scattering-source moments, or NULL */
  float* source;         /* [V][8*dc][N][G][4] explicit source chunk, or NULL (in/out when write_psi) */
  float* phi;            /* [V][G][N][nsh][4] fluence moments (output) */
  const float* Y;        /* [8*D8][nsh] */
  const float* Yw;       /* [8*D8][nsh] */
  const float* cdir;     /* [8*D8][3] */
  float* faces;          /* scratch of citrus_sweep_face_floats(plan, G, dc, V) floats */
} CitrusSweepArgs;

/* Sweep one direction chunk of every octant, all groups and views, through the grid with vacuum boundaries, and
 * reduce it to SH moments. The source of direction m in a cell is sum_n Y[m][n] T[n] (if T) plus source[m] (if
 * source). Deterministic: the order of every floating-point reduction is fixed by the schedule. */
int citrus_sweep_run(const CitrusSweepPlan* plan, const CitrusSweepArgs* args, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Scattering-source moments T[go][n][k] = sum_{gi} mu_s_l[gi][go][l(n)] phi[gi][n][k] per cell.
 *
 * Replaces `_scattering_moments_kernel` / `scattering_moments` (citrus/scatter_solver/kernels/
 * scattering_kernels.py:162-252) and `scatter_source_moments` (citrus/scatter_solver/angular_quadrature.py:600-677).
 * --------------------------------------------------------------------------------------------------------------- */

/* Dense transfer volume: transfer [Gi][Go][Lt+1][N] (the reference's mu_s_l_volume layout), phi and T
 * [V][G][N][nsh][4] with Lt >= sqrt(nsh) - 1. */
int citrus_scatter_moments_dense(const float* transfer, int Lt, const float* phi, float* T, int G, int nsh,
                                 int64_t N, int V, void* stream);

/* Factored transfer: mu_s_l(cell) = sum_m rho[m][cell] sigma[m][Gi][Go][Lt+1] (mass transfer moments, cm^2/g,
 * times densities, g/cm^3), formed on the fly; the counterpart of transfer_moments_volume
 * (citrus/scatter_solver/cross_sections.py:886-902) fused into the contraction. rho is [M][N] (shared by the V
 * views). downscatter_only = 1 asserts sigma[m][gi][go] == 0 for gi > go, and the upper triangle is skipped. */
int citrus_scatter_moments_factored(const float* rho, int num_materials, const float* sigma, int Lt,
                                    int downscatter_only, const float* phi, float* T, int G, int nsh, int64_t N,
                                    int V, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CITRUS_HIP_H */
