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
  float* source;         /* explicit source in source_layout, or NULL (in/out when write_psi) */
  float* phi;            /* [V][G][N][nsh][4] fluence moments (output) */
  const float* Y;        /* [8*D8][nsh] */
  const float* Yw;       /* [8*D8][nsh] */
  const float* cdir;     /* [8*D8][3] */
  float* faces;          /* scratch of citrus_sweep_face_floats(plan, G, dc, V) floats */
  int source_layout;     /* 0: cell-major chunk [V][8*dc][N][G][4] (the reference's sweep buffer);
                          * 1: group-major [V][G][N][8*D8][4] of every direction, indexed by the reordered direction
                          *    m = o*D8 + d (not with write_psi);
                          * 2: group-major cell means [V][G][N][8*D8], slopes zero (not with write_psi) */
  const void* T_bf16;    /* in place of T, the mixed-precision operand (P3 and P5) of citrus_bf16_operand or
                          * citrus_scatter_moments_mixed: the moments of degree l >= 2 (n = 4 .. nsh-1) as bfloat16
                          * [V][G][N][4 components][KH], KH = nsh - 4 rounded up to 16, zero padded. The source
                          * expansion then runs as T_low Y (float32 matrix product) + T_bf16 Y (bfloat16 matrix
                          * product, float32 accumulation) */
  const float* T_low;    /* with T_bf16: the moments of degree l <= 1 (n = 0..3) as float32 [V][G][N][4 comp][4] */
} CitrusSweepArgs;

/* Sweep one direction chunk of every octant, all groups and views, through the grid with vacuum boundaries, and
 * reduce it to SH moments. The source of direction m in a cell is sum_n Y[m][n] T[n] (if T) plus source[m] (if
 * source). Deterministic: the order of every floating-point reduction is fixed by the schedule. */
int citrus_sweep_run(const CitrusSweepPlan* plan, const CitrusSweepArgs* args, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Persistent sweep: the same sweep (same tasks, same order of every floating-point reduction, bit-identical moments)
 * in ONE launch of resident waves, for up to CITRUS_PERSISTENT_MAX_VIEWS views whose grids may differ in shape and
 * cell size (views cropped to their beams, coarse and fine stages of a multigrid schedule). Each (view, group) is
 * pinned to one XCD; waves take its block tasks from that XCD's queue and wait only on the tasks they depend on.
 *
 * Replaces the same reference functions as citrus_sweep_run (the per-view sweep of transport_sweep.py:457-537, called
 * once per view by citrus.py:1920-1935); the reference has no multi-grid batching.
 * --------------------------------------------------------------------------------------------------------------- */

#define CITRUS_PERSISTENT_MAX_VIEWS 16
#define CITRUS_PERSISTENT_TIMEOUT 1 /* status: a dependency wait passed its spin limit; the sweep was abandoned */
#define CITRUS_PERSISTENT_BAD_XCC 2 /* status: a wave ran on an XCD id outside 0..7 */

/* Opaque schedule of a list of view grids: per-XCD task queues on the device. */
typedef struct CitrusPersistentPlan CitrusPersistentPlan;

/* One view of a persistent sweep: its grid's buffers (N = R*C*S of the view's grid). */
typedef struct CitrusSweepView {
  const float* mu_t;         /* [G][N] */
  const float* T;            /* [G][N][nsh][4] scattering-source moments, or NULL with T_bf16 */
  const void* T_bf16;        /* mixed operand: bfloat16 degrees l >= 2 [G][N][4][KH] (P3, P5), or NULL */
  const float* T_low;        /* with T_bf16: float32 degrees l <= 1 [G][N][4][4] */
  float* source;             /* group-major source of every direction ([G][N][8*D8][4] or [G][N][8*D8]), source_layout
                              * 3: the block-order cell means [G][blocks*8*64*dp + pad] (CITRUS_FC_BLOCK_ORDER), or
                              * NULL */
  float* phi;                /* [G][N][nsh][4] fluence moments (output) */
  const float* cdir;         /* [8*D8][3] |Omega_a| / Delta_a on this view's grid */
  const float* fc_scale;     /* source_layout 3: [G][blocks*64 + pad] scale of the first-collision sampling correction */
  const float* fc_beta;      /* source_layout 3: [G][blocks*64 + pad] beta of the correction */
  const float* fc_u;         /* source_layout 3: [blocks][64][4] Omega_UC of each cell (citrus_fc_primary_directions
                              * in block order) */
} CitrusSweepView;

typedef struct CitrusPersistentArgs {
  int G, nsh, D8, d0, dc;    /* as in CitrusSweepArgs */
  int accumulate;            /* 0: phi is overwritten; 1: phi += this chunk's moments */
  int source_layout;         /* -1: no explicit source; 1 or 2 as in CitrusSweepArgs (the same for every view); 3: the
                              * cell means of layout 2 before the first-collision sampling correction, in the block
                              * order of CITRUS_FC_BLOCK_ORDER (citrus_fc_sums_blocked), which the sweep applies as it
                              * reads them: mean scale (1 + beta eta), eta = u[cell] . omega[m], the operations of
                              * citrus_fc_correct, so the moments equal those of layout 2 after
                              * citrus_fc_correct_means bit for bit (needs fc_scale, fc_beta, fc_u and fc_omega) */
  const float* Y;            /* [8*D8][nsh], shared by the views */
  const float* Yw;           /* [8*D8][nsh] */
  const CitrusSweepView* views; /* HOST array of the plan's nviews views, read before the launch */
  int* status;               /* DEVICE int or NULL: a sweep that aborts ORs its CITRUS_PERSISTENT_* code into it. The
                              * word is never cleared by the library, so one word can collect the status of every
                              * chunk and sweep until the caller reads and clears it (the workspace's own status word
                              * is cleared by each chunk's launch) */
  int spin_limit;            /* sleeps of one dependency wait before the sweep aborts; 0: the default (about a
                              * second). A small value forces the abort path from the host, for tests */
  const float* fc_omega;     /* source_layout 3: [8*D8][4] the sweep directions (x, y, z, 0) in float32 */
  int waves_per_cu;          /* resident single-wave workgroups per CU: 0 as many as fit (the kernel's occupancy),
                              * else at most that many (a launch shape to tune: fewer waves contend less for L2 and
                              * HBM; any count >= 1 per XCD completes, since every task's dependencies precede it
                              * in its own queue) */
} CitrusPersistentArgs;

/* Schedule of nviews view grids, shapes[v] = (R, C, S), with G groups. Uploads the task queues (hipMalloc). */
int citrus_persistent_plan_create(int nviews, const int* shapes, int G, CitrusPersistentPlan** plan);
void citrus_persistent_plan_destroy(CitrusPersistentPlan* plan);

/* Bytes of the device workspace (done flags, tickets, status, inter-block faces) of a sweep with dc directions per
 * octant; the workspace must be 256-byte aligned. */
int64_t citrus_persistent_workspace_bytes(const CitrusPersistentPlan* plan, int dc);

/* Block tasks of one chunk sweep (8 octants x groups x blocks of every view). */
int64_t citrus_persistent_num_tasks(const CitrusPersistentPlan* plan);

/* Host-only view of the schedule a plan of these shapes would run (no device needed): returns the task count n and,
 * where the pointers are non-NULL, the tasks in queue order (tasks[n][4] = {view | group << 8 | octant << 16, block
 * row, block col, block slice}), the queue offsets qoff[9] (queue q = XCD q holds tasks [qoff[q], qoff[q+1])) and,
 * per task, the indices of the tasks it waits for (deps[n][7], -1 padded). Returns -1 on bad arguments. */
int64_t citrus_persistent_schedule_table(int nviews, const int* shapes, int G, int* tasks, int* qoff, int* deps);

/* Sweep one direction chunk of every view in one persistent launch (preceded by a small kernel that clears the
 * workspace's flags; no memset, so the call can be captured in a graph). */
int citrus_persistent_sweep_run(const CitrusPersistentPlan* plan, const CitrusPersistentArgs* args, void* workspace,
                                int64_t workspace_bytes, void* stream);

/* After a sweep (synchronizes the stream): *status is 0 or a CITRUS_PERSISTENT_* code, *unfinished the number of
 * block tasks that did not complete (0 when the sweep ran through). */
int citrus_persistent_check(const CitrusPersistentPlan* plan, const void* workspace, int* status,
                            int64_t* unfinished, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Spherical-harmonic expansion at a set of directions: C[k][j] = sum_n Y[k][n] B[n][j] (float32 on the matrix
 * cores, four moments per step), the multiple-collision source of the collapsed-cone image at a chunk of cone
 * directions from the moment-major cell means (replaces `torch.mm(harmonics, moments_by_cell)`,
 * citrus/scatter_solver/scatter_image.py:1225,1246).
 * --------------------------------------------------------------------------------------------------------------- */

/* Y [K][nsh], B [nsh][ncol], C [K][ncol]; nsh = (L+1)^2 with L <= 7; the directions' harmonics, K x nsh padded to
 * 16-direction tiles, must fit 64 KB. */
int citrus_sh_expand(const float* Y, int K, int nsh, const float* B, int64_t ncol, float* C, void* stream);

/* The multiple-collision part of the signal-only image's emission at one point per cell, expanded, attenuated and
 * summed over the groups in one pass (the multiple-collision term of `attenuated_emission` weighted by the signal
 * response, uncollided_kernels.py:1138, scatter_image.py:1246-1262): with S_g = sum_l Y[k][l] means[l][g][n] and
 * tau_g the cell's group optical depth (as in citrus_attenuated_emission: tau [K][G][N] with tau_mix NULL, or
 * [K][tau_channels][N] with tau_mix [G][tau_channels], tau_channels 1 or 2, +inf channel 0 = no path),
 *   out[k][0][n] = sum_g signal_weight[k][g] exp(-tau_g) S_g,   out[k][1][n] = sum_g exp(-tau_g) S_g.
 * Y [K][nsh], means [nsh][G][N], signal_weight [K][G], out [K][2][N]; G <= 32 and (3 G + 1) N x 4 bytes < 4 GiB. */
int citrus_multiple_emission(const float* Y, int K, int nsh, const float* means, int G, int64_t N, const float* tau,
                             int tau_channels, const float* tau_mix, const float* signal_weight, float* out,
                             void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Density model: the reference materials' densities that reproduce an attenuation volume X (1/cm at the model's
 * energy), interpolating between the two materials whose nominal attenuations bracket max(X, 0); below the first or
 * above the last, that material alone with density X / m. Replaces attenuation_to_densities
 * (citrus/utilities.py:710-753) as DensityModel applies it (citrus/citrus.py:466-509).
 * --------------------------------------------------------------------------------------------------------------- */

/* X [n] -> rho [M][n] (library material order). material, mu_nominal, nominal_density and mass_attenuation are HOST
 * arrays of M <= 8 slots in bracketing order (mu_nominal strictly increasing, mu_nominal = nominal_density *
 * mass_attenuation as float32): each slot's library material index and its parameters. */
int citrus_attenuation_to_densities(const float* X, int64_t n, int M, const int* material, const float* mu_nominal,
                                    const float* nominal_density, const float* mass_attenuation, float* rho,
                                    void* stream);

/* Conservative resampling of a voxel-wise constant field between two grids: each target cell gets the exact average
 * of the field over its volume (zero outside the source box), so every channel's integral is conserved. Replaces
 * resample_conservative (citrus/scatter_solver/geometry.py:304-327) with its per-axis overlap matrices. field
 * [channels][R][C][S] -> out [channels][Rt][Ct][St]; per axis (rows, cols, slices) the DEVICE arrays first [n_t],
 * count [n_t] (int32) and share [n_t][width] (float32) give, for target cell j, the source cells first[j] ..
 * first[j] + count[j] - 1 it overlaps and the share of j each one covers (the nonzero band of the overlap matrix). */
int citrus_resample_conservative(const float* field, int channels, int R, int C, int S, int Rt, int Ct, int St,
                                 const int* row_first, const int* row_count, const float* row_share, int row_width,
                                 const int* col_first, const int* col_count, const float* col_share, int col_width,
                                 const int* slice_first, const int* slice_count, const float* slice_share,
                                 int slice_width, float* out, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Compact per-view accumulators: the LD cell mean of every (view, group, moment) as float16 scaled by the
 * (group, moment)'s maximum magnitude over 2^14 (scale 1 where that maximum is 0).
 *
 * Replaces StateStore.put / StateStore.get (citrus/citrus.py:686-737), batched over views and reading the sweep's
 * group-major moments in place.
 * --------------------------------------------------------------------------------------------------------------- */

/* Floats of the workspace citrus_state_pack needs (per-block maxima). */
int64_t citrus_state_pack_workspace_floats(int V, int G, int nsh, int64_t N);

/* phi [V][G][N][nsh][4] (component 0 = cell mean) -> scale [V][G][nsh] float32 and q [V][G][nsh][N] float16. */
int citrus_state_pack(const float* phi, int V, int G, int nsh, int64_t N, float* scale, void* q, float* workspace,
                      int64_t workspace_floats, void* stream);

/* Size of the device scratch array of citrus_state_pack_static, in floats (32 MiB per GPU, held until the process
 * ends). */
#define CITRUS_STATE_PACK_SCRATCH_FLOATS (1 << 23)

/* citrus_state_pack without a workspace argument and without allocating: the per-block maxima go to a file-scope
 * device array of CITRUS_STATE_PACK_SCRATCH_FLOATS floats, one per GPU, whose address is looked up on the first call
 * on each device. No counterpart in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). Same inputs, outputs and results (bit for
 * bit) as citrus_state_pack; returns CITRUS_ERR_ARG when citrus_state_pack_workspace_floats(V, G, nsh, N) exceeds
 * the array. Every call on one GPU shares the array: calls that overlap on one GPU (two streams, or two threads)
 * race and must be serialized by the caller. */
int citrus_state_pack_static(const float* phi, int V, int G, int nsh, int64_t N, float* scale, void* q,
                             void* stream);

/* q [V][G][nsh][N] float16 and scale [V][G][nsh] -> phi [V][G][N][nsh][4] with phi[..][0] = q * scale and zero
 * LD slopes. */
int citrus_state_unpack(const void* q, const float* scale, int V, int G, int nsh, int64_t N, float* phi,
                        void* stream);

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
 * views). downscatter_only = 1 asserts sigma[m][gi][go] == 0 for gi > go, and the upper triangle is skipped.
 * cell_means (or NULL) receives the cell means of T again, moment-major [V][nsh][G][N]: the operand of the
 * collapsed-cone image's multiple-collision expansion (the moments_by_cell operand, scatter_image.py:1225,1246),
 * with no separate transposition. */
int citrus_scatter_moments_factored(const float* rho, int num_materials, const float* sigma, int Lt,
                                    int downscatter_only, const float* phi, float* T, int G, int nsh, int64_t N,
                                    int V, float* cell_means, void* stream);

/* citrus_scatter_moments_factored of stored accumulators (the state store's format, citrus/citrus.py:623-737): q
 * [V][G][nsh][N] float16 and scale [V][G][nsh], the cell means q scale with zero LD slopes decoded as
 * citrus_state_unpack decodes them, so T [V][G][N][nsh][4] equals the factored contraction of the unpacked
 * accumulators bit for bit, without the float32 accumulators being written and read. */
int citrus_scatter_moments_stored(const float* rho, int num_materials, const float* sigma, int Lt,
                                  int downscatter_only, const void* q, const float* scale, float* T, int G, int nsh,
                                  int64_t N, int V, void* stream);


/* ---------------------------------------------------------------------------------------------------------------
 * Ray traversal (Siddon / Amanatides-Woo): exact line integrals of a channel-last field [R][C][S][nch] (nch <= 16).
 *
 * box = {lo_x, lo_y, lo_z, dx, dy, dz} (cm, float64) of the traced grid; axes x = cols, y = rows, z = slices.
 * box, panel, panel_ints, point and lattice_* are HOST parameter blocks (read before the launch); volume, field,
 * p0, p1, omega, out and the seq_* arrays are DEVICE buffers.
 * panel = {source[3], panel_center[3], u_hat[3], v_hat[3], n_hat[3], SDD, row pitch, channel pitch, row offset,
 * channel offset} (float64, recon frame) and panel_ints = {curved, rows, channels}.
 * Replaces the reference's `_ray_kernel` MODE 0-3 and `_downstream_shared_kernel`
 * (citrus/scatter_solver/kernels/ray_tracing_kernels.py:330-639) and `line_integrals` / `_trace_chunk`
 * (citrus/scatter_solver/ray_tracing.py:141-284), behind the `Projectors` protocol (ray_tracing.py:27-115).
 * --------------------------------------------------------------------------------------------------------------- */

/* Integrals along the segments p0[i] -> p1[i] ([n][3] float64; p0_shared: one start point for every ray) into
 * out[c][i]. Counterpart of line_integrals (ray_tracing.py:205) and SiddonProjectors.upstream_to_points (:426). */
int citrus_ray_segments(const float* volume, int R, int C, int S, int nch, const double* box, const double* p0,
                        const double* p1, int p0_shared, int64_t n, float* out, void* stream);

/* Voxel sequence of every segment: flat voxel indices with nonzero chord, in traversal order, and their parameter
 * lengths (up to max_len per ray; seq_count holds the full count). Test and diagnostic entry point. */
int citrus_ray_trace_indices(int R, int C, int S, const double* box, const double* p0, const double* p1, int64_t n,
                             int max_len, int32_t* seq_idx, float* seq_len, int32_t* seq_count, void* stream);

/* Integrals from `point` to every point origin + (j dx, i dy, k dz) of a lattice of shape (rows, cols, slices) into
 * out[c][(i*cols + j)*slices + k]: SiddonProjectors.upstream_cumulative (ray_tracing.py:405) when the lattice is the
 * voxel centres. */
int citrus_ray_lattice(const float* volume, int R, int C, int S, int nch, const double* box, const double* point,
                       const double* lattice_origin, const double* lattice_delta, const int* lattice_shape,
                       float* out, void* stream);

/* Focal spot to every pixel centre, out[c][row][col]; samples > 1 traces the samples^2 midpoint sub-rays of each
 * pixel and returns -log(mean(exp(-tau))). SiddonProjectors.primary_project (ray_tracing.py:447). */
int citrus_ray_primary(const float* volume, int R, int C, int S, int nch, const double* box, const double* panel,
                       const int* panel_ints, int samples, float* out, void* stream);

/* Line through every pixel centre along omega[k] ([K][3]), clipped to the box, of the field of direction k
 * (field + k * field_k_stride; stride 0 shares one field) into out[k * out_k_stride + c * out_c_stride + pixel].
 * SiddonProjectors.oblique_project (ray_tracing.py:548). */
int citrus_ray_oblique(const float* field, int64_t field_k_stride, int R, int C, int S, int nch, const double* box,
                       const double* panel, const int* panel_ints, const double* omega, int K, float* out,
                       int64_t out_k_stride, int64_t out_c_stride, void* stream);

/* From every voxel centre to the panel along omega[k] into out[k][c][voxel]; +inf where omega . (-n_hat) <= 0.
 * SiddonProjectors.downstream_cumulative (ray_tracing.py:501). */
int citrus_ray_downstream(const float* volume, int R, int C, int S, int nch, const double* box, const double* panel,
                          const int* panel_ints, const double* omega, int K, float* out, void* stream);

/* Crossing sequences of K directions ([K][3] float64, device) through the lattice of the grid's voxels, one record of
 * citrus_ray_crossing_sequence_bytes(R, C, S) bytes per direction into `sequences` (device, 16-byte aligned). Rays
 * from voxel centres along one direction are whole-voxel translates of each other, so one float64 traversal per
 * direction (the plane-stepping and tie rules of the other traversals) serves every voxel. Only the voxel sizes of
 * `box` enter. At most 511 voxels per axis. The records depend on the directions and voxel sizes alone: compute them
 * once per direction set and reuse them for every field traced along those directions. */
int64_t citrus_ray_crossing_sequence_bytes(int R, int C, int S);
int citrus_ray_crossing_sequences(int R, int C, int S, const double* box, const double* omega, int K, void* sequences,
                                  void* stream);

/* citrus_ray_downstream from the crossing-sequence records of the same K directions (citrus_ray_crossing_sequences):
 * every voxel sums its field over the samples of the record that stay inside the grid, cut at its own panel distance
 * (Kahan-compensated float32). With mix ([nout][nch] device, nout <= 64) the outputs are out[k][g][voxel] =
 * sum_c mix[g][c] * integral of channel c, e.g. the optical depth of every group from the material densities and
 * mu_t per unit density; without mix, nout == nch and out is as in citrus_ray_downstream. At most 65535 directions
 * per launch. Same contract as SiddonProjectors.downstream_cumulative (ray_tracing.py:501, shared-sequence kernel
 * ray_tracing_kernels.py:518-639) up to float32 rounding. */
int citrus_ray_downstream_shared(const float* volume, int R, int C, int S, int nch, const double* box,
                                 const double* panel, const int* panel_ints, const double* omega, int K,
                                 const void* sequences, const float* mix, int nout, float* out, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Group transmission and uncollided fluence.
 * --------------------------------------------------------------------------------------------------------------- */

/* T[g][r] = sum_{e: fine_group[e] = g} fine_weight[e] exp(-sum_m fine_mu[m][e] paths[m][r] - filter_mu[e]
 * filter_path[r]) for M <= 8 path channels, G <= 32 groups (filter_path and filter_mu both NULL or both given).
 * Replaces `_group_transmission_kernel` / `group_transmission` (citrus/scatter_solver/kernels/
 * transmission_kernels.py:190-275, citrus/scatter_solver/cross_sections.py:797-878). All pointers are device. */
int citrus_group_transmission(const float* paths, int M, int64_t nrays, const float* fine_mu, const int* fine_group,
                              const float* fine_weight, int E, int G, const float* filter_path,
                              const float* filter_mu, float* out, void* stream);

/* Uncollided fluence phi[g][point] = b(Omega) / |r - r0|^2 T_g(paths[:, point], filter path) at the lattice points
 * r = origin + (j dx, i dy, k dz), lattice = {ox, oy, oz, dx, dy, dz} and lattice_shape = {rows, cols, slices}
 * (host). source = the 20 values of the ray-traversal panel block, then {field margin rows, field margin channels,
 * filter density} (host, 23 doubles); source_ints = {curved, rows, channels, off-focal nodes K <= 16} (host);
 * off_focal = {K distances, K fractions} (host, or NULL for K = 0); profile and thickness are device [rows][channels]
 * images or NULL. Counterpart of uncollided_from_path_lengths (citrus/scatter_solver/uncollided.py:160-197) with
 * SourceModel.angular_weight and filter_path (citrus/scatter_solver/source.py:289-446). */
int citrus_uncollided_fluence(const float* paths, int M, const double* lattice, const int* lattice_shape,
                              const float* fine_mu, const int* fine_group, const float* fine_weight, int E, int G,
                              const double* source, const int* source_ints, const double* off_focal,
                              const float* profile, const float* thickness, const float* filter_mu, float* out,
                              void* stream);

/* ---- first-collision source (citrus/scatter_solver/uncollided.py:200-656; uncollided_kernels.py:404-700) ---------
 * Shared inputs of the three entry points below. Points form the lattice r = lattice[0:3] + (c dx, r dy, s dz) of
 * the R x C x S grid (flat (r C + c) S + s). Each point carries up to 4 material slots (index into the library,
 * density in g/cm^3; unused slots density 0), the counterpart of the reference's MaterialPairs
 * (cross_sections.py:656). The transfer table is node-major [n_eta][M][G][GP] with GP = 4 for G <= 4, 8 for G <= 8,
 * 16 for G <= 16 and 32 otherwise (padding entries 0), and the entries g_out < g_in zeroed (downscatter only,
 * uncollided.py:258); eta_map, when given, is the adaptive node map of the table (cross_sections.py:905-945).
 * Pointers are device unless marked host. */
typedef struct CitrusFcArgs {
  int M, G, n_eta, n_map;       /* materials, groups, table nodes, map knots (ignored without eta_map) */
  const float* table;           /* [n_eta][M][G][GP] per-steradian transfer in cm^2/g/sr */
  const float* eta_map;         /* [n_map] or NULL (uniform table) */
  int n_slots;                  /* material slots per point, 1..4 */
  const int* slot_material;     /* [n_slots][P] */
  const float* slot_density;    /* [n_slots][P] */
  const float* phi_uc;          /* [G][P] uncollided fluence */
  int R, C, S;                  /* point grid */
  double lattice[6];            /* host values: first point centre (x, y, z), pitches (dx, dy, dz) in cm */
  double focal[3];              /* host values: focal spot, or the beam direction when parallel != 0 */
  int parallel;
  const float* omegas;          /* [K][3] unit directions */
  int K;
} CitrusFcArgs;

/* S_FC at every point and direction: out [K][G][P] (first_collision_source_at_directions, uncollided.py:200). */
int citrus_fc_source(const CitrusFcArgs* args, float* out, void* stream);

/* Sampling-correction inputs per point: out [5][G][P] = quadrature sums I0, I1, I2 (sum_k w_k eta_k^j S_k, float32
 * in direction order) and the exact moments E0, E1 = sum_{g'<=g} sum_m rho_m mu_s_l^{g'->g}[l] phi_UC,g'
 * (first_collision_quadrature_sums and first_collision_exact_moments, uncollided.py:357-384, :538-608; the fused
 * first_collision_sums of uncollided_kernels.py:637). weights [K]; mu_s_l [M][G][G][lp1], lp1 >= 2. samples (or
 * NULL): also stores the uncorrected S_FC of every point and direction, group-major [G][P][K]. */
int citrus_fc_sums(const CitrusFcArgs* args, const float* weights, const float* mu_s_l, int lp1, float* out,
                   float* samples, void* stream);

/* CITRUS_FC_BLOCK_ORDER: the first-collision cell means as one persistent sweep task reads them. Per group
 * [blocks][8 octants][64 cells][dp], dp = D8 rounded up to 4 (the padding zero), then CITRUS_FC_GROUP_PAD unused
 * floats; blocks of 4^3 cells over (ceil(R/4), ceil(C/4), ceil(S/4)) row-major; in octant o's run the cells come in
 * the task order of that octant: sorted by diagonal plane a + b + c, then a, then b, where (a, b, c) are the cell's
 * (column, row, slice) within the block counted from the octant's upstream corner (o & 4: columns reversed, o & 2:
 * rows, o & 1: slices), cells past the grid unused. Per-cell quantities per group (scale, beta) [blocks][64] then the
 * pad, per view (Omega_UC) [blocks][64][4], a block's cells (row, column, slice) row-major. The pad skews the regions
 * of consecutive groups and views across HBM channels. */
#define CITRUS_FC_BLOCK_ORDER 2
#define CITRUS_FC_GROUP_PAD 1088

/* citrus_fc_sums with the samples in block order: samples [G][blocks*8*64*dp + pad] of the K = 8 D8 octant-major sweep
 * directions (K <= 256). The sums are those of citrus_fc_sums bit for bit. scale, beta [G][blocks*64 + pad] and
 * u [blocks*64][4] (all three, or all NULL): the sampling-correction factors of the sums (the float64 formula of
 * first_collision_correction_factors, uncollided.py:611-640, rounded to float32) and each cell's primary direction
 * (as citrus_fc_primary_directions), in the per-cell block order of CITRUS_FC_BLOCK_ORDER, cells past the grid zero.
 * No counterpart in the reference: the layout in which the persistent sweep reads the deferred first-collision
 * source. */
int citrus_fc_sums_blocked(const CitrusFcArgs* args, const float* weights, const float* mu_s_l, int lp1, float* out,
                           float* samples, int D8, float* scale, float* beta, float* u, void* stream);

/* In place, samples [G][P][K] (from citrus_fc_sums) *= scale (1 + beta eta) with scale, beta [G][P] and eta the
 * cosine between Omega_UC of the point and omegas[k]. At one point per cell (factor 1) the result equals
 * citrus_fc_moments layout 3 bit for bit, without a second evaluation of the source (uncollided.py:420-433). */
int citrus_fc_correct(const CitrusFcArgs* args, const float* scale, const float* beta, float* samples, void* stream);

/* out [P][4]: Omega_UC of every point as the first-collision kernels form it (FP64, rounded to float32), w = 0. With
 * citrus_fc_correct_means or the persistent sweep's source layout 3 it applies the sampling correction away from the
 * kernel that stored the samples (the split of uncollided.py:420-433 between the sums and the sweep). */
int citrus_fc_primary_directions(const CitrusFcArgs* args, float* out, void* stream);

/* citrus_fc_correct at one point per cell for V views at once, the directions given as tables: samples [V][G][N][M]
 * (cell means, in place) *= scale (1 + beta eta), scale and beta [V][G][N], eta from u4 [V][N][4]
 * (citrus_fc_primary_directions) and omega4 [M][4]. The same operations as citrus_fc_correct, so the values are equal
 * bit for bit (uncollided.py:420-433). */
int citrus_fc_correct_means(float* samples, const float* scale, const float* beta, const float* u4,
                            const float* omega4, int V, int G, int64_t N, int M, void* stream);

/* LD moments of the corrected first-collision source on the cells of the factor-coarsened point grid
 * (first_collision_moments_at_directions, uncollided.py:464-535): samples S scale (1 + beta eta) (scale, beta
 * [G][P], or both NULL for no correction) at the f^3 sub-points of each cell, restricted to [mean, x, y, z] slopes
 * (restrict_to_linear_discontinuous_moments, uncollided.py:311). layout 0: out [K][G][4][cells]; 1: cell-major
 * [K][cells][G][4] (the sweep's source chunk); 2: group-major [G][cells][K][4]; 3 (factor 1 only, where the slopes
 * are zero): group-major cell means [G][cells][K]. */
int citrus_fc_moments(const CitrusFcArgs* args, const float* scale, const float* beta, int factor, int layout,
                      float* out, void* stream);

/* ---- collapsed-cone scatter image (citrus/scatter_solver/scatter_image.py:829-1323) ----------------------------
 * Attenuated total emission at the quadrature points of one direction chunk, the fused form of
 * first_collision_source_at_directions + the cell's multiple-collision source + supervoxel_subpoint_emission
 * (uncollided_kernels.py:713, attenuated_emission at :1138): e_g = (S_FC,g + S_MC,g) exp(-tau_j,g) with
 * tau_j = max(0, tau_cell - mu_cell w . (r_j - r_cell)). args: the first-collision inputs on the quadrature grid,
 * omegas = the chunk's K directions; multiple [K][G][cells] of the factor-coarsened grid; tau [K][G][cells] with
 * tau_mix NULL, or [K][tau_channels][cells] (tau_channels <= 8) with tau_mix [G][tau_channels], the cell's optical
 * depth of group g then being sum_c tau_mix[g][c] tau[k][c][cell] (+inf where tau[k][0][cell] is +inf); mu_t_coarse
 * [G][cells]; heights [P] panel heights of the points (or NULL: 0); signal_weight [K][G] (or NULL). out [K][G][P], or
 * [K][P] = sum_g signal_weight e_g with signal_weight. partial [K][ceil(P / CITRUS_EMISSION_BLOCK)][2]: per block of
 * CITRUS_EMISSION_BLOCK consecutive points the sums of sum_g e_g and h sum_g e_g (unweighted), which the caller
 * reduces. */
#define CITRUS_EMISSION_BLOCK 32
int citrus_attenuated_emission(const CitrusFcArgs* args, const float* tau, int tau_channels, const float* tau_mix,
                               const float* mu_t_coarse, const float* multiple, int factor, const float* heights,
                               const float* signal_weight, float* out, float* partial, void* stream);

/* citrus_attenuated_emission at one point per cell (factor 1) with signal weights, the multiple-collision part
 * taken from citrus_multiple_emission: collapsed [K][2][cells]; out [K][P] = sum_g signal_weight e_g as above and the
 * same partial sums. */
int citrus_attenuated_emission_collapsed(const CitrusFcArgs* args, const float* tau, int tau_channels,
                                         const float* tau_mix, const float* collapsed, const float* heights,
                                         const float* signal_weight, float* out, float* partial, void* stream);

/* Footprint box-filter matrices out [k][N][N], B[c][j] = |[j, j+1] n W_c| / |W_c|, W_c the window of width
 * widths[m] (pixels) centred on pixel c clipped to [0, N] (box_matrix, scatter_image.py:952). */
int citrus_box_matrix(const float* widths, int k, int N, float* out, void* stream);

/* The finite-cone footprint filters of one direction chunk straight from citrus_attenuated_emission's partial sums
 * [K][nblocks][2] (_Footprint of scatter_image.py:1033-1060 with box_matrix, :952): heights[k] = sum e h / sum e
 * (center_height where sum e is 0), widths (heights[k] cell_width[k][a]) scale_a pixels (a = 0 channels, 1 rows),
 * channel_filter_t [K][Nc][Nc] the transposed channel matrices B^T, each column c scaled by channel_weight[k][c]
 * (NULL: 1), row_filter [K][Nr][Nr] the row matrices B. heights [K] receives the mean heights. */
int citrus_footprint_filters(const float* partial, int K, int nblocks, float center_height, const float* cell_width,
                             float scale_channel, float scale_row, int Nc, int Nr, const float* channel_weight,
                             float* heights, float* channel_filter_t, float* row_filter, void* stream);

/* The signal-only image's footprint blur of one direction chunk and its sum over the chunk's directions, banded
 * (the box filters of citrus_footprint_filters applied window by window instead of as dense matrices:
 * box_blur_rows_cols and the direction sum of scatter_image.py:897,1033-1060,1200-1230): heights[k] from the
 * emission kernel's partial sums as in citrus_footprint_filters; work [K][Nr][Nc] = the channel blur of image
 * [K][Nr][Nc] (times channel_weight [K][Nc] when given); acc [Nr][Nc] += sum_k the row blur of work[k]. */
int citrus_footprint_blur(const float* partial, int K, int nblocks, float center_height, const float* cell_width,
                          float scale_channel, float scale_row, int Nc, int Nr, const float* channel_weight,
                          const float* image, float* heights, float* work, float* acc, void* stream);

/* Mixed-precision operand of the scattering-source contraction (the mixed-precision form of the source expansion of
 * transport_sweep.py:524-528 / source_iteration.py:92-102): T [V][G][N][nsh][4] float32 -> T_low [V][G][N][4][4]
 * float32 (moments n = 0..3, degrees l <= 1, component-major) and T16 [V][G][N][4][KH] bfloat16 (moments n >= 4,
 * round to nearest even, KH = nsh - 4 rounded up to 16, zero padded). nsh 16 or 36. */
int citrus_bf16_operand(const float* T, int G, int nsh, int64_t N, int V, float* T_low, uint16_t* T16, void* stream);

/* citrus_scatter_moments_factored that writes its result directly as the mixed-precision operand (T_low, T_bf16 as
 * above) instead of float32 T, with the same optional cell_means: the sweep of the next step reads it in place. */
int citrus_scatter_moments_mixed(const float* rho, int num_materials, const float* sigma, int Lt, int downscatter_only,
                                 const float* phi, float* T_low, void* T_bf16, int G, int nsh, int64_t N, int V,
                                 float* cell_means, void* stream);

/* ---- Phi_{>2} surrogate network ------------------------------------------------------------------------------
 * No counterpart in the reference: the learned part of this path, a 3D U-Net that predicts the collided moments of
 * collision order >= 3 from the densities and the second-order moments. Activations are float32 [B][C][D][H][W]
 * with the channels of a sample contiguous; *_bstride is the element stride between samples (a tensor may be a
 * channel slice of a larger one). Weights are bfloat16 (uint16_t bit patterns) [Co][Ci][k][k][k], k = 3 or 1, zero
 * padding (k-1)/2, stride 1; every sum accumulates in float32. The leaky ReLU has slope 0.01 below zero. */

/* y = conv(x, w) + bias (bias may be NULL), then the leaky ReLU when leaky != 0. */
int citrus_nn_conv3d_forward(const float* x, int64_t x_bstride, int B, int Ci, int D, int H, int W, const uint16_t* w,
                             const float* bias, int Co, int k, int leaky, float* y, int64_t y_bstride, void* stream);

/* dx = gradient of sum(g * conv(x, w)) with respect to x: g [B][Co] -> dx [B][Ci] (overwritten). */
int citrus_nn_conv3d_backward_data(const float* g, int64_t g_bstride, int B, int Co, int D, int H, int W,
                                   const uint16_t* w, int Ci, int k, float* dx, int64_t dx_bstride, void* stream);

/* dw [Co][Ci][k^3] and db [Co] (db may be NULL) of sum(g * conv(x, w)), summed over the samples in a fixed order;
 * workspace holds citrus_nn_wgrad_workspace_floats(B, Ci, Co, D, H, W, k) floats. */
int64_t citrus_nn_wgrad_workspace_floats(int B, int Ci, int Co, int D, int H, int W, int k);
int citrus_nn_conv3d_backward_weight(const float* x, int64_t x_bstride, const float* g, int64_t g_bstride, int B,
                                     int Ci, int Co, int D, int H, int W, int k, float* dw, float* db,
                                     float* workspace, void* stream);

/* g *= leaky'(y) elementwise over n_per_sample contiguous elements of each sample (y is the layer's output). */
int citrus_nn_leaky_backward(float* g, int64_t g_bstride, const float* y, int64_t y_bstride, int64_t n_per_sample,
                             int B, void* stream);

/* 2x2x2 average pooling of x [B][C][D][H][W] (even sizes) into y; its adjoint adds (accumulate) or writes
 * g / 8 into every voxel of the block. */
int citrus_nn_avg_pool2(const float* x, int64_t x_bstride, int B, int C, int D, int H, int W, float* y,
                        int64_t y_bstride, void* stream);
int citrus_nn_avg_pool2_backward(const float* g, int64_t g_bstride, int B, int C, int D, int H, int W, float* dx,
                                 int64_t dx_bstride, int accumulate, void* stream);

/* Nearest 2x upsampling of x [B][C][D][H][W] into y [B][C][2D][2H][2W]; its adjoint sums each 2x2x2 block. */
int citrus_nn_upsample2(const float* x, int64_t x_bstride, int B, int C, int D, int H, int W, float* y,
                        int64_t y_bstride, void* stream);
int citrus_nn_upsample2_backward(const float* g, int64_t g_bstride, int B, int C, int D, int H, int W, float* dx,
                                 int64_t dx_bstride, void* stream);

/* One bias-corrected Adam step on float32 master weights p (m, v the moment estimates, step >= 1); p_bf16 (or NULL)
 * receives p rounded to bfloat16 (nearest, ties to even). */
int citrus_nn_adam(float* p, const float* g, float* m, float* v, uint16_t* p_bf16, int64_t n, float lr, float beta1,
                   float beta2, float eps, int step, void* stream);

/* y = x rounded to bfloat16 (nearest, ties to even). */
int citrus_nn_to_bf16(const float* x, uint16_t* y, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CITRUS_HIP_H */
