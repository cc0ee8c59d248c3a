/*
 * ouranos_amd.h -- C ABI of the MI355X DG-Sn transport-sweep engine.
 *
 * The library replaces the CPU sweep behind the reference's
 * RadiativeTransfer<dim,tensor_dim> Epetra_Operator (Rombur/ouranos,
 * src/RadiativeTransfer.hh:46-127): the operator y = x - D2M L^-1 M2D S(x)
 * that AztecOO GMRES/BiCGStab and source iteration call.  Plain pointers and
 * sizes only; no C++ types cross this boundary.  Every entry point returns
 * OU_OK (0) or a negative ou_status; ou_last_error() gives the message of the
 * last failure on the calling thread.
 *
 * Vector layout (all FP64):  x[m * n_dofs + cell * N + l]
 *   cell = i + nx*(j + ny*k),  l = lx + (p+1)*(ly + (p+1)*lz),  N = (p+1)^dim
 */
#ifndef OURANOS_AMD_H
#define OURANOS_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  OU_OK = 0,
  OU_ERR_ARG = -1,         /* invalid argument or inconsistent descriptor */
  OU_ERR_SIZE = -2,        /* vector / plan size mismatch */
  OU_ERR_HIP = -3,         /* HIP runtime error */
  OU_ERR_UNSUPPORTED = -4, /* valid request the library does not implement */
  OU_ERR_NONFINITE = -5,   /* NaN/Inf detected in an input */
  OU_ERR_NOT_CONVERGED = -6,
  OU_ERR_NOMEM = -7,
  OU_ERR_TIMEOUT = -8
} ou_status;

typedef enum { OU_FP64 = 0, OU_FP32 = 1 } ou_precision;
/* How a sweep is launched (ou_set_sweep_mode). */
typedef enum { OU_SWEEP_LEVELS = 0, OU_SWEEP_PERSISTENT = 1, OU_SWEEP_TILES = 2 } ou_sweep_mode_kind;
/* OU_BC_REFLECTIVE is what a parameter file may name (Parameters.cc:59-70);
 * the reference never implements it in the sweep, and plans reject it with
 * OU_ERR_UNSUPPORTED. */
typedef enum {
  OU_BC_VACUUM = 0,
  OU_BC_ISOTROPIC = 1,
  OU_BC_MOST_NORMAL = 2,
  OU_BC_REFLECTIVE = 3
} ou_bc;
typedef enum { OU_QUAD_LS = 0, OU_QUAD_GLC = 1, OU_QUAD_PRODUCT_GLC = 2 } ou_quad_kind;
typedef enum {
  OU_SOLVER_SI = 0,
  OU_SOLVER_GMRES = 1,
  OU_SOLVER_BICGSTAB = 2,
  OU_SOLVER_FGMRES = 3
} ou_solver_kind;

typedef struct ou_plan ou_plan;
typedef struct ou_quadrature ou_quadrature;

/* ------------------------------------------------------------------------ */
/* Errors                                                                    */
/* ------------------------------------------------------------------------ */

/* Message of the last failing call on this thread ("" if none). */
const char *ou_last_error(void);

/* ------------------------------------------------------------------------ */
/* Angular quadrature and moment operators                                  */
/* Replaces RTQuadrature / GLC / LS (src/RTQuadrature.hh:26-98,             */
/* src/GLC.hh:21-41, src/LS.hh:19-30): ctor(sn, L_max, galerkin) +          */
/* build_quadrature(weight_sum, dim) (src/RTQuadrature.cc:27-56).           */
/* kind OU_QUAD_LS / OU_QUAD_GLC use `sn`; OU_QUAD_PRODUCT_GLC uses         */
/* n_polar (Gauss-Legendre nodes on (-1,1)) x n_azim (Chebyshev azimuths). */
/* dim 3 deploys all eight octants (the reference deploys four, defect A.1).*/
/* ------------------------------------------------------------------------ */
int ou_quadrature_create(int kind, int dim, int sn, int n_polar, int n_azim,
                         int L_max, int galerkin, double weight_sum,
                         ou_quadrature **quad);
void ou_quadrature_destroy(ou_quadrature *quad);
/* get_n_dir / get_n_mom (src/RTQuadrature.hh:100-108) */
int ou_quadrature_n_dir(const ou_quadrature *quad);
int ou_quadrature_n_mom(const ou_quadrature *quad);
/* Copies get_omega (n_dir x 3), the weights (n_dir), get_M2D (n_dir x
 * n_mom), get_D2M (n_mom x n_dir), get_l (n_mom) and is_most_normal_direction
 * (2*dim x n_dir flags) out (src/RTQuadrature.hh:39-59,100-133).  Any
 * output pointer may be NULL. */
int ou_quadrature_copy(const ou_quadrature *quad, double *omega, double *weight,
                       double *M2D, double *D2M, int *mom_l,
                       unsigned char *most_normal);

/* ------------------------------------------------------------------------ */
/* Reference-element DG matrices                                            */
/* Replaces FECell (src/FECell.hh:25-79, src/FECell.cc:10-78) on an axis-   */
/* aligned cell of size h[0..dim): M (N x N), G (dim x N x N), E and U      */
/* (2*dim x N x N), normals (2*dim x 3).  Faces ordered -x,+x,-y,+y,-z,+z.  */
/* ------------------------------------------------------------------------ */
int ou_fe_matrices(int dim, int fe_order, const double *h, double *M, double *G,
                   double *E, double *U, double *normals);

/* ------------------------------------------------------------------------ */
/* Problem description and plan                                             */
/* Replaces the RadiativeTransfer constructor + setup()                     */
/* (src/RadiativeTransfer.hh:50-77, src/RadiativeTransfer.cc:49-78) on a    */
/* structured Cartesian mesh.  All arrays are copied; the caller may free  */
/* them after ou_plan_create returns.                                       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int dim;               /* 2 or 3 */
  int fe_order;          /* p: 1..3 (N = (p+1)^dim <= 64) */
  int n_cells[3];        /* cells per axis (n_cells[2] ignored in 2D) */
  double cell_size[3];   /* cell sizes */
  int n_dir;
  const double *omega;   /* n_dir x 3 */
  int n_mom;             /* <= 16 */
  const double *M2D;     /* n_dir x n_mom */
  const double *D2M;     /* n_mom x n_dir */
  const int *mom_l;      /* n_mom Legendre orders */
  int n_mat;
  int L;                 /* sigma_s holds orders 0..L per material */
  const double *sigma_t; /* n_mat          (get_sigma_t, RTMaterialProperties.hh:84-90) */
  const double *sigma_s; /* n_mat x (L+1)  (get_sigma_s, RTMaterialProperties.hh:92-101) */
  const int *material_id;   /* n_cells */
  const double *source;     /* n_cells isotropic source intensity, or NULL */
  double weight_sum;        /* get_weight_sum (Parameters.hh:76) */
  int bc[6];                /* ou_bc per face (get_bc_type, Parameters.hh:89) */
  double inc_flux[6];       /* get_inc_flux (Parameters.hh:73) */
  const unsigned char *most_normal; /* 2*dim x n_dir, or NULL */
  int precision;            /* ou_precision of the sweep */
} ou_problem_desc;

int ou_plan_create(const ou_problem_desc *desc, ou_plan **plan);
void ou_plan_destroy(ou_plan *plan);

/* set_group (src/RadiativeTransfer.hh:73,251-255).  One group: g must be 0. */
int ou_set_group(ou_plan *plan, int g);

/* Sizes of the FP64 vectors the operator maps: n_mom * n_dofs
 * (OperatorDomainMap / OperatorRangeMap, src/RadiativeTransfer.hh:120-126). */
long ou_domain_size(const ou_plan *plan);
long ou_range_size(const ou_plan *plan);
long ou_n_dofs(const ou_plan *plan);

/* Label() (src/RadiativeTransfer.hh:106): "radiative_transfer". */
const char *ou_label(const ou_plan *plan);

/* Apply (src/RadiativeTransfer.hh:94, src/RadiativeTransfer.cc:780-809):
 * y = x - D2M L^-1 M2D S(x) on host vectors (copied in and out). */
int ou_apply(ou_plan *plan, const double *x, double *y);

/* Apply to n_rhs columns at once, x and y host arrays [rhs][m][dof] (the
 * column-major layout of a multi-column Epetra_MultiVector).  The reference
 * Apply reads column 0 only (src/RadiativeTransfer.cc:784-806); every column
 * here gets the operator.  Columns are applied in order on the plan's stream,
 * each equal to its own ou_apply result bit-for-bit. */
int ou_apply_multi(ou_plan *plan, const double *x, double *y, int n_rhs);

/* Apply on device vectors of ou_domain_size doubles, enqueued on `stream`
 * (a hipStream_t, or NULL for the plan's stream).  No host synchronisation. */
int ou_apply_device(ou_plan *plan, const double *d_x, double *d_y, void *stream);

/* compute_scattering_source (src/RadiativeTransfer.hh:90,
 * src/RadiativeTransfer.cc:811-843): stores S(x) in the plan. */
int ou_scatter_source(ou_plan *plan, const double *x);
int ou_scatter_source_device(ou_plan *plan, const double *d_x, void *stream);

/* The task loop initialize_scheduler/get_next_task/sweep/free_buffers
 * (src/RadiativeTransfer.hh:59-87, src/RadiativeTransfer.cc:605-1012) as one
 * call: y += D2M L^-1 (M2D S [+ M q/ws + inflow if with_source]), S from the
 * last ou_scatter_source call (zero after ou_plan_create). */
int ou_sweep(ou_plan *plan, double *y, int with_source);
int ou_sweep_device(ou_plan *plan, double *d_y, int with_source, void *stream);

/* SetUseTranspose / ApplyInverse / UseTranspose / HasNormInf / NormInf
 * (src/RadiativeTransfer.hh:97-114).  The reference returns 0 without doing
 * anything (and UseTranspose() returns true, defect A.8); these report
 * OU_ERR_UNSUPPORTED / false / 0 explicitly. */
int ou_set_use_transpose(ou_plan *plan, int use_transpose);
int ou_apply_inverse(ou_plan *plan, const double *x, double *y);
int ou_use_transpose(const ou_plan *plan);
int ou_has_norm_inf(const ou_plan *plan);
double ou_norm_inf(const ou_plan *plan);

/* ------------------------------------------------------------------------ */
/* Solver drivers (src/ouranos.cc:43-209: source iteration, GMRES and       */
/* BiCGStab around the operator; FGMRES has no reference counterpart).      */
/* Vectors stay in HBM; Arnoldi, inner products and the residual test are   */
/* FP64.  phi holds the initial guess on entry, the solution on exit.       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int solver;        /* ou_solver_kind */
  double tol;        /* relative tolerance: ||b-Ax||/||b|| (SI: ||dphi||/||phi||) */
  int max_it;        /* maximum operator applications */
  int restart;       /* GMRES / FGMRES restart length (30 if 0) */
  int verbose;
} ou_solver_opts;

typedef struct {
  int iterations;      /* operator applications used (FP32 ones for ou_solve_refined) */
  double rel_residual; /* final relative residual (FP64 test) */
  double seconds;      /* wall time of the solve */
  int converged;
  int fp64_applications; /* ou_solve_refined: FP64 residual evaluations */
  int preconditioner_applications;
} ou_solve_info;

int ou_solve(ou_plan *plan, const ou_solver_opts *opts, double *phi,
             ou_solve_info *info);

/* ------------------------------------------------------------------------ */
/* Learned right preconditioner (no reference counterpart: the reference    */
/* solves with PreconditionIdentity, src/ouranos.cc:161).                   */
/*                                                                          */
/* A 3-level convolutional U-Net (32 base channels, BF16 weights and        */
/* activations, FP32 accumulation on MFMA) on a G^3 coarse cell grid.  Per  */
/* coarse cell it reads 12 FP32 features: r_hat[8] (the residual's octant   */
/* means, see ou_coarse_features, divided by their RMS), tau_hat,           */
/* kappa_hat, g and one unused slot; it predicts the moment-0 correction at */
/* the cell's 8 corners (corner k = kx + 2 ky + 4 kz):                      */
/* pred = gain (skip r_hat + (N(+r_hat) - N(-r_hat)) / 2), with gain = g or */
/* 1 and skip = 1 or 0 by the network's mode.  Feature and target arrays    */
/* are [batch][G^3][12] and [batch][G^3][8], cell index x fastest.          */
/* ------------------------------------------------------------------------ */
typedef struct ou_unet ou_unet;

int ou_unet_create(int grid, int max_batch, unsigned long long seed, ou_unet **net);
void ou_unet_destroy(ou_unet *net);
long ou_unet_n_params(const ou_unet *net);
int ou_unet_grid(const ou_unet *net);
/* mode bit 0: add the skip r_hat; bit 1: scale by the prior g (default 0). */
int ou_unet_set_mode(ou_unet *net, int mode);
int ou_unet_mode(const ou_unet *net);
/* FP32 master parameters (layout: per convolution W[27][cout][cin], bias[cout];
 * head w[8][32], b[8]). */
int ou_unet_get_params(const ou_unet *net, float *params);
int ou_unet_set_params(ou_unet *net, const float *params);
/* Host arrays in, host arrays out. */
int ou_unet_predict(ou_unet *net, const float *features, int batch, float *pred);
/* loss = mean_b |pred_b - target_b|^2 / |target_b|^2 and its gradient with
 * respect to the parameters (grad may be NULL). */
int ou_unet_loss_grad(ou_unet *net, const float *features, const float *target, int batch,
                      double *loss, float *grad);
/* One Adam step (lr, beta1 0.9, beta2 0.999) on that loss; *loss is the loss
 * before the update. */
int ou_unet_train_step(ou_unet *net, const float *features, const float *target, int batch,
                       float lr, double *loss);

/* Coarse view of a 3D n^3 plan on the G^3 grid (n a multiple of G):
 *   coarse[C][k]    = mean of the moment-0 nodal values of v in octant
 *                     k = kx + 2 ky + 4 kz of coarse cell C (nodes on a
 *                     mid-plane count half to each side; skipped if v or
 *                     coarse is NULL),
 *   material[C][3]  = tau_hat, kappa_hat, g (skipped if NULL).
 * v is a host vector of ou_domain_size doubles. */
int ou_coarse_features(ou_plan *plan, int grid, const double *v, double *coarse,
                       float *material);

/* The preconditioner's prolongation: z[moment 0] += P(corners), where the
 * corner values corners[C][k] (k = kx + 2 ky + 4 kz) are averaged over the
 * coarse cells sharing each coarse vertex and interpolated trilinearly to
 * every fine node.  z is a host vector of ou_domain_size doubles. */
int ou_coarse_prolong(ou_plan *plan, int grid, const float *corners, double *z);

/* ou_coarse_prolong on device arrays, allocating nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  d_corners ([G^3][8] FP32)
 * and d_z (ou_domain_size doubles, updated in place) are device pointers;
 * the work is ordered on `stream` (NULL: the plan's stream) and not
 * synchronised.  Its scratch (the (G+1)^3 vertex averages and the scale) is
 * a fixed array of the library on each device, sized for grid <= 64;
 * larger grids return OU_ERR_SIZE.  Calls that overlap on one device share
 * that array and race: order them on one stream. */
int ou_coarse_prolong_noalloc(ou_plan *plan, int grid, const float *d_corners, double *d_z,
                              void *stream);

/* Training pairs from the operator's own error propagation: starting from
 * e_0 (host vector, or NULL for e_0 = the source-driven right-hand side
 * D L^-1 (q + inflow)), runs e_{k+1} = e_k - A e_k for n_steps steps on the
 * GPU (FP64 as the plan's precision) and, for every k with bit k of
 * record_mask set, stores the coarse pair (R A e_k, R e_k) into
 * r_out[j][G^3][8], e_out[j][G^3][8] (j counts the recorded steps).  Returns the
 * number of recorded pairs, or a negative ou_status. */
int ou_learned_pairs(ou_plan *plan, int grid, const double *e0, int n_steps,
                     unsigned long long record_mask, double *r_out, double *e_out);

/* Training pairs from plain GMRES(restart) on the source problem: every
 * Arnoldi vector v_j comes with A v_j, so (R A v_j, R v_j) is an exact
 * (residual, correction) pair of the kind FGMRES hands its preconditioner.
 * Runs n_steps applications and records every stride-th pair, at most
 * max_records, into r_out/e_out [j][G^3][8].  Returns the number recorded. */
int ou_learned_pairs_krylov(ou_plan *plan, int grid, int n_steps, int restart, int stride,
                            int max_records, double *r_out, double *e_out);

/* The same from FGMRES(restart) right-preconditioned by `net`: every step's
 * preconditioned vector z_j comes with A z_j, so (R A z_j, R z_j) is an exact
 * pair drawn from the residuals the network meets inside FGMRES (training on
 * the solver's own distribution).  grid must be the network's grid. */
int ou_learned_pairs_fgmres(ou_plan *plan, ou_unet *net, int grid, int n_steps, int restart,
                            int stride, int max_records, double *r_out, double *e_out);

/* FGMRES(restart) on the plan's operator, right-preconditioned by the
 * network (opts->solver must be OU_SOLVER_FGMRES; net NULL = unpreconditioned). */
int ou_solve_learned(ou_plan *plan, ou_unet *net, const ou_solver_opts *opts, double *phi,
                     ou_solve_info *info);

/* Mixed precision: FP64 residuals on plan64, inner FGMRES(restart) solves of
 * the correction on the FP32 plan32 (same problem) to inner_tol relative,
 * right-preconditioned by net (NULL = plain GMRES).  The inner solves keep
 * their Krylov basis in FP32 (FP64 inner products and iterate).  opts->tol
 * is the FP64 relative residual to reach; opts->max_it bounds the FP32
 * applications. */
int ou_solve_refined(ou_plan *plan64, ou_plan *plan32, ou_unet *net,
                     const ou_solver_opts *opts, double inner_tol, double *phi,
                     ou_solve_info *info);

/* ------------------------------------------------------------------------ */
/* Sweep workgroup shape and the batched multi-instance launch.             */
/*                                                                          */
/* The reference sweeps one problem per process (RadiativeTransfer::Apply,  */
/* src/RadiativeTransfer.cc:781-809, called once per Krylov step); these    */
/* calls have no counterpart there.  They replace a host loop of ou_apply   */
/* calls over many independent plans: the members' wavefront levels are     */
/* merged (one sweep launch per level of the deepest member), the           */
/* element-wise stages are one launch each for the whole batch, and the     */
/* sequence is captured once as a HIP graph and replayed.                   */
/* ------------------------------------------------------------------------ */

/* Waves per sweep workgroup (directions in flight per chunk of 64 cells) and
 * direction groups per octant (workgroups per chunk, each with a range of the
 * octant's directions and its own FP64 partial moments, summed in the
 * epilogue).  Shapes change the order of FP64 additions only.  dir_groups
 * lies in [1, 8] (OU_ERR_ARG otherwise); waves must name a kernel built for
 * the element (OU_ERR_UNSUPPORTED otherwise).  A new plan starts with the
 * groups its level parallelism calls for. */
int ou_set_sweep_shape(ou_plan *plan, int waves, int dir_groups);
int ou_sweep_shape(const ou_plan *plan, int *waves, int *dir_groups);

/* Streams of one sweep: n_streams = 2 (the default) launches each level as
 * two kernels, one per octant half, on the plan's stream and a side stream
 * forked from and joined back into it, so that one half's level fills the
 * CUs the other half's leaves idle while its last workgroups drain; 4 does
 * the same with octant quarters on four streams; 1 launches each level as
 * one kernel.  Results are bit-for-bit the same.  No reference
 * counterpart (the reference's scheduler runs tasks as their inputs arrive,
 * RadiativeTransfer.cc:605-778). */
int ou_set_sweep_streams(ou_plan *plan, int n_streams);
int ou_sweep_streams(const ou_plan *plan, int *n_streams);

/* How a sweep is launched.  OU_SWEEP_LEVELS: one launch per wavefront level
 * (octant parts per ou_set_sweep_streams).  OU_SWEEP_PERSISTENT: one launch
 * for the whole sweep; resident workgroups take (chunk, direction group)
 * items in level order and start one as soon as the chunks upwind of it in
 * its own octant and group are done, the dataflow of the reference's
 * scheduler (RadiativeTransfer.cc:605-778, Task.cc:52-62) with the chunk as
 * the unit of readiness.  OU_SWEEP_TILES (2D Q1-Q2 with one material per 8x8
 * tile of cells, 3D Q1 with one material per 4x4x4 tile): one launch per tile
 * level; a workgroup sweeps a tile's cell levels with the traces in registers
 * (OU_ERR_UNSUPPORTED for other plans).  These modes run the same per-cell
 * arithmetic in the same order: at equal direction groups the results are
 * bit-for-bit the same.  3D Q2 plans in FP32 with at most 4 moments and one
 * material per 4x4x4 tile also accept OU_SWEEP_TILES: four directions per
 * wave are solved together on the matrix cores, with results within FP32
 * rounding of the level launches (not the same bits), 3 or 4 waves.
 * ou_set_sweep_shape / ou_sweep_shape set and read the shape of the mode in
 * use (for OU_SWEEP_TILES: waves of 8 directions each in 2D and 4 in 3D,
 * and direction groups per octant).  A new plan starts in OU_SWEEP_TILES
 * where it has a tile sweep, except 3D Q2 plans, whose level launches are
 * faster, and in OU_SWEEP_LEVELS otherwise. */
int ou_set_sweep_mode(ou_plan *plan, int mode);
int ou_sweep_mode(const ou_plan *plan, int *mode);

/* Watchdog of the persistent sweep: a workgroup that waits longer than
 * `seconds` for its upwind level gives up, every workgroup of the launch
 * leaves, the operator's output is NaN and the call that synchronises next
 * (ou_apply, ou_sweep, ou_solve*, ou_sweep_status) returns OU_ERR_TIMEOUT.
 * The reference has no such bound: a missing message makes get_next_task
 * spin forever (RadiativeTransfer.cc:620-623).  Default 10 s; 0 gives up at
 * the first wait that finds its upwind level unfinished. */
int ou_set_sweep_wait_limit(ou_plan *plan, double seconds);
/* OU_ERR_TIMEOUT (and clears the condition) if a persistent sweep of the
 * plan gave up since the last check, else OU_OK.  Synchronises the plan's
 * stream. */
int ou_sweep_status(ou_plan *plan);

typedef struct ou_batch ou_batch;

typedef struct {
  int waves;
  int dir_groups;
  double ms; /* mean event time of one batched application */
  int mode;  /* OU_SWEEP_LEVELS (the chunk sweep's shape) or OU_SWEEP_TILES */
} ou_tune_entry;

/* A batch of n_plans distinct plans sharing dimension, element order, moment
 * count and precision (meshes, materials, quadratures and sources may all
 * differ).  waves / dir_groups <= 0: the autotuner picks the sweep mode and
 * its shape, from its cache (OURANOS_TUNE_CACHE names a file that persists
 * it) or by timing every supported shape of the chunk sweep and, when every
 * member has one, of the tile sweep on this batch; waves / dir_groups > 0 set
 * the chunk sweep's shape.  The batch sets every member's sweep shape;
 * members keep their own state (scattering source, traces, partials), so
 * batched and single calls on one plan must be ordered by the caller, as
 * must batched calls on one batch issued to different streams.  The plans
 * must outlive the batch. */
int ou_batch_create(ou_plan *const *plans, int n_plans, int waves, int dir_groups,
                    ou_batch **batch);
void ou_batch_destroy(ou_batch *batch);
int ou_batch_info(const ou_batch *batch, int *waves, int *dir_groups, int *n_levels,
                  long *n_chunks);
int ou_batch_set_shape(ou_batch *batch, int waves, int dir_groups);
/* use_graph = 0 launches the same kernels eagerly (for comparison). */
int ou_batch_set_graph(ou_batch *batch, int use_graph);
/* n_streams = 2 (the default) launches each merged level as two kernels, the
 * members' first octant halves on the batch's stream and their second halves
 * on a side stream forked from and joined back into it (inside the captured
 * graph as well); 1 launches each level as one kernel.  Results are
 * bit-for-bit the same. */
int ou_batch_set_streams(ou_batch *batch, int n_streams);
/* OU_SWEEP_PERSISTENT sweeps every member's levels in one persistent launch
 * (inside the captured graph); OU_SWEEP_LEVELS launches the merged levels;
 * OU_SWEEP_TILES launches the members' merged tile levels (every member needs
 * a tile sweep, see ou_set_sweep_mode; OU_ERR_UNSUPPORTED otherwise).  At
 * equal direction groups the results are bit-for-bit the same, and equal each
 * member's own ou_apply in the same mode and shape.  ou_batch_set_shape and
 * ou_batch_info set and read the shape of the mode in use.  A watchdog trip
 * (the first member's wait limit) makes the next synchronising batched call
 * return OU_ERR_TIMEOUT. */
int ou_batch_set_sweep_mode(ou_batch *batch, int mode);
int ou_batch_sweep_mode(const ou_batch *batch, int *mode);

/* Apply of every member: d_y[i] = d_x[i] - D L^-1 M S d_x[i] (device vectors;
 * d_x / d_y are host arrays of n_plans device pointers; all d_y distinct and
 * distinct from the d_x).  stream NULL = the batch's stream.  No host
 * synchronisation beyond the upload of the pointer table. */
int ou_batch_apply_device(ou_batch *batch, const double *const *d_x, double *const *d_y,
                          void *stream);
/* ou_sweep_device of every member: d_y[i] += D2M L^-1 (M2D S_i [+ M q_i/ws +
 * inflow if with_source]). */
int ou_batch_sweep_device(ou_batch *batch, double *const *d_y, int with_source,
                          void *stream);

/* Source iteration (ouranos.cc:52-130) of every member at once from the
 * initial guesses d_phi (n_plans device vectors, updated in place):
 * phi_i <- D L^-1 (M2D S(phi_i) + M q_i/ws + inflow), one captured graph per
 * iteration for the whole batch, until every member's change
 * |phi_new - phi| / |phi_new| is below tol or max_it iterations.
 * iterations[i] = the iteration at which member i first met tol (-1 if never),
 * change[i] = its last change; *n_iterations = iterations run.  A member that
 * met tol is not swept again: d_phi[i] returns that iterate.  Each member's
 * iterates equal its own ou_solve SI iterates bit-for-bit.  verbose != 0
 * prints the progress every 25 iterations. */
int ou_batch_solve_si(ou_batch *batch, double *const *d_phi, double tol, int max_it,
                      int *iterations, double *change, int *n_iterations, int verbose);

/* Re-runs the autotuner (reps timed applications per shape), keeps the
 * fastest shape and writes up to max_entries timings. */
int ou_batch_tune(ou_batch *batch, int reps, ou_tune_entry *entries, int max_entries,
                  int *n_entries);

/* `warmup` untimed then `steps` timed batched applications on the batch's
 * stream; *ms = mean event time of one. */
int ou_batch_time_apply(ou_batch *batch, const double *const *d_x, double *const *d_y,
                        int warmup, int steps, double *ms);

/* ------------------------------------------------------------------------ */
/* Device memory and timing plumbing (for callers without a HIP runtime of  */
/* their own, e.g. ctypes).                                                 */
/* ------------------------------------------------------------------------ */
int ou_device_alloc(size_t bytes, void **ptr);
int ou_device_free(void *ptr);
int ou_memcpy_htod(void *dst, const void *src, size_t bytes);
int ou_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int ou_memcpy_dtod(void *dst, const void *src, size_t bytes);
int ou_device_synchronize(void);
int ou_device_name(char *buf, int len);

typedef struct {
  double ms_total;        /* event time over all timed applications */
  double ms_sweep_kernel; /* summed event time of the sweep-level launches */
  double ms_other;        /* scatter source + trace init + epilogue */
  int n_sweep_launches;   /* sweep-level launches per application */
  int applications;
} ou_timing;

/* Runs `warmup` untimed then `steps` timed applications of the operator on
 * device vectors, with HIP events on the plan's stream around each launch
 * group.  Inputs are resident in HBM before the first event. */
int ou_time_apply(ou_plan *plan, const double *d_x, double *d_y, int warmup,
                  int steps, ou_timing *timing);

/* ------------------------------------------------------------------------ */
/* Input files                                                              */
/* Readers of the reference's parameter, geometry and cross-section files,  */
/* a subset of deal.II's ParameterHandler format (`set name = value`,       */
/* `subsection name` ... `end`, `#` comments, a trailing backslash          */
/* continues a line).  They replace Parameters (src/Parameters.cc:10-302),  */
/* the reading half of Geometry (src/Geometry.cc:10-128) and                */
/* RTMaterialProperties (src/RTMaterialProperties.cc:10-81), with the same  */
/* entry names, defaults and patterns.  A missing file, an undeclared entry */
/* or a value outside its pattern is OU_ERR_ARG, with the file and line in  */
/* ou_last_error() (the reference throws "Invalid input parameters file."). */
/* ------------------------------------------------------------------------ */
typedef struct ou_parameters ou_parameters;
typedef struct ou_geometry ou_geometry;
typedef struct ou_material_properties ou_material_properties;

/* Parameters::get_* (src/Parameters.hh:40-104) */
typedef struct {
  int dim;                  /* get_dimension */
  int fe_order;             /* get_fe_order */
  int n_refinements;        /* get_n_refinements (adaptive cycles) */
  double refinement_factor; /* get_refinement_factor */
  double coarsening_factor; /* get_coarsening_factor */
  int solver_type;          /* OU_SOLVER_SI, OU_SOLVER_GMRES or OU_SOLVER_BICGSTAB */
  int max_outer_it;         /* get_max_outer_it */
  int max_inner_it;         /* get_max_inner_it */
  double outer_tol;         /* get_outer_tolerance */
  double inner_tol;         /* get_inner_tolerance */
  int sn_order;             /* get_sn_order */
  int quad_type;            /* OU_QUAD_LS or OU_QUAD_GLC (get_quad_type) */
  double weight_sum;        /* get_weight_sum: 1, 2 pi or 4 pi */
  int galerkin;             /* get_galerkin */
  int bc_type[6];           /* ou_bc per face -x,+x,-y,+y,-z,+z (vacuum past 2*dim) */
  int n_groups;             /* get_n_groups */
  int n_src;                /* get_n_src */
} ou_parameters_info;

int ou_parameters_read(const char *path, ou_parameters **prm);
void ou_parameters_destroy(ou_parameters *prm);
int ou_parameters_get_info(const ou_parameters *prm, ou_parameters_info *info);
/* get_geometry_filename (which = 0), get_xs_filename (1), get_output_filename
 * (2): copies at most len-1 characters and a terminating zero into buf (buf
 * may be NULL); *needed = the name's length + 1. */
int ou_parameters_get_filename(const ou_parameters *prm, int which, char *buf, int len,
                               int *needed);
/* get_inc_flux(face, group) and get_src(source, group) */
int ou_parameters_get_inc_flux(const ou_parameters *prm, int face, int group, double *value);
int ou_parameters_get_src(const ou_parameters *prm, int source, int group, double *value);

/* Geometry's reader: a subdivided hyper-rectangle with a material and a
 * source id per coarse cell, globally refined n_global_refinements times
 * (every child keeps its parent's ids). */
typedef struct {
  int dim;
  double bottom_left[3];    /* Bottom left point */
  double upper_right[3];    /* Upper right point */
  int n_subdivisions[3];    /* get_n_subdivisions */
  int n_global_refinements;
  int n_materials;          /* get_n_materials: largest material id + 1 */
  int n_cells[3];           /* cells per axis after the refinements (1 past dim) */
  double cell_size[3];
} ou_geometry_info;

int ou_geometry_read(const char *path, int dim, ou_geometry **geo);
void ou_geometry_destroy(ou_geometry *geo);
int ou_geometry_get_info(const ou_geometry *geo, ou_geometry_info *info);
/* get_material_ids(i, j[, k]) and get_source_ids(i, j[, k]) of coarse cell
 * (i, j, k) (k = 0 in 2D; src/Geometry.hh:45-52); either output may be NULL. */
int ou_geometry_get_ids(const ou_geometry *geo, int i, int j, int k, int *material_id,
                        int *source_id);
/* Ids of every cell of the refined mesh, x fastest (the plans' cell order):
 * n_cells[0] * n_cells[1] * n_cells[2] values each; either output may be NULL. */
int ou_geometry_cell_ids(const ou_geometry *geo, int *material_id, int *source_id);

/* RTMaterialProperties(file, n_materials, n_groups) and its getters
 * (src/RTMaterialProperties.hh:31-46, layout :84-101). */
int ou_material_properties_read(const char *path, int n_materials, int n_groups,
                                ou_material_properties **mp);
void ou_material_properties_destroy(ou_material_properties *mp);
int ou_material_properties_get_L_max(const ou_material_properties *mp, int *L_max);
int ou_material_properties_get_sigma_t(const ou_material_properties *mp, int mat, int group,
                                       double *value);
/* get_sigma_s(mat, group, group_to, l) */
int ou_material_properties_get_sigma_s(const ou_material_properties *mp, int mat, int group,
                                       int group_to, int l, double *value);

/* ------------------------------------------------------------------------ */
/* Several energy groups (src/ouranos.cc:62-205): one plan per group (its    */
/* sigma_t, sigma_s(g->g), source and inflow), coupled by the scattering     */
/* from the other groups.                                                    */
/* ------------------------------------------------------------------------ */
/* compute_outer_scattering_source (src/RadiativeTransfer.cc:845-876), one
 * other group at a time: S_m = sigma_s[mat][l(m)] M x_m with the caller's
 * table sigma_s (n_mat x (L+1), e.g. sigma_s(g'->g) of each material), added
 * to the stored scattering source when accumulate != 0.  x is group g''s
 * moments (host, n_mom * n_dofs).  A following ou_sweep(plan, y, 1) sweeps
 * the stored source with the group's own source and inflow: with every other
 * group accumulated, y = D L^-1 (sum_g' S_g'->g + M q_g / ws + inflow_g), the
 * right-hand side of group g.  (The reference uses the moment index as the
 * Legendre order here too; this uses l(m), defect A.5.) */
int ou_scatter_source_with(ou_plan *plan, const double *x, const double *sigma_s,
                           int accumulate);
/* Solves A phi = rhs for the plan's operator A with the solver of opts
 * (SI: phi <- phi - A phi + rhs, stopping on the relative change; GMRES /
 * BICGSTAB on the residual), from the initial guess in phi (host vectors). */
int ou_solve_rhs(ou_plan *plan, const ou_solver_opts *opts, const double *rhs, double *phi,
                 ou_solve_info *info);

/* output_results (src/ouranos.cc:211-253): legacy VTK of the flux moments.
 * Writes `<filename>0000.vtk` (one VTK cell per mesh cell with its own 2^dim
 * corner points; point fields flux_<ggg>g_<mmm>m with the corner values of
 * every group and moment; cell field partitioning = 0) and the VisIt record
 * `<filename>.vtk` naming it.  flux is [group][moment][dof] on the mesh of
 * n_cells cells of size cell_size from origin, cells x fastest and each
 * cell's (fe_order+1)^dim dofs lexicographic (the plans' layout). */
int ou_write_vtk(const char *filename, int dim, int fe_order, const int *n_cells,
                 const double *cell_size, const double *origin, int n_groups, int n_mom,
                 const double *flux);

#ifdef __cplusplus
}
#endif
#endif
