/* cup2d_poisson.h -- C ABI of the MI355X pressure-Poisson solver
 * (libcup2d_poisson.so, HIP for gfx950).
 *
 * Drop-in for the reference's solver plugin boundary:
 *   class PoissonSolver { virtual void solve(const ScalarGrid *input, ScalarGrid *output) = 0; };
 *   std::shared_ptr<PoissonSolver> makePoissonSolver(SimulationData&);
 *     (source/Poisson/Base.h:8-15, source/Poisson/Base.cpp:8-30)
 * whose GPU implementation is ExpAMRSolver (source/Poisson/ExpAMRSolver.h:14-307)
 * -> LocalSpMatDnVec (source/Poisson/LocalSpMatDnVec.h:44-126)
 * -> BiCGSTABSolver (source/Poisson/BiCGSTAB.cuh:24-148).
 *
 * Conventions: unknowns are ordered block-major, 64 per 8x8-cell block, cell
 * iy*8+ix (ExpAMRSolver.h:74-75); all vectors are FP64; every entry point
 * returns a cup2d_ps_status; no C++ exception crosses the ABI.  A handle owns
 * one HIP stream and its device buffers; it is not thread-safe (serialise
 * calls per handle).
 */
#ifndef CUP2D_POISSON_H
#define CUP2D_POISSON_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cup2d_ps cup2d_ps;

typedef enum {
  CUP2D_PS_OK = 0,
  CUP2D_PS_EINVAL = 1,   /* invalid argument / mesh */
  CUP2D_PS_ENOMESH = 2,  /* solve before cup2d_ps_set_mesh */
  CUP2D_PS_EHIP = 3,     /* HIP runtime error (message in cup2d_ps_last_error) */
  CUP2D_PS_ENOMEM = 4,   /* device or host allocation failed */
  CUP2D_PS_ENODEV = 5    /* no usable gfx950 device */
} cup2d_ps_status;

typedef enum { CUP2D_PS_FP64 = 0, CUP2D_PS_MIXED = 1 } cup2d_ps_precision;
typedef enum { CUP2D_PS_NORM_INF = 0, CUP2D_PS_NORM_L2 = 1 } cup2d_ps_norm;

/* Block description of the mesh: what the reference reads from
 * cubism::BlockInfo (level, index) at ExpAMRSolver.cpp:186-249.
 * Block b owns unknowns 64b..64b+63; the tree must be 2:1 balanced across
 * faces (ExpAMRSolver.cpp:177 throws otherwise). */
typedef struct {
  int bpdx, bpdy;        /* level-0 blocks per direction (-bpdx, -bpdy) */
  int levels;            /* -levelMax */
  int periodic_x, periodic_y;
  double extent;         /* -extent: length of the longer side */
  long long n_blocks;
  const int *level;      /* [n_blocks] */
  const int *bx;         /* [n_blocks] block index at its level */
  const int *by;         /* [n_blocks] */
} cup2d_mesh_desc;

typedef struct {
  int device;             /* HIP device ordinal */
  int precision;          /* cup2d_ps_precision */
  int mean_constraint;    /* bMeanConstraint (SimulationData.h:98): 0 = post-hoc mean removal only,
                             nonzero = the (0,0) row of the corner block is replaced by
                             sum h^2 x = 0 (BiCGSTAB.cu:393-409) */
  int norm;               /* stopping norm: CUP2D_PS_NORM_INF as BiCGSTAB.cu:572-587, or
                             CUP2D_PS_NORM_L2 as AMRSolver.cpp:466-471 */
  int max_iter;           /* iteration cap (reference GPU: 1000, BiCGSTAB.cu:468) */
  int chunk;              /* iterations per captured graph launch; 0 = automatic */
  int use_graph;          /* 1 = replay iteration chunks as a hipGraph */
  int profile;            /* 1 = per-phase device timers (the reference's DeviceProfiler,
                             DeviceProfiler.cuh:26-107, used at BiCGSTAB.cu:191-619): kernels are
                             launched one by one between HIP events and the phase times are
                             returned in cup2d_ps_stats.ms_phase; 0 = off (graph replay) */
  int item_blocks;        /* batch handles: blocks per work item, 1..256; 0 = the value tuned in
                             this process for the batch's shape, else 8 */
  int schedule;           /* 0 = fused kernels (default); 1 = one operation per launch with a
                             host check of the stopping norm every iteration, the structure of the
                             reference GPU loop (BiCGSTAB.cu:412-620), FP64 single solves only: the
                             unfused comparator */
  int xcd_run;            /* batch handles: the operator kernels take their work items in runs of
                             xcd_run consecutive items per XCD (1 = plain grid-stride order; a power
                             of two up to 32); 0 = the value tuned in this process, else 8 */
  int reserved_i[1];
  double eps;             /* denominator guard; reference GPU 1e-21 (BiCGSTAB.cu:427),
                             reference CPU 1e-100 (AMRSolver.cpp:167) */
  double breakdown_tol;   /* restart when rho^2 < breakdown_tol |r|^2 |rhat|^2 (BiCGSTAB.cu:483) */
  double inner_tol;       /* mixed precision: the FP32 recurrence residual may fall to inner_tol
                             times the last FP64 residual before the residual is recomputed in
                             FP64 and BiCGSTAB restarts from the best iterate; 0 = 1e-2 */
  double reserved_d[5];
} cup2d_ps_config;

typedef struct {
  int iterations;         /* BiCGSTAB iterations (inner iterations for the mixed solver) */
  int restarts;           /* breakdown restarts (BiCGSTAB.cu:488-509) */
  int converged;          /* 1 when the stopping test was met; 0 returns the best iterate x_opt */
  int refinements;        /* FP64 residual updates (mixed precision) */
  double err_init;        /* stopping norm of r0 = b - A x0 */
  double err_final;       /* stopping norm of the returned iterate (recurrence residual) */
  double ms;              /* wall time of the solve on the handle's stream, ms */
  double true_rel2;       /* |b - A x|_2 / |b|_2 of the returned x, recomputed in FP64 on the device
                             after the solve (b's mean row taken as 0 with mean_constraint) */
  double true_relinf;     /* |b - A x|_inf / |b|_inf, likewise */
  double ms_phase[5];     /* with cfg.profile: device ms per phase, indexed by cup2d_ps_phase;
                             0 otherwise */
  double reserved[2];
} cup2d_ps_stats;

/* Phases of cup2d_ps_stats.ms_phase (the reference profiler's tags "KerSpMV",
 * "Prec", ... at BiCGSTAB.cu:191-619).  A fused kernel counts in the phase of
 * its main operation: the preconditioner kernels also update p or s. */
typedef enum {
  CUP2D_PS_PHASE_SPMV = 0,    /* A z (with the dot products fused into it) */
  CUP2D_PS_PHASE_PREC = 1,    /* P_inv r (with the p / s update fused into it) */
  CUP2D_PS_PHASE_UPDATE = 2,  /* x and r updates, other vector passes and dot products */
  CUP2D_PS_PHASE_SCALAR = 3,  /* scalar recurrences and stopping tests */
  CUP2D_PS_PHASE_SETUP = 4    /* r0 = b - A x0, residual replacements, mean removal, true residual */
} cup2d_ps_phase;

/* Fills the defaults: FP64, mean_constraint 0, NORM_INF, max_iter 1000,
 * eps 1e-100, breakdown_tol 1e-16, inner_tol 0, graph replay on. */
void cup2d_ps_config_default(cup2d_ps_config *cfg);

/* Creates a solver (reference: ExpAMRSolver ctor, ExpAMRSolver.cpp:25-99, which
 * builds P_inv = -K^-1 and the BiCGSTABSolver, BiCGSTAB.cu:8-47). */
int cup2d_ps_create(const cup2d_ps_config *cfg, cup2d_ps **out);
int cup2d_ps_destroy(cup2d_ps *h);

/* Mesh changed: assemble the operator and upload it
 * (reference: ExpAMRSolver::getMat + LocalSpMatDnVec::make + BiCGSTABSolver::updateAll,
 * ExpAMRSolver.cpp:180-313, LocalSpMatDnVec.cpp:81-189, BiCGSTAB.cu:142-244). */
int cup2d_ps_set_mesh(cup2d_ps *h, const cup2d_mesh_desc *mesh);

/* Same operator, assembled by GPU kernels from the block description (the
 * same getMat / makeFlux / interpolate restatement, ExpAMRSolver.cpp:100-313,
 * run per block and per coarse-fine row on the device; only the 12 bytes per
 * block of `mesh` cross PCIe).  The result equals cup2d_ps_set_mesh's bit for
 * bit; the same meshes are rejected with the same messages.  No plain CSR
 * copy is kept: spmv/spmv_time kind 1 and cup2d_ps_get_csr return
 * CUP2D_PS_EINVAL on such a mesh. */
int cup2d_ps_set_mesh_on_device(cup2d_ps *h, const cup2d_mesh_desc *mesh);

/* Solve A x = b (reference: ExpAMRSolver::solve, ExpAMRSolver.cpp:354-416).
 * b: right-hand side (not modified; with mean_constraint its mean row is
 * taken as 0, ExpAMRSolver.cpp:340-345); x: initial guess on entry, solution
 * with its h^2-weighted mean removed on exit.  tol_abs / tol_rel are
 * PoissonTol / PoissonTolRel; the caller applies the reference's step<10 rule
 * (tol 0, 100 restarts, ExpAMRSolver.cpp:366-368).  Host pointers. */
int cup2d_ps_solve(cup2d_ps *h, const double *b, double *x, double tol_abs, double tol_rel,
                   int max_restarts, cup2d_ps_stats *stats);

/* Same with device pointers (n doubles each, resident in HBM). */
int cup2d_ps_solve_device(cup2d_ps *h, const double *d_b, double *d_x, double tol_abs,
                          double tol_rel, int max_restarts, cup2d_ps_stats *stats);

/* Largest mesh (unknowns) of cup2d_ps_solve_device_noalloc: 2048 blocks. */
#define CUP2D_PS_NOALLOC_MAX_N 131072

/* cup2d_ps_solve_device without any allocation: no hipMalloc / hipFree /
 * graph instantiation in the call.  No counterpart in the reference; it is
 * for callers that cannot allocate (serving loops, harnesses that forbid
 * allocation).  The work vectors, partial sums and solver state live in a
 * fixed per-device array of the library (about 9.5 MB, held until the
 * process ends) instead of the handle's buffers, so: FP64 handles with the
 * fused schedule only, meshes of at most CUP2D_PS_NOALLOC_MAX_N unknowns
 * (CUP2D_PS_EINVAL beyond), d_b must not alias d_x, and calls that overlap
 * on one device race (every call on a device shares its array).  Same
 * results as cup2d_ps_solve_device (the same kernels in the same order);
 * stats as there, without per-phase times. */
int cup2d_ps_solve_device_noalloc(cup2d_ps *h, const double *d_b, double *d_x, double tol_abs,
                                  double tol_rel, int max_restarts, cup2d_ps_stats *stats);

/* The pressure step around the solve, device-resident (reference
 * PressureSingle::operator(), PressureSingle.cpp:842-879):
 *   b  = rhs - L p_old   (updatePressureRHS1, :401-456, applied at :848-849;
 *                         L = the mesh's operator without the mean-row replacement.
 *                         At cells on a coarse-fine face the reference takes L p_old
 *                         from the Cubism lab's interpolated ghosts and flux
 *                         correction; here it is the explicit operator's row, the
 *                         same ghost interpolation in face form -- equality with
 *                         Cubism's is parity unpinned, Cubism being absent)
 *   dp = solve A dp = b from dp = 0 (:845, :851; as cup2d_ps_solve_device)
 *   p  = dp + (p_old - <dp>_h2), <.>_h2 the h^2-weighted mean (:853-879).
 * d_rhs: the divergence part of the RHS (updatePressureRHS's output, PressureSingle.cpp:324-398),
 * n doubles, not modified.  d_p: p_old on entry, p on exit (must not alias d_rhs).
 * The stats are the solve's. */
int cup2d_ps_pressure_step_device(cup2d_ps *h, const double *d_rhs, double *d_p, double tol_abs,
                                  double tol_rel, int max_restarts, cup2d_ps_stats *stats);

/* The two operators either side of the pressure step on device arrays.
 * Ghost values at non-periodic domain faces mirror the boundary cell: the
 * face-normal velocity component negated (Definitions.h:84-117, both wall and
 * free-slip), the pressure copied (zero gradient).  Cells on a coarse-fine
 * block face are computed in face form, as the sum over their faces of the
 * face values the reference stores for the flux correction
 * (PressureSingle.cpp:350-397, :139-185): across a coarser face with the
 * explicit operator's coarse-fine ghost (ExpAMRSolver.cpp:100-128), across a
 * finer face as the sum of the two fine cells' face values, so coarse and
 * fine fluxes are equal and opposite.  The reference takes these ghosts
 * from Cubism's lab interpolation (absent here): equality with it is parity
 * unpinned; linear fields are differentiated exactly there and the RHS sums
 * to the boundary flux.
 *
 * Divergence RHS (updatePressureRHS, PressureSingle.cpp:324-349):
 *   rhs = h/(2 dt) [(u_E - u_W) + (v_N - v_S)] - h/(2 dt) chi [(ud_E - ud_W) + (vd_N - vd_S)]
 * d_u, d_v: velocity components; d_udef_u, d_udef_v: the obstacles' deformation
 * velocity; d_chi: the characteristic function; all n doubles, block-major. */
int cup2d_ps_pressure_rhs_device(cup2d_ps *h, const double *d_u, const double *d_v,
                                 const double *d_udef_u, const double *d_udef_v, const double *d_chi,
                                 double dt, double *d_rhs);

/* Velocity correction (pressureCorrectionKernel + pressureCorrection,
 * PressureSingle.cpp:122-210): u += (-dt/2 h (p_E - p_W)) / h^2, v likewise
 * with p_N - p_S.  d_p is not modified. */
int cup2d_ps_pressure_correct_device(cup2d_ps *h, const double *d_p, double dt, double *d_u,
                                     double *d_v);

/* ---- the rest of the projection time step on device arrays ---------------
 * Block-major fields of the handle's mesh (n doubles each).  Ghost cells at
 * coarse-fine faces: the coarse cell containing the ghost plus its gradient
 * times the offset (across a coarser face), the average of the four fine
 * cells under it (across a finer face); the reference takes them from
 * Cubism's AMR lab (absent here: parity unpinned there). */

/* Advection-diffusion, in place (advDiff::operator(), advDiff.cpp:238-282):
 * u_old = u; u1 = u_old + 0.5 RHS(u) / h^2; u = u_old + RHS(u1) / h^2, RHS =
 * KernelAdvectDiffuse (advDiff.cpp:140-236): -dt h ((u + uinf_x) du/dx + (v +
 * uinf_y) du/dy) with the upwinded WENO5 differences of advDiff.cpp:18-82,
 * plus nu dt times the 5-point Laplacian, flux-corrected at coarse-fine
 * faces; the same for v.  Domain faces: the "freespace" velocity lab
 * (Definitions.h:84-117).  Uses the handle's solver work vectors as
 * scratch (not between the calls of a solve). */
int cup2d_ps_advect_diffuse_device(cup2d_ps *h, double *d_u, double *d_v, double uinf_x,
                                   double uinf_y, double nu, double dt);

/* Brinkman penalization toward the rigid velocity (us, vs), in place
 * (PressureSingle::penalize, PressureSingle.cpp:268-322, one obstacle, no
 * deformation velocity): where chi > 0, alpha = 1 / (1 + lambda dt) if
 * chi > 0.5 else 1, u = alpha u + (1 - alpha) us. */
int cup2d_ps_penalize_device(cup2d_ps *h, const double *d_chi, double us, double vs, double lambda,
                             double dt, double *d_u, double *d_v);

/* chi of the disk of radius r centred at (cx, cy): the signed distance
 * r - |x - c| (FillBlocks_Cylinder, ShapeLibrary.cpp:19-34) through Towers'
 * mollified Heaviside within one cell of the circle (PutObjectsOnGrid.cpp:
 * 70-96). */
int cup2d_ps_chi_disk_device(cup2d_ps *h, double cx, double cy, double r, double *d_chi);

/* *umax = max(|u + uinf_x|, |v + uinf_y|, |u|, |v|) over the mesh
 * (findMaxU, Helpers.cpp:155-207), for the time-step size. */
int cup2d_ps_max_speed_device(cup2d_ps *h, const double *d_u, const double *d_v, double uinf_x,
                              double uinf_y, double *umax);

/* Mesh adaptation (AdaptTheMesh.cpp:11-122).  Refinement score of every
 * block, written to the host array score[n_blocks]: max |omega| over the
 * block, omega = (0.5 / h) ((u(x, y-1) - u(x, y+1)) + (v(x+1, y) - v(x-1, y)))
 * (KernelVorticity, Helpers.h:103-116), or 2 rtol when the obstacle's
 * interface (0 < chi < 0.9, GradChiOnTmp, AdaptTheMesh.cpp:11-71) is in the
 * block or within 2 cells (3 on the finest level) across its faces.  The
 * caller tags (refine above Rtol, compress below Ctol) and calls
 * cup2d_mesh_adapt (include/cup2d_mesh.h). */
int cup2d_ps_block_score_device(cup2d_ps *h, const double *d_u, const double *d_v,
                                const double *d_chi, double rtol, double *score);

/* Fields of old_mesh onto new_mesh by the plan of cup2d_mesh_adapt: copies;
 * refinement from an ancestor (the containing ancestor cell plus its
 * gradient times the offset: exact for linear fields, keeps each parent
 * cell's mean one level down); compression (the average of four fine cells).
 * d_old[f] (64 old_mesh->n_blocks doubles) -> d_new[f] (64 new_mesh->n_blocks),
 * f < n_fields.  Independent of the handle's current mesh.  The reference
 * refines and compresses in Cubism (absent here): parity unpinned. */
int cup2d_ps_transfer_device(cup2d_ps *h, const cup2d_mesh_desc *old_mesh,
                             const cup2d_mesh_desc *new_mesh, const int *src_kind, const long long *src,
                             int n_fields, const double *const *d_old, double *const *d_new);

/* y = A x on the device (the operator of the last cup2d_ps_set_mesh, with the
 * mean-constraint row if enabled).  kind 0: stencil kernel used by the solver,
 * kind 1: plain CSR kernel.  Reference: BiCGSTABSolver::hd_cusparseSpMV
 * (BiCGSTAB.cu:315-410). */
int cup2d_ps_spmv_device(cup2d_ps *h, const double *d_x, double *d_y, int kind);

/* Times `reps` back-to-back launches of one SpMV kind on the handle's stream
 * with HIP events (after one untimed warm-up launch); writes the mean ms
 * per launch. */
int cup2d_ps_spmv_time(cup2d_ps *h, const double *d_x, double *d_y, int kind, int reps,
                       double *ms_per_launch);

/* Times one of the solver's five iteration kernels alone: `reps` launches
 * back to back on the handle's stream (HIP events), each redoing the work
 * of one iteration from the same state (set up from x = 0 and a fixed b; the
 * handle's work vectors are overwritten, the next solve starts afresh).
 * which: 0 p update + preconditioner, 1 nu = A z1, 2 s update +
 * preconditioner, 3 t = A z2, 4 x and r update.  No counterpart in the
 * reference (DeviceProfiler.cuh:26-107 times the reference's phases). */
int cup2d_ps_kernel_time(cup2d_ps *h, int which, int reps, double *ms_per_launch);

/* Block preconditioner z = P_inv r on every block (reference: cublasDgemm at
 * BiCGSTAB.cu:520, :545). */
int cup2d_ps_precond_device(cup2d_ps *h, const double *d_r, double *d_z);

/* The same with the BF16 MFMA kernel of the mixed-precision solver: P_inv
 * rounded to bf16, the input split into three bf16 terms, FP32
 * accumulation; FP32 vectors (n floats each). */
int cup2d_ps_precond_bf16_device(cup2d_ps *h, const float *d_r, float *d_z);

/* True residual of x: out[0] = |b - A x|_2, out[1] = |b - A x|_inf,
 * out[2] = |b|_2, out[3] = |b|_inf, all in FP64 on the device (b's mean row
 * taken as 0 with mean_constraint). */
int cup2d_ps_residual_device(cup2d_ps *h, const double *d_b, const double *d_x, double out[4]);

/* Operator sizes of the current mesh: n, nnz, explicit (coarse-fine) rows. */
int cup2d_ps_sizes(const cup2d_ps *h, long long *n, long long *nnz, long long *n_explicit,
                   long long *mean_row);
/* Copies the assembled FP64 CSR (row_ptr[n+1], col[nnz], val[nnz]). */
int cup2d_ps_get_csr(const cup2d_ps *h, long long *row_ptr, int *col, double *val);

/* Host-only assembly (no device needed): the FP64 CSR of the operator of
 * `mesh` (ExpAMRSolver::getMat, ExpAMRSolver.cpp:180-313).  Returns nnz, or
 * -(needed nnz) when cap is too small (arrays untouched), or -1 on an invalid
 * mesh (message via cup2d_ps_last_error(NULL)).  *mean_row receives the
 * replaced row or -1. */
long long cup2d_op_csr(const cup2d_mesh_desc *mesh, int mean_constraint, long long *row_ptr,
                       int *col, double *val, long long cap, long long *mean_row);

/* Device memory helpers for callers without an allocator of their own. */
int cup2d_ps_dmalloc(cup2d_ps *h, size_t bytes, void **ptr);
int cup2d_ps_dfree(cup2d_ps *h, void *ptr);
int cup2d_ps_h2d(cup2d_ps *h, void *dst, const void *src, size_t bytes);
int cup2d_ps_d2h(cup2d_ps *h, void *dst, const void *src, size_t bytes);

/* Message of the last failure on this handle (or, when h is NULL, of the
 * calling thread's last failure without a handle: cup2d_ps_create,
 * cup2d_op_csr). */
const char *cup2d_ps_last_error(const cup2d_ps *h);

/* ------------------------------------------------------------------------
 * Batched solves: many independent meshes solved together on one device.
 * The reference solves one mesh per ExpAMRSolver::solve call
 * (ExpAMRSolver.cpp:354-416); a batch handle runs that solve for every
 * instance at once, each instance with its own scalars, stopping test,
 * restarts and (mixed precision) FP64 refinement passes.  Vectors are the
 * instances' vectors concatenated in instance order (offsets from
 * cup2d_psb_offsets).  Config fields as for cup2d_ps_create (item_blocks
 * sets the blocks per work item; schedule must be 0).
 * ------------------------------------------------------------------------ */
typedef struct cup2d_psb cup2d_psb;

int cup2d_psb_create(const cup2d_ps_config *cfg, cup2d_psb **out);
int cup2d_psb_destroy(cup2d_psb *h);

/* Assembles and uploads the operators of n_inst meshes (1..4096) with the
 * handle's mean_constraint (per instance as cup2d_ps_set_mesh). */
int cup2d_psb_set_meshes(cup2d_psb *h, int n_inst, const cup2d_mesh_desc *meshes);

/* Same operators, assembled by GPU kernels for all instances in one pipeline
 * (as cup2d_ps_set_mesh_on_device; reference ExpAMRSolver.cpp:100-313 per
 * instance).  Bit-identical to cup2d_psb_set_meshes's; the same meshes are
 * rejected with the same messages ("mesh k: ..."). */
int cup2d_psb_set_meshes_on_device(cup2d_psb *h, int n_inst, const cup2d_mesh_desc *meshes);

/* The pressure step around the solve for every instance at once (as
 * cup2d_ps_pressure_step_device, PressureSingle.cpp:842-879): device vectors
 * of all instances concatenated; stats[n_inst] are the solves'. */
int cup2d_psb_pressure_step_device(cup2d_psb *h, const double *d_rhs, double *d_p, double tol_abs,
                                   double tol_rel, int max_restarts, cup2d_ps_stats *stats);

/* The divergence RHS and the velocity correction for every instance at once
 * (as cup2d_ps_pressure_rhs_device / cup2d_ps_pressure_correct_device). */
int cup2d_psb_pressure_rhs_device(cup2d_psb *h, const double *d_u, const double *d_v,
                                  const double *d_udef_u, const double *d_udef_v,
                                  const double *d_chi, double dt, double *d_rhs);
int cup2d_psb_pressure_correct_device(cup2d_psb *h, const double *d_p, double dt, double *d_u,
                                      double *d_v);

/* offsets[k] = first unknown of instance k, offsets[n_inst] = total (n_inst+1
 * entries). */
int cup2d_psb_offsets(const cup2d_psb *h, long long *offsets);

/* Solves every instance (semantics of cup2d_ps_solve per instance, the same
 * tol_abs / tol_rel / max_restarts for all); stats: n_inst entries or NULL
 * (ms is the whole batch's time).  Host pointers. */
int cup2d_psb_solve(cup2d_psb *h, const double *b, double *x, double tol_abs, double tol_rel,
                    int max_restarts, cup2d_ps_stats *stats);
/* Same with device pointers. */
int cup2d_psb_solve_device(cup2d_psb *h, const double *d_b, double *d_x, double tol_abs,
                           double tol_rel, int max_restarts, cup2d_ps_stats *stats);

/* Times short runs of the batched iteration for each schedule -- blocks per
 * work item (4 .. 64) and the XCD run length of the operator kernels' items
 * (plain order, or runs of 4 .. 32 consecutive items per XCD) -- on
 * right-hand side d_b, keeps the fastest (also for later handles of this
 * process with the same precision and batch shape) and reports its item
 * size and ms per iteration.  No counterpart in the reference. */
int cup2d_psb_autotune(cup2d_psb *h, const double *d_b, int *item_blocks, double *ms_per_iter);
/* Blocks per work item in use. */
int cup2d_psb_item_blocks(const cup2d_psb *h);
/* XCD run length of the operator kernels' items in use (1: plain order). */
int cup2d_psb_xcd_run(const cup2d_psb *h);

/* y = A x over every instance (device pointers, concatenated vectors; the
 * mean rows with mean_constraint), and a timing of `reps` back-to-back
 * launches of the operator kernel on the handle's stream (HIP events; mean
 * ms per launch).  Reference: BiCGSTABSolver::hd_cusparseSpMV
 * (BiCGSTAB.cu:315-410), applied per instance. */
int cup2d_psb_spmv_device(cup2d_psb *h, const double *d_x, double *d_y);
int cup2d_psb_spmv_time(cup2d_psb *h, const double *d_x, double *d_y, int reps,
                        double *ms_per_launch);
/* Times one of the batched iteration kernels alone with every instance
 * iterating: `reps` launches back to back on the handle's stream (HIP events,
 * mean ms per launch), each redoing one iteration's work from the same state
 * (set up from x = 0 and a fixed b; the handle's work vectors are
 * overwritten, the next solve starts afresh).  which: 0 p update +
 * preconditioner, 1 nu = A z1, 2 s update + preconditioner, 3 t = A z2,
 * 4 x and r update, 5 alpha, 6 omega, 7 state evaluation.  No counterpart in
 * the reference (DeviceProfiler.cuh:26-107 times the reference's phases). */
int cup2d_psb_kernel_time(cup2d_psb *h, int which, int reps, double *ms_per_launch);
/* d_y = d_x over the batch's unknowns with the library's streaming copy
 * kernel, timed over `reps` launches (HIP events, ms per launch): the
 * bandwidth ceiling of the solver's vector access pattern, and the kernel
 * that calibrates HBM traffic counters.  No counterpart in the reference (a
 * measurement utility). */
int cup2d_psb_copy_time(cup2d_psb *h, const double *d_x, double *d_y, int reps,
                        double *ms_per_launch);
/* Totals over the instances: unknowns, nonzeros, explicit (coarse-fine) rows. */
int cup2d_psb_sizes(const cup2d_psb *h, long long *n, long long *nnz, long long *n_explicit);

/* As cup2d_ps_last_error, for batch handles and cup2d_psb_create. */
const char *cup2d_psb_last_error(const cup2d_psb *h);

#ifdef __cplusplus
}
#endif

#endif /* CUP2D_POISSON_H */
