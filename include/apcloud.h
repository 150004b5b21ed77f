/*
 * apcloud.h — C ABI of the MI355X-native AP-Cloud solve-and-move path.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (APC_OK or a negative APC_ERR_* code; apc_last_error() gives the text).  No
 * exception crosses the ABI.  A context owns all device memory and one HIP
 * stream; caller pointers are read or written only during the call.  One host
 * thread per context.
 *
 * The reference has no C ABI: its "operator API" is the C++ class surface of
 * Octree / Solver / Mover / Controller.  Each function below names the
 * reference member (file:line) whose work it performs; the C++ classes in
 * ap-cloud-cosmology_amd/host/ keep the reference's signatures and call these.
 */
#ifndef APCLOUD_H
#define APCLOUD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  APC_OK = 0,
  APC_ERR_ARG = -1,          /* bad argument / state */
  APC_ERR_HIP = -2,          /* HIP runtime error */
  APC_ERR_NOMEM = -3,        /* device allocation failed */
  APC_ERR_NEIGH_MANY = -4,   /* solver.cpp:535-539 "Too many neighbours in an octant" */
  APC_ERR_NEIGH_FEW = -5,    /* solver.cpp:544-550 "Not enough neighbours" */
  APC_ERR_NEIGH_INVALID = -6,/* solver.cpp:511-518 "Invalid neighbour" */
  APC_ERR_RANK = -7,         /* solver.cpp:107-115,343-348 rank-deficient WLSQ */
  APC_ERR_BALANCE = -8,      /* 2:1 balance violated while emitting Neigh24 (internal) */
  APC_ERR_UNSUPPORTED = -9,  /* input outside what this build implements */
  APC_ERR_NO_DEVICE = -10
};

/* Parameters of one simulation box (the .param tags of parameter.cpp:45-149
 * that the path reads, plus the optional tags this build adds). */
typedef struct {
  double bbox[6];          /* BoundingBox_min_x, _max_x, _min_y, _max_y, _min_z, _max_z */
  int depth;               /* MaxOctreeLevel (1..21) */
  int bc;                  /* Periodic: 1 periodic, 0 non-periodic */
  double c;                /* CellSizeParameter */
  int order;               /* criterion order, reference mover.cpp:86 passes 2 */
  int mindepth;            /* reference mover.cpp:86 passes 0 */
  int maxdepth;            /* reference mover.cpp:86 passes MaxOctreeLevel */
  int leaf_capacity;       /* >0: open a cell iff n_p > leaf_capacity (added tag); 0: n*h^(2o-2) > c */
  int lm;                  /* NumberOfNeighbors */
  int dirichlet_point;     /* solver.cpp:17: pin row 0 of the Laplacian when periodic */
  int boundary_at_faces;   /* added tag: 0 = reference (+-1, octree.cpp:2772...), 1 = box faces */
  int bv_kind;             /* added tag: 0 = reference bv=0, 1 = uniform-sphere potential */
  double bv_mass, bv_radius, bv_center[3], bv_G;
} apc_params;

/* PETSc KSP options honoured by the GMRES (-ksp_rtol, -ksp_atol,
 * -ksp_divtol, -ksp_max_it, -ksp_gmres_restart, -ksp_monitor). */
typedef struct {
  double rtol, atol, dtol;
  int max_it, restart;
  int monitor;             /* print "%3d KSP Residual norm %14.12e" like -ksp_monitor */
  int precision;           /* 0: FP64 GMRES; 1: FP32 inner GMRES in FP64 iterative refinement */
  int pc_type;             /* 0: Jacobi, left (the reference's PCJACOBI, solver.cpp:201,273);
                              1: aggregation multigrid built on the device, right
                              preconditioning (what -pc_type gamg -ksp_pc_side right selects
                              through the reference's KSPSetFromOptions, solver.cpp:202,274);
                              convergence is then tested on the true residual */
} apc_ksp_opts;

typedef struct {
  int its_rho, reason_rho; double rnorm_rho;
  int its_phi, reason_phi; double rnorm_phi;
  int refine_rho, refine_phi;   /* outer refinement sweeps (precision form) */
  int nc, nv, nb;               /* sizes of the system that was solved */
  int fit_qr_rows;              /* least-squares rows fitted by Householder QR (condition
                                   estimate above the threshold, see apc_solver_set_fit_threshold) */
  double fit_max_cond;          /* largest condition estimate among the Cholesky-fitted rows */
} apc_solve_info;

/* Host view of the computational particles (particle.h:23-53), filled by
 * apc_get_comp.  Arrays are caller-allocated with the sizes from
 * apc_get_comp_sizes; NULL members are skipped. */
typedef struct {
  double *x, *y, *z, *cell_length;
  int *np, *first, *type, *first_neigh, *number_neigh, *neigh24;
  double *density, *potential;
  double *p_x, *p_y, *p_z, *p_xx, *p_yy, *p_zz, *p_xy, *p_xz, *p_yz;
} apc_comp_view;

typedef struct apc_ctx apc_ctx;

const char *apc_last_error(const apc_ctx *ctx);
const char *apc_version(void);
int apc_device_count(int *n);

/* Octree::Octree + Solver::Solver + Mover::Mover (octree.h:224,
 * solver.h:53, mover.h:31): allocate a context on the current HIP device. */
int apc_create(const apc_params *p, apc_ctx **out);
int apc_destroy(apc_ctx *ctx);

/* Launch-shape and cycle-shape knobs of this build's kernels, per context (no
 * counterpart in the reference).  Defaults below; an APC_* environment
 * variable of the same meaning (APC_NEIGH_XCHUNK, APC_MDOT_BLOCKS,
 * APC_MAXPY_BLOCKS, APC_SEG_BLOCKS, APC_AMG_G8, APC_AMG_G4, APC_AMG_RG8,
 * APC_AMG_NOPRE, APC_AMG_WLEVELS) overrides a default at apc_create; the
 * autotuner (apcloud.tune) sets them with apc_set_knob.  Every value >= 1 is
 * valid (>= 0 for APC_KNOB_AMG_NOPRE).  The launch-shape knobs change only
 * speed, never results beyond the summation order of the reductions; the two
 * multigrid cycle knobs change the preconditioner, so the Krylov step count
 * and the solution within the solve's tolerance. */
enum {
  APC_KNOB_SEARCH_CHUNK = 0, /* nodes per chunk the neighbour search deals to the XCDs (256) */
  APC_KNOB_MDOT_BLOCKS,      /* grid of the GMRES multi-dot (1024) */
  APC_KNOB_MAXPY_BLOCKS,     /* grid of the GMRES orthogonalisation update (1024) */
  APC_KNOB_SEG_BLOCKS,       /* grid of the packed boxes' tile kernels (2048) */
  APC_KNOB_AMG_G8_ROWS,      /* multigrid levels of at most this many rows: 8 lanes per row (65536) */
  APC_KNOB_AMG_G4_ROWS,      /* ... else of at most this many rows: 4 lanes per row (2^30) */
  APC_KNOB_AMG_RG8_ROWS,     /* restrictions into at most this many rows: 8 lanes (2^30) */
  APC_KNOB_AMG_NOPRE,        /* multigrid: finest levels without pre-smoothing (2) */
  APC_KNOB_AMG_WLEVELS,      /* multigrid: finest levels visited by a W-cycle (3) */
  APC_KNOB_COUNT
};
int apc_set_knob(apc_ctx *ctx, int knob, int value);
int apc_get_knob(apc_ctx *ctx, int knob, int *value);

/* Packed boxes: nbox (<= 64) independent periodic boxes of the same
 * parameters in one context.  No counterpart in the reference, which runs one
 * box per Controller (controller.cpp:16-219); a packed context runs every box
 * through each kernel of the path in one launch.  counts[b] particles of box
 * b are the rows pstart[b] .. pstart[b+1]-1 of apc_set_particles (pstart the
 * running sum of counts).  The box id sits above the Morton bits of the
 * 64-bit sort key, which lowers the depth to (64 - ceil(log2 nbox)) / 3 when
 * p->depth is deeper; apc_select then refuses (APC_ERR_UNSUPPORTED) a tree
 * that p->depth would refine beyond the lowered depth.  Computational rows
 * are box-major: box b's rows are row_start[b] .. row_start[b+1]-1 (see
 * apc_pack_rows), its computational nodes first, each box in the order a
 * context of that box alone gives.  Every box's potential system pins its
 * first row; the density right-hand side of box b carries mass[b]
 * (apc_pack_set_masses, default 1) times c_dens.  Each GMRES solve keeps one
 * Krylov space, restart cycle and convergence test per box, every box in
 * the same launches; the solve info reports the largest iteration count.
 * bc must be 1 (APC_ERR_UNSUPPORTED otherwise); apc_set_comp and
 * apc_freeze_topology refuse a packed context. */
int apc_create_packed(const apc_params *p, int nbox, const int *counts, apc_ctx **out);
int apc_pack_set_masses(apc_ctx *ctx, const double *mass);
/* Rows of the last selection (apc_select or apc_step): row_start nbox + 1
 * entries, comp_rows nbox entries (either may be NULL). */
int apc_pack_rows(apc_ctx *ctx, int *row_start, int *comp_rows);
/* Per box, the density and potential solves of the last apc_solver_solve /
 * apc_step on a packed context: iterations, KSP reasons and final residual
 * norms (nbox entries each; NULL pointers are skipped).  Each box's solve has
 * its own Krylov space and convergence test, so these are what a context of
 * that box alone reports (up to the summation order of the reductions). */
int apc_pack_solve_info(apc_ctx *ctx, int *its_rho, int *its_phi, int *reason_rho,
                        int *reason_phi, double *res_rho, double *res_phi);

/* initialization.cpp:41-91 hands the path a Physical_particle; this copies
 * the caller's SoA arrays (particle.h:5-21) to HBM. */
int apc_set_particles(apc_ctx *ctx, int n, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz);
/* Copies positions, velocities and accelerations back in the caller's
 * original particle order.  NULL pointers are skipped. */
int apc_get_particles(apc_ctx *ctx, double *x, double *y, double *z, double *vx, double *vy,
                      double *vz, double *ax, double *ay, double *az);

/* Octree::buildOctree (octree.cpp:111-494): Morton keys (computeKey64,
 * octree.cpp:676-852), stable device radix sort and the move of the particle
 * arrays into key order.  The sorted (key, original index) pairs are
 * readable with apc_get_sorted. */
int apc_octree_build(apc_ctx *ctx);
int apc_get_sorted(apc_ctx *ctx, uint64_t *keys, int *index);
/* The octree level arrays of buildOctree for one level: node count, and per
 * node the masked key, first sorted particle and particle count
 * (octree.cpp:177-313).  Pass NULL arrays to get the count only. */
int apc_octree_level(apc_ctx *ctx, int level, int *count, uint64_t *key, int *first, int *np);

/* Octree::select_computational_particles_2to1_3d (octree.cpp:1050-3001):
 * top-down refinement, 2:1 face balance and Neigh24.  *S_out receives
 * nc+nv+nb.  Output order: computational leaves in the reference's order,
 * vacuum leaves by (level, Morton), boundary nodes in leaf order. */
int apc_select(apc_ctx *ctx, double c, int order, int mindepth, int maxdepth, int *S_out);
int apc_get_comp_sizes(apc_ctx *ctx, int *nc, int *nv, int *nb, int *nneigh);
int apc_get_comp(apc_ctx *ctx, const apc_comp_view *view);

/* Solver::Solver over a caller-filled Computational_particle (solver.h:53,
 * particle.h:23-53): loads nc computational (Type 0), nv vacuum (Type -1) and
 * nb boundary (Type 1) nodes with their centres, CellLength,
 * NumberOfPhysicalParticles and Neigh24 (row i: neigh24[first_neigh[i] ..
 * + number_neigh[i]), at most 24 entries, nneigh entries in all) in place
 * of a selection; the solver entry points below then work on them.  Neigh24
 * entries are checked where the neighbour search reads them
 * (APC_ERR_NEIGH_INVALID, solver.cpp:511-518).  There are no physical
 * particles: apc_interp_kick refuses until the next apc_select. */
int apc_set_comp(apc_ctx *ctx, int nc, int nv, int nb, const double *x, const double *y,
                 const double *z, const double *cell_length, const int *np, const int *type,
                 const int *first_neigh, const int *number_neigh, const int *neigh24, int nneigh);

/* Solver::searchneighboroctant (solver.cpp:456-570) for every computational
 * and vacuum node: lm neighbour indices and distances, row-major (M x lm). */
int apc_neighbors(apc_ctx *ctx, int *nbr, double *dist);

/* Solver::Set_NumberOfGFDNeighbors (solver.h:56): lm >= 9. */
int apc_solver_set_neighbors(apc_ctx *ctx, int lm);
/* Solver::Set_Constant_For_{Density_Estimator,Potential_Solver,Field_Solver}
 * (solver.h:57-59). */
int apc_solver_set_constants(apc_ctx *ctx, double c_dens, double c_pot, double c_field);
/* Solver::Build_Linear_System (solver.cpp:24-180).  Optional export of the
 * rows (M x lm, neighbour order): cols, A1 and A2 off-diagonal values (both
 * 0 on boundary columns, whose A2 weights go into the boundary RHS term
 * pbnd), diagonals, rhs1 and pbnd. */
int apc_build_system(apc_ctx *ctx, int *cols, double *a1, double *a2, double *d1, double *d2,
                     double *rhs1, double *pbnd);
/* Solver::solve (solver.cpp:382-454): system build, Solve_Density,
 * Solve_Potential, Differentiate. */
int apc_solver_solve(apc_ctx *ctx, const apc_ksp_opts *rho, const apc_ksp_opts *phi,
                     apc_solve_info *info);
/* The least-squares fits of Build_Linear_System and Differentiate
 * (LAPACKE_dgels, solver.cpp:106,342).  Each row is fitted through the
 * Cholesky factor R of its Gram matrix (BᵀB = RᵀR, B the weighted basis) when
 * its condition estimate ||R||_F ||R^-1||_F (at most 9 cond_2(B)) is at most
 * cond_qr, and by Householder QR of B in FP64 (dgels' method: error growing
 * with cond(B), not its square) otherwise or when a Cholesky pivot fails.
 * Default cond_qr = 1e5; 0 sends every row to QR. */
int apc_solver_set_fit_threshold(apc_ctx *ctx, double cond_qr);
/* After a system build: per row the condition estimate (M values, infinity
 * where the Cholesky factorisation failed), the number of rows fitted by QR
 * and the largest estimate among the others.  Any pointer may be NULL. */
int apc_fit_conditions(apc_ctx *ctx, double *cond, int *qr_rows, double *max_cond);
/* Multigrid hierarchy of the last solve run with pc_type 1: *nlev levels, the
 * first min(cap, *nlev) row counts and stored nonzeros (level 0 = the GFD
 * matrix).  *nlev = 0 if no multigrid was built.  No reference counterpart
 * (PETSc's -ksp_view prints the same for gamg). */
int apc_solver_amg_levels(apc_ctx *ctx, int *rows, long long *nnz, int cap, int *nlev);
/* Level `level` >= 1 of that multigrid as CSR (rp: rows + 1, col and val:
 * nnz entries, sizes from apc_solver_amg_levels) and the aggregate map from
 * level `level` to the next (rows entries; level 0 included, the coarsest
 * excluded).  Any pointer may be NULL.  For inspection and tests. */
int apc_solver_amg_level(apc_ctx *ctx, int level, int *rp, int *col, double *val, int *agg);
/* Solver::Differentiate alone on a caller-supplied potential (M values). */
int apc_differentiate(apc_ctx *ctx, const double *phi);
/* Node fields after a solve: rho, phi (M), g (M x 3), H (M x 6: xx yy zz xy xz yz). */
int apc_get_fields(apc_ctx *ctx, double *rho, double *phi, double *g, double *H);
/* Overwrites g and H (same layout) — lets tests drive the mover alone. */
int apc_set_gradients(apc_ctx *ctx, const double *g, const double *H);

/* Mover::move_velocity (mover.cpp:51-89) with the interpolation of
 * octree.cpp:3003-3031: a = g + H dx, v += a Cv dt; *max_dv = max |a| Cv dt. */
int apc_interp_kick(apc_ctx *ctx, double Cv, double dt, double *max_dv);
/* Mover::move_position (mover.cpp:91-153): x = wrap(x + v Cx dt);
 * *max_dx = max |v| Cx dt. */
int apc_drift(apc_ctx *ctx, double Cx, double dt, double *max_dx);
/* max over computational particles of n_p / h^3 (controller.cpp:146-162). */
int apc_max_density(apc_ctx *ctx, double *max_density);

/* One whole solve-and-move step (controller.cpp:119-208 without output):
 * octree + selection, solve with the given coefficients, kick with Cv, drift
 * with Cx.  Reports max changes and, if timing != NULL, per-phase seconds
 * [select, solve, move]. */
int apc_step(apc_ctx *ctx, double c_dens, double c_pot, double c_field, double Cv, double Cx,
             double dt, const apc_ksp_opts *rho, const apc_ksp_opts *phi, double *max_dv,
             double *max_dx, apc_solve_info *info, double *timing);

/* ---- Multi-box form: many independent boxes stepped in one call ----
 * The reference steps one box per process (controller.cpp:106-208); a batch
 * holds one context per box (apc_batch_box gives it for the per-box entry
 * points above) and apc_batch_run steps every box through the same
 * coefficient schedule on a pool of host threads, largest box first.  Each
 * box's result equals stepping it alone with apc_step. */
typedef struct apc_batch apc_batch;
/* nthreads <= 0: min(nbox, 8). */
int apc_batch_create(const apc_params *p, int nbox, int nthreads, apc_batch **out);
int apc_batch_destroy(apc_batch *b);
int apc_batch_size(const apc_batch *b);
apc_ctx *apc_batch_box(apc_batch *b, int i);
const char *apc_batch_last_error(const apc_batch *b);
int apc_batch_set_particles(apc_batch *b, int i, int n, const double *x, const double *y,
                            const double *z, const double *vx, const double *vy, const double *vz);
/* coef: nsteps x 5 (c_pot, c_field, Cv, Cx, dt) per step, as apc_step;
 * c_dens: per box.  info (nbox, may be NULL) gets each box's last solve,
 * status (nbox, may be NULL) each box's return code; the first failing
 * box's code is returned and apc_batch_last_error names it. */
int apc_batch_run(apc_batch *b, int nsteps, const double *coef, const double *c_dens,
                  const apc_ksp_opts *rho, const apc_ksp_opts *phi, apc_solve_info *info,
                  int *status);

/* ---- Differentiation form: adjoints with the topology held fixed ----
 * No reference counterpart (the reference has no derivatives); used by
 * apcloud.adjoint (torch.autograd.Function) and its finite-difference tests.
 * After apc_freeze_topology(ctx, 1) the tree, leaves, neighbour lists, GFD
 * systems and the particle-to-leaf map stay as built; apc_step then only
 * solves (fields), kicks and drifts. */
int apc_freeze_topology(apc_ctx *ctx, int on);
/* Particle state in caller order into the frozen device order. */
int apc_set_state(apc_ctx *ctx, const double *x, const double *y, const double *z,
                  const double *vx, const double *vy, const double *vz);
/* Save (restore = 0) or restore device-side copies of (x, v), slot 0..63.
 * Outside the frozen form the slot also keeps the particle order (caller
 * indices), and a restore drops the tree built since: the next apc_step or
 * apc_octree_build starts from the saved state, as if it had been uploaded. */
int apc_checkpoint(apc_ctx *ctx, int slot, int restore);
/* Loss = sum over a periodic n^3 grid of (CIC density / mean - 1)^2 of the
 * current positions; seeds the adjoint state with lambda_x = dL/dx,
 * lambda_v = 0.  The deposit is sorted by cell and summed with compensation,
 * so the loss is reproducible bit for bit. */
int apc_adjoint_seed_cic(apc_ctx *ctx, int n, double *loss);
/* Per-particle weights of the CIC loss's density (caller order, >= 0, not
 * all zero; the mean is then the weight sum); NULL restores equal weights. */
int apc_adjoint_cic_weights(apc_ctx *ctx, const double *w);
/* Adjoint state (lambda_x, lambda_v) in caller order. */
int apc_adjoint_get(apc_ctx *ctx, double *lx, double *ly, double *lz, double *lvx, double *lvy,
                    double *lvz);
int apc_adjoint_set(apc_ctx *ctx, const double *lx, const double *ly, const double *lz,
                    const double *lvx, const double *lvy, const double *lvz);
/* Adjoint of the drift (mover.cpp:91-153): lambda_v += Cx dt lambda_x. */
int apc_adjoint_drift(apc_ctx *ctx, double Cx, double dt);
/* Adjoint of the kick with the interpolation (mover.cpp:51-89,
 * octree.cpp:3003-3031) using the node fields of the last solve:
 * lambda_x += Cv dt H lambda_v; acc != 0 also accumulates the leaf-field
 * adjoints dg, dH (cleared first when clear != 0). */
int apc_adjoint_kick(apc_ctx *ctx, double Cv, double dt, int acc, int clear);
/* Accumulated leaf-field adjoints, layouts as apc_get_fields (M x 3, M x 6). */
int apc_adjoint_fields(apc_ctx *ctx, double *dg, double *dH);
/* Transpose of Differentiate (solver.cpp:295-362): dphi = J^T (dg, dH). */
int apc_gradient_adjoint(apc_ctx *ctx, const double *dg, const double *dH, double *dphi);
/* out = A^-T rhs, A the density (which = 1) or Poisson (which = 2) matrix of
 * the last solve (solver.cpp:182-293), by FP64 Jacobi GMRES on the explicit
 * transpose. */
int apc_solve_adjoint(apc_ctx *ctx, int which, const apc_ksp_opts *opts, const double *rhs,
                      double *out, int *its, int *reason);

/* ---- No allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the Morton keys and stable key sort of apc_octree_build
 * (computeKey64 octree.cpp:676-852, the sort of buildOctree octree.cpp:111-
 * 494), read back by apc_get_sorted, on caller device arrays: x, y, z (n
 * doubles each) in, keys (n) and index (n, the caller position of each
 * sorted entry) out, bbox (6 doubles, host memory) and depth (1..21) as in
 * apc_params.  Launches on `stream` (a hipStream_t; NULL = the default
 * stream) and returns without waiting.  Its workspace is a fixed device
 * array of the library, about 25 MB per GPU, held until the process ends:
 * n <= APC_NOALLOC_MAX_N, larger n is rejected with APC_ERR_ARG
 * (apc_last_error(NULL) gives the text).  Every call on one GPU shares that
 * array: calls that overlap on one GPU race. */
#define APC_NOALLOC_MAX_N (1 << 21)
int apc_morton_sort_noalloc(int n, const double *x, const double *y, const double *z,
                            const double *bbox, int depth, uint64_t *keys, int *index,
                            void *stream);

/* Per-kernel device time of the timed calls (HIP events on the context's
 * stream): average ms and algorithmic bytes per launch, launch counts; names
 * are NUL-separated.  Returns the number of kernels. */
int apc_kernel_times(apc_ctx *ctx, char *names, int names_len, double *avg_ms, int *launches,
                     double *avg_bytes, int cap);
/* Copies `bytes` bytes device-to-device with a plain streaming kernel `reps`
 * times and reports the achieved GB/s (read + write), the measured HBM
 * reference beside the 8 TB/s datasheet peak. */
int apc_stream_copy(apc_ctx *ctx, double bytes, int reps, double *gbs);
int apc_reset_kernel_times(apc_ctx *ctx);
int apc_set_timing(apc_ctx *ctx, int on);
int apc_sync(apc_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
