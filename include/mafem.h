/*
 * mafem.h -- C ABI of the MI355X (gfx950) implementation of MembraneAleFem.jl's
 * per-time-step hot path: element residual/tangent assembly and the Newton
 * sparse linear solve.
 *
 * All arrays crossing this boundary use the reference's own layouts: Julia
 * column-major matrices of Int64 (1-based indices) and Float64, so a Julia
 * caller passes `mesh.IX`, `mesh.ID`, `xms`, `cps`, ... by pointer with no
 * conversion.  Plain pointers and sizes only.
 *
 * Return codes: MAFEM_OK (0), MAFEM_NOT_CONVERGED (1: Newton did not reach
 * the tolerance, the reference raises AssertionError), negative on error
 * (invalid argument, HIP runtime error, device memory exhausted, unsupported
 * configuration).  mafem_last_error() describes the most recent failure.
 *
 * Threading: one context per host thread at a time; calls on a context are
 * blocking and not re-entrant.
 */
#ifndef MAFEM_H
#define MAFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MAFEM_OK 0
#define MAFEM_NOT_CONVERGED 1
#define MAFEM_EINVAL (-1)
#define MAFEM_EHIP (-2)
#define MAFEM_ENOMEM (-3)
#define MAFEM_EUNSUPPORTED (-4)

typedef struct mafem_ctx mafem_ctx;

/* Mesh tables as the reference's `Mesh` struct holds them
 * (src/input/Mesh.jl:48-77).  Boundary arrays are indexed by the Boundary
 * enum minus one: BOTTOM, RIGHT, TOP, LEFT (src/input/Enums.jl:89-94). */
typedef struct {
  int64_t num1el, num2el, numel;   /* Mesh.jl:50-52 */
  int64_t num1np, num2np, numnp;   /* Mesh.jl:53-55 */
  int64_t ndf, nmdf;               /* Mesh.jl:71,74 */
  const int64_t *IX;               /* NEN x numel, 1-based node ids (Mesh.jl:56) */
  const int64_t *ID;               /* ndf x numnp, 1-based equation ids, 0 = prescribed (Mesh.jl:72) */
  /* mesh.dofs as column indices of cps (get_v_order/get_m_order/get_λ_order/
   * get_p_order, Mesh.jl:477-521); 0 = unknown absent */
  int32_t v_order[3], m_order[3], lam_order, p_order;
  /* area basis tables (mesh.area_gp_fns, GpBasisFn.jl:300-356): ufns is the
   * Matrix{GpBasisFnsζα}(n_uel, 9) in Julia memory order, each record 55
   * Float64 = w, N[9], ∂Nα[9x2], ∂∂Nαβ[9x3] (GpBasisFn.jl:96-112) */
  int64_t area_n_uel;
  const int64_t *area_uel_ids;     /* numel, 1-based */
  const double *area_ufns;         /* 55 * n_uel * 9 */
  /* boundary elements and basis tables (mesh.bdry_elems, mesh.bdry_gp_fns,
   * Mesh.jl:57,66; GpBasisFn.jl:231-277): ufns is Matrix(n_uel, 3) */
  int64_t bdry_n_el[4];
  const int64_t *bdry_elems[4];    /* 1-based area element ids */
  int64_t bdry_n_uel[4];
  const int64_t *bdry_uel_ids[4];  /* bdry_n_el[b] entries, 1-based */
  const double *bdry_ufns[4];      /* 55 * n_uel * 3 */
  /* mesh.inh_neu_bcs (Mesh.jl:77): (Boundary, Neumann, value) triples */
  int64_t n_neu;
  const int32_t *neu_bdry;         /* Boundary enum, 1..4 */
  const int32_t *neu_type;         /* Neumann enum: SHEAR 1, STRETCH 2, MOMENT 3 */
  const double *neu_val;
  /* mesh.crnr_gp_fns (Mesh.jl:228-233): 4 records of 55 Float64 in Corner
   * order BOTTOM_LEFT, BOTTOM_RIGHT, TOP_LEFT, TOP_RIGHT (Enums.jl:109-114);
   * NULL makes mafem_generate_output unavailable */
  const double *crnr_ufns;
} mafem_mesh_desc;

/* The hot-path subset of `Params` (src/input/Params.jl:36-55) plus the
 * `bend_tm` keyword argument used by MOMENT edges (FiniteElement.jl:379). */
typedef struct {
  int32_t motion;                  /* Motion enum: STATIC 1, EUL 2, LAG 3, ALEV 4, ALEVB 5 */
  int32_t scenario;                /* Scenario enum: F_CAVI 1 .. F_BEND 5 */
  double kb, kg, zv, pn, adb, am;  /* kb, kg, ζv, pn, αdb, αm */
  double ek, enr;                  /* εk (unused: the tangent is exact), εnr */
  double bend_tm;
} mafem_params;

/* Linear solver of the Newton step.
 * MAFEM_SOLVER_DIRECT: band LU of the current tangent every Newton iteration,
 *   then FP64 iterative refinement (the analogue of the reference's sparse LU).
 * MAFEM_SOLVER_KRYLOV: FGMRES(restart) in FP64 against the current tangent,
 *   preconditioned by a band LU that is kept across Newton iterations and time
 *   steps and refactored only when a solve needed more than refactor_iters
 *   iterations; factor_precision 32 stores and factors it in FP32 (the
 *   mixed-precision path: FP32 preconditioner inside FP64 residuals, update
 *   and convergence test). */
#define MAFEM_SOLVER_DIRECT 0
#define MAFEM_SOLVER_KRYLOV 1
typedef struct {
  int32_t mode;              /* MAFEM_SOLVER_DIRECT or MAFEM_SOLVER_KRYLOV */
  int32_t factor_precision;  /* 64 or 32 */
  int32_t restart;           /* Krylov restart length, 1..64 */
  int32_t max_iters;         /* Krylov iterations per linear solve */
  int32_t refactor_iters;    /* refactor when a solve used more iterations than this */
  int32_t refine_sweeps;     /* DIRECT: maximum refinement sweeps */
  double rtol;               /* target |K du + r| / |r| (FP64) */
} mafem_solver_opts;

/* Per-phase device timings (microseconds) of the last mafem_time_step or
 * mafem_assemble call, measured with HIP events on the context's stream. */
typedef struct {
  double assemble_us;              /* element + edge kernels and the gathers */
  double area_kernel_us;           /* the area-element kernel alone */
  double factor_us, solve_us, update_us, total_us;
  int32_t newton_iters, refine_sweeps;
  int32_t n_assemble, n_band_solve;  /* launches of the assembly / triangular-solve sweeps */
  int32_t n_factor, gmres_iters;     /* band factorizations / Krylov iterations */
  double step_kernel_us;           /* mean duration of one factorization-step kernel
                                      (trailing update, next diagonal-tile LU and next
                                      panel; band: the middle step of each factorization,
                                      nested dissection: every step launch of every
                                      level), first block start to last block end on the
                                      device's constant clock */
  double last_relres;              /* |K du + r| / |r| of the last solve */
  int32_t sweep_groups;            /* workgroups of each triangular sweep (autotuned) */
  int32_t pivot_perturbations;     /* diagonal-tile pivots below eps max|K| (numerically
                                      zero) replaced by +-eps max|K|: static pivoting,
                                      switched on for a context once a factorization
                                      without it gave a failed solve */
  int32_t factor_chains;           /* 2: two-ended band LU (top part downward and bottom
                                      part upward at once, separator last); 1: one chain */
  int32_t separator_rows;          /* 64-row tile rows of the separator (0 with one chain) */
  int32_t fronts;                  /* MAFEM_ORDER_ND: fronts of the elimination tree (else 0) */
  int32_t front_steps;             /* MAFEM_ORDER_ND: dependent tile steps of one
                                      factorization (sum over tree levels of the
                                      largest front's fully-summed tile rows) */
  double front_flops;              /* MAFEM_ORDER_ND: algorithmic FLOPs of one
                                      factorization (tile products 2*64^3, diagonal
                                      LUs 2/3*64^3, panel solves 64^3; else 0) */
  double step_flops;               /* algorithmic FLOPs of the launches step_kernel_us
                                      averages, per launch (same operation counts) */
  double step_bytes;               /* MAFEM_ORDER_ND: their algorithmic bytes per launch
                                      (tiles read and written once; else 0) */
  double schur_kernel_us;          /* MAFEM_ORDER_ND: mean duration of one tree level's
                                      Schur-block launch (every front's boundary block
                                      minus its fully-summed products), first block
                                      start to last block end on the device's constant
                                      clock (else 0) */
  double schur_flops;              /* algorithmic FLOPs per such launch (per front
                                      (FT - fT)^2 fT tile products of 2*64^3) */
  double schur_bytes;              /* algorithmic bytes per such launch (each operand
                                      tile read once, each output tile read and written
                                      once) */
} mafem_stats;

/* Builds the device context: uploads the tables, builds the block-sparse
 * pattern, the deterministic gather lists and the band layout once per mesh.
 * Replaces the per-call set-up of calc_r_K (FiniteElement.jl:85-95). */
int mafem_create(mafem_ctx **ctx, const mafem_mesh_desc *mesh, const mafem_params *p);
void mafem_destroy(mafem_ctx *ctx);
/* Text of the last error on ctx (or of the last failed mafem_create when ctx
 * is NULL). */
const char *mafem_last_error(const mafem_ctx *ctx);

/* Device-resident state: xms is numnp x 3, cps numnp x ndf, column-major,
 * exactly the reference's `xms::Matrix{Float64}`, `cps::Matrix{Float64}`. */
int mafem_set_state(mafem_ctx *ctx, const double *xms, const double *cps);
int mafem_get_state(mafem_ctx *ctx, double *xms, double *cps);

/* Sets prescribed values in the device state: cps[node, dof] = val for n
 * (1-based node, 1-based dof column) triples, as prepare_input applies
 * mesh.inh_dir_bcs (src/Input.jl:103-108). */
int mafem_set_prescribed(mafem_ctx *ctx, int64_t n, const int64_t *node,
                         const int32_t *dof, const double *val);

/* Replaces the equation numbering (the reference's `mesh.ID`, ndf x numnp,
 * column-major, 1-based, 0 = prescribed; Mesh.jl:72, built by
 * generate_scenario, Mesh.jl:262-302) of a context whose mesh is otherwise
 * unchanged, e.g. when the set of prescribed unknowns moves with a tether
 * that translates across the patch.  The reference rebuilds its Mesh for a
 * new Dirichlet set (restart with new arguments, src/MembraneAleFem.jl:75-107);
 * the state stays on the device, prescribed values are set with
 * mafem_set_prescribed.  Not while the context belongs to a batch. */
int mafem_set_dof_ids(mafem_ctx *ctx, const int64_t *ID, int64_t nmdf);

/* update_xms!(motion, xms, cps, Δt, dofs): the predictor x += Δt u_motion
 * applied before each step (FiniteElement.jl:408-423, Analysis.jl:70). */
int mafem_predict(mafem_ctx *ctx, double dt);

/* Number of stored entries of the exported tangent (the structural superset
 * of the reference's value-dependent pattern: every active DOF pair of two
 * nodes sharing an element). */
int mafem_nnz(mafem_ctx *ctx, int64_t *nnz);

/* calc_r_K (FiniteElement.jl:75-200) on the device state: writes r_gl
 * (nmdf) and K_gl as 1-based CSC (colptr nmdf+1, rowval/nzval nnz).  Any
 * output pointer may be NULL to skip it. */
int mafem_assemble(mafem_ctx *ctx, double time, double dt, double *r,
                   int64_t *colptr, int64_t *rowval, double *nzval);

/* Solves K du = rhs with the tangent of the last mafem_assemble (direct band
 * LU on the device plus FP64 refinement); rhs, du have nmdf entries.
 * Replaces `K_gl \ r_gl` (FiniteElement.jl:38). */
int mafem_solve(mafem_ctx *ctx, const double *rhs, double *du, double *relres);

/* time_step! (FiniteElement.jl:11-63): Newton on the device state, at most
 * max_iter iterations (the reference allows 14); eps_hist receives
 * |Δu|_2 / nmdf per iteration (may be NULL), iters the iteration count.
 * Returns MAFEM_NOT_CONVERGED if the last eps >= εnr. */
int mafem_time_step(mafem_ctx *ctx, double time, double dt, int32_t max_iter,
                    int32_t *iters, double *eps_hist);

int mafem_get_stats(const mafem_ctx *ctx, mafem_stats *out);

/* Pull force of the device state (calc_pull_force, src/analysis/PullForce.jl:
 * 61-80, with the adjacent elements and node maps of get_adj_maps,
 * PullForce.jl:25-53, for poly = 2) and the mean position of the pulled
 * element's 9 control points (src/Analysis.jl:82-84): f_pull[3] = sum over
 * the pulled element's nodes of the velocity rows of the area-element
 * residuals of the 25 elements around it (taken by element id, as the
 * reference does), x_pull[3].  The pulled element is ceil(numel / 2)
 * (get_pull_el_id, PullForce.jl:8-13).  MAFEM_EINVAL if an adjacent id falls
 * outside the mesh (the reference would raise a BoundsError). */
int mafem_pull_force(mafem_ctx *ctx, double *x_pull, double *f_pull);

/* mafem_pull_force without allocating device memory.  No counterpart in the
 * reference: for callers that cannot allocate (code recorded into HIP graphs,
 * tight serving loops).  mafem_pull_force keeps the adjacent-element maps and
 * residuals in buffers it allocates on its first call; this variant keeps
 * them in a 64 KiB array the library holds on each device until the process
 * ends.  Every call on a device shares that array (its contents are those of
 * the last call): calls that overlap on one device race.  Results equal
 * mafem_pull_force's bit for bit.  MAFEM_EUNSUPPORTED if the adjacent
 * elements' residuals would not fit (at most 25 elements of 72 rows, 15 KB,
 * for the reference's biquadratic meshes). */
int mafem_pull_force_noalloc(mafem_ctx *ctx, double *x_pull, double *f_pull);

/* generate_output(mesh, xms, cps) (src/Output.jl:30-118) of the device
 * state: the surface position and the unknowns sampled at every area Gauss
 * point, at the 3 Gauss points of each boundary element on the four edges and
 * at the four corners, on a (3 num1el + 2) x (3 num2el + 2) grid.  xout is
 * that grid x 3 and uout that grid x ndf, both column-major like the
 * reference's Array{Float64,3}; either may be NULL.  MAFEM_EINVAL when the
 * corner basis tables were not given at create time. */
int mafem_generate_output(mafem_ctx *ctx, double *xout, double *uout);

/* Selects the linear solver (no counterpart in the reference, which always
 * uses a sparse direct solve: this chooses how FiniteElement.jl:38 is done).
 * Defaults: DIRECT, FP64, restart 30, 200 iterations, refactor above 20,
 * 3 refinement sweeps, rtol 1e-13. */
int mafem_set_solver(mafem_ctx *ctx, const mafem_solver_opts *opts);

/* Fill-reducing ordering of the context's factorization (both solver modes).
 * Default at mafem_create: MAFEM_ORDER_ND from about 16x16 elements up (where
 * the band would run two-ended), MAFEM_ORDER_BAND below; the environment
 * variable MAFEM_ORDERING=band|nd at mafem_create overrides it.
 * MAFEM_ORDER_BAND: natural node order, band LU (two-ended on long patches),
 *   the tile rows one dependency chain.
 * MAFEM_ORDER_ND: nested dissection of the node grid (separators of two node
 *   lines, recursively), factored as a multifrontal LU: the fronts of each
 *   level of the elimination tree together, from the leaves up, each a dense
 *   tile matrix with partial pivoting inside its 64-row diagonal tiles.
 *   Fewer operations and a far shorter dependency chain than the band (64x64
 *   ALEVB: 48 dependent tile steps instead of 281), at about twice the
 *   band's factor storage.
 * This is the role of UMFPACK's fill-reducing ordering inside `K_gl \ r_gl`
 * (src/analysis/FiniteElement.jl:38); results agree with the band's to the
 * solve tolerance, not bit for bit.  Takes effect at the next factorization. */
#define MAFEM_ORDER_BAND 0
#define MAFEM_ORDER_ND 1
int mafem_set_ordering(mafem_ctx *ctx, int32_t ordering);
int mafem_get_solver(const mafem_ctx *ctx, mafem_solver_opts *opts);

/* Free and total memory of the current device, in bytes (no counterpart in
 * the reference: used to admit work by HBM budget). */
int mafem_device_memory(size_t *free_bytes, size_t *total_bytes);

/* ------------------------------------------------------------------ batches
 *
 * No counterpart in the reference (it advances one mesh per `solve` call,
 * src/MembraneAleFem.jl:44-60).  A batch advances many independent contexts
 * (patches of any size and formulation) through one time step together:
 * each phase is one launch covering every patch, and the Newton iteration's
 * launch sequence is recorded once in HIP graphs and replayed.  The Newton
 * step of each patch keeps the reference's definition (assembly, solve,
 * update, ε = |Δu|_2 / nmdf < εnr, at most max_iter iterations;
 * FiniteElement.jl:11-63); its linear solve is FP64 GMRES(m), right-
 * preconditioned by a band LU stored in `factor_precision` (32: the
 * mixed-precision path) and kept across iterations and steps; a patch is
 * refactored when a solve needs more than `refactor_iters` Arnoldi steps or
 * does not converge in `max_cycles` restart cycles, and a patch whose fresh
 * FP32 factorization cannot precondition the solve moves to FP64 storage.  The batch borrows
 * its members: they must outlive it, and while they belong to it only their
 * state calls (set_state, get_state, set_prescribed) may be used on them. */
typedef struct mafem_batch mafem_batch;

typedef struct {
  int32_t factor_precision;  /* 32 (default) or 64 */
  int32_t restart;           /* GMRES basis size m (default 20, at most 64) */
  int32_t max_cycles;        /* restart cycles per solve before refactoring (default 1); a
                                failed solve moves its member up one rung: stale factors ->
                                fresh -> FP64 storage -> static pivoting, where it keeps
                                cycling until the solve has taken 400 Arnoldi steps */
  int32_t refactor_iters;    /* refactor when a solve needed more Arnoldi steps (default 6;
                                -1: factor afresh for every Newton iteration) */
  int32_t use_graphs;        /* 1 (default): replay recorded HIP graphs */
  int32_t escalate_iters;    /* move a patch to FP64 when a solve on a fresh FP32
                                factorization needed more Arnoldi steps (default 12) */
  double rtol;               /* target |b - K x| / |b| of each solve (default 1e-10) */
} mafem_batch_opts;

typedef struct {
  double total_us, assemble_us, factor_us, solve_us, update_us;  /* last time step */
  int32_t newton_iters;      /* largest per-patch Newton iteration count */
  int32_t factor_passes;     /* batched factorization passes */
  int32_t gmres_cycles;      /* batched GMRES restart cycles */
  int32_t hard_failures;     /* solves left unconverged on fresh, statically pivoted FP64
                                factors after 400 Arnoldi steps (the correction is kept) */
  int64_t patch_iters;       /* sum over patches of Newton iterations */
  int64_t patch_factors;     /* sum over patches of factorizations */
  int64_t gmres_iters;       /* sum over patches of Arnoldi steps */
  int64_t elements;          /* sum over patches of area elements */
  int32_t fp64_patches;      /* members whose band LU is stored in FP64 */
  int32_t reserved;
} mafem_batch_stats;

/* opts may be NULL (defaults).  All members must live on the current device. */
int mafem_batch_create(mafem_batch **out, int32_t n, mafem_ctx *const *members,
                       const mafem_batch_opts *opts);
void mafem_batch_destroy(mafem_batch *b);
const char *mafem_batch_last_error(const mafem_batch *b);
/* One backward-Euler step of the members selected by `active` (NULL: all):
 * the predictor update_xms! when `predict` is nonzero (Analysis.jl:70), then
 * Newton.  time and dt hold one value per member; iters[n], eps_hist[n *
 * max_iter] (row per member) and status[n] (0 converged, 1 not converged,
 * 2 not selected) may be NULL.  Members not selected are left untouched.
 * Returns 1 if any selected member did not converge. */
int mafem_batch_time_step(mafem_batch *b, const double *time, const double *dt,
                          const uint8_t *active, int32_t predict, int32_t max_iter, int32_t *iters,
                          double *eps_hist, int32_t *status);
int mafem_batch_get_stats(const mafem_batch *b, mafem_batch_stats *out);
/* mafem_pull_force for every member, ordered after the batch's last step:
 * out[6 p .. 6 p + 5] = x_pull[3], f_pull[3] of member p (no counterpart in
 * the reference, which evaluates calc_pull_force per mesh,
 * src/Analysis.jl:81-90). */
int mafem_batch_pull_force(mafem_batch *b, double *out);

#ifdef __cplusplus
}
#endif

#endif /* MAFEM_H */
