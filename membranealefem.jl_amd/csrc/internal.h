// Internal declarations shared by the HIP translation units of libmafem.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "launch_limits.h"
#include "mafem.h"

namespace mafem {

enum Motion { STATIC = 1, EUL = 2, LAG = 3, ALEV = 4, ALEVB = 5 };
enum Boundary { BOTTOM = 1, RIGHT = 2, TOP = 3, LEFT = 4 };
enum Neumann { SHEAR = 1, STRETCH = 2, MOMENT = 3 };
enum Scenario { F_CAVI = 1, F_COUE = 2, F_POIS = 3, F_PULL = 4, F_BEND = 5 };

constexpr int NEN = 9;
constexpr int REC = 55;      // doubles per tabulated 2-D basis record
constexpr int RECP = 56;     // padded record stride on the device
constexpr int TILE = 64;     // band-LU tile edge
constexpr int SOLVE_RING = 72;  // solution tiles a sweep keeps in LDS (needs tb + 1 <= this)
// Lanes per tangent column of the area-element kernel, by Motion (STATIC 1 ..
// ALEVB 5; index 0 unused): each lane accumulates the rows of ceil(9 / parts)
// nodes.  With the Gauss points' values shared through LDS one lane per column
// fits the registers without spills and is the fastest split for every motion
// at 64x64 (ALEVB 338 us against 463 us with 3 and 918 us with 9).
constexpr int AREA_PARTS[6] = {1, 1, 1, 1, 1, 1};
inline int area_parts(int motion) { return motion >= 1 && motion <= 5 ? AREA_PARTS[motion] : 1; }

// Device-side physical parameters (by value into kernels).
struct PhysParams {
  double kb, kg, zv, pn, adb, am;
  double bend_tm;
  int32_t scenario;
};

// Per-edge-element work item: area element, unique-basis slot, boundary,
// Neumann type and value.
struct EdgeItem {
  int32_t el;       // 0-based area element
  int32_t rec;      // 0-based record offset (in records) of its 3 Gauss points
  int32_t bdry;     // Boundary enum
  int32_t ntype;    // Neumann enum
  double nval;
};

struct NdDev;  // nested-dissection fronts (band.hip)

struct Ctx {
  std::string err;
  int dev = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[12] = {};   // 0-6 phases, 8-9 area kernel

  // sizes
  int64_t num1el = 0, num2el = 0, numel = 0, num1np = 0, num2np = 0, numnp = 0;
  int ndf = 0, ne = 0;          // ne = 9 ndf
  int64_t nmdf = 0;
  int motion = 0;
  PhysParams phys{};
  double enr = 0;
  int32_t v_order[3], m_order[3], lam_order, p_order, mm_order[3];

  // state (column-major like the host arrays)
  double *d_xms = nullptr, *d_cps = nullptr;

  // mesh tables
  int32_t *d_ix = nullptr;       // 9 x numel, 0-based nodes
  int32_t *d_uel = nullptr;      // numel, 0-based unique element
  uint8_t *d_active = nullptr;   // ndf x numnp (1 = unknown, 0 = prescribed)
  double *d_ufns = nullptr;      // area records [uel][gp][RECP]
  double *d_dbt = nullptr;       // Dohrmann--Bochev G^T H^-1 G per unique element [uel][9][9]
  double *d_bufns = nullptr;     // boundary records, 3 per unique boundary element
  EdgeItem *d_edges = nullptr;
  int64_t n_edges = 0;

  // element buffers: K_el row-major [elem][row][col], r_el [elem][row];
  // edge elements follow the area elements
  double *d_kel = nullptr, *d_rel = nullptr;

  // block-sparse (node blocks ndf x ndf, row-major blocks) tangent of the
  // extended system (prescribed DOFs -> identity rows/columns)
  int64_t nnzb = 0;
  int32_t *d_rowptr = nullptr, *d_colidx = nullptr, *d_brow = nullptr;
  double *d_bsr = nullptr;
  // gather lists: for BSR block k, entries gptr[k]..gptr[k+1] of base offsets
  // into d_kel (element e, local nodes a, b) -> e ne^2 + a ndf ne + b ndf
  int64_t *d_gptr = nullptr, *d_gsrc = nullptr;
  // residual gather: for node p, entries rptr[p]..rptr[p+1] of e ne + a ndf
  int64_t *d_rptr = nullptr, *d_rsrc = nullptr;

  // extended vectors (numnp ndf)
  int64_t nx = 0;
  double *d_r = nullptr, *d_x = nullptr, *d_w = nullptr, *d_t = nullptr, *d_b = nullptr;
  double *d_t2 = nullptr;        // band-solve intermediate
  double *d_red = nullptr;       // reduction scratch
  double *h_red = nullptr;       // pinned host mirror of reduction results

  // CSC export (1-based) over active DOFs, built on first use (nnz < 0 before)
  std::vector<int64_t> h_ID, h_eq;
  std::vector<int32_t> h_rowptr;
  int64_t nnz = -1;
  std::vector<int64_t> h_colptr, h_rowval;
  int64_t *d_csc_src = nullptr;  // BSR value index of each CSC entry
  double *d_csc_val = nullptr;
  int64_t *d_eq = nullptr;       // extended index of each equation (nmdf)

  // band LU (storage precision band_prec: 64 or 32; allocated on first use)
  int64_t nt = 0;                // tile rows
  int tb = 0;                    // tile half bandwidth
  void *d_band = nullptr;        // [nt][2tb+1][TILE][TILE] row-major tiles
  void *d_linv = nullptr, *d_uinv = nullptr;  // [nt][TILE][TILE]
  int32_t *d_piv = nullptr;      // [nt][TILE] row permutation of each tile row
  int64_t band_bytes = 0;
  int band_prec = 0;
  bool have_factor = false;      // the band holds a factorization usable as preconditioner
  int32_t n_factor = 0;          // running count of factorizations
  bool step_timed = false;       // d_span holds a timed factorization-step launch not yet read
  unsigned long long *d_span = nullptr;  // its first block start, last block end (constant clock)
  double wall_khz = 1e5;         // frequency of that clock

  // linear solver options and Krylov workspace
  mafem_solver_opts opts{};
  double *d_V = nullptr, *d_Z = nullptr;   // (m+1) and m basis vectors of length nt TILE
  double *d_h = nullptr;                   // multi-dot results (device)
  double *d_coef = nullptr;                // combination weights (device)
  double *d_mdot_part = nullptr;           // [65][256] partial dots
  int kry_m = 0;
  int32_t n_gmres = 0;           // running count of GMRES iterations

  mafem_stats stats{};
  int32_t n_band_solve = 0;     // running count of triangular-solve sweeps

  // factorization left in the band by a batch (kept when the context joins
  // another batch): its precision, 0 = none; batch_prec = precision a batch
  // last chose for this context (an FP64 fallback stays FP64)
  int batch_factor_prec = 0;
  int batch_prec = 0;

  // triangular sweeps split over sweep_groups workgroups (1: one workgroup)
  int sweep_groups = 1;
  bool sweep_tuned = false;      // the autotuner has chosen sweep_groups
  double *d_s1 = nullptr, *d_s2 = nullptr;  // sweep outputs (sentinel-filled, atomic access only)
  int *d_sweep_err = nullptr;    // count of sweep waits that ran out (must stay 0)
  // static pivoting of the single-context factorization: d_amax[0] = max |K|
  // (as the bits of a non-negative double), d_amax[1] = count of pivots
  // replaced by +-sqrt(eps) max |K|
  unsigned long long *d_amax = nullptr;
  bool perturb_pivots = false;   // set once a factorization without it failed

  // Two-ended band LU of the single-context solve (tw_a > 0).  The tile rows
  // split into a top part T = [0, tw_a), a separator S of tb rows and a bottom
  // part B of tw_nb rows; T and B do not couple.  T is factored downward in
  // the band BT (rows T and S, in original order), B upward in the band BR
  // (rows B then S, both in reversed order: tile row I' = nt - 1 - I, element
  // (i, j) -> (TILE-1-i, TILE-1-j)), step k of both chains in the same
  // launches; both add their Schur complements into S, which BT factors last.
  // The solve runs both chains' sweeps in the same launches too.  Storage:
  // BT at tile row 0 of d_band / d_linv / d_uinv / d_piv, BR at tile row
  // tw_a + tb.  tw_a = 0: one band factored top to bottom.
  int64_t tw_a = 0, tw_nb = 0;
  // look-ahead staging of the factorization: per chain (top, bottom) two
  // buffers of TILE^2 + TILE words, sentinel-filled before each chain runs
  unsigned long long *d_pub = nullptr;
  double *d_tw_z = nullptr;      // separator right-hand side (original order, nt TILE)
  double *d_tw_br = nullptr;     // right-hand side in reversed order ((tw_nb + tb) TILE)
  double *d_tw_cr = nullptr, *d_tw_dr = nullptr;  // BR forward input / backward input (tw_nb TILE)
  double *d_tw_s1 = nullptr, *d_tw_s2 = nullptr;  // BR sweep outputs (atomic access only)
  int64_t band_rows() const { return tw_a > 0 ? nt + tb : nt; }

  // fill-reducing ordering of the single-context factorization: 0 = the band
  // (above), 1 = nested dissection (multifrontal, fronts built on first use)
  int ordering = 0;
  NdDev *nd = nullptr;
  int nd_fill_prec = 64;  // precision band_fill chose for the next front factorization

  // staging of mafem_set_prescribed (grown on demand, kept)
  int64_t *d_presc_idx = nullptr;
  double *d_presc_val = nullptr;
  int64_t presc_cap = 0;

  // pull force (PullForce.jl:8-80): the pulled element, its adjacent elements
  // in the reference's order and, per adjacent element and local node, the
  // local index in the pulled element of the same node (-1: not a pulled node)
  int32_t pf_pull_el = -1;
  std::vector<int32_t> h_pf_el;
  std::vector<int8_t> h_pf_slot;
  std::string pf_err;            // why the maps do not exist (empty: they do)
  int32_t *d_pf_el = nullptr;    // device copies, allocated on first use
  int8_t *d_pf_slot = nullptr;
  double *d_pf_rel = nullptr;    // r_el of the adjacent elements [n][ne]
  double *d_pf_out = nullptr;    // x_pull[3], f_pull[3]

  // output sampling (Output.jl:30-118): per boundary, its elements in order as
  // (0-based element, first record in d_bufns) pairs from h_obel[2 h_boff[b]];
  // corner records (RECP doubles each) and their elements
  std::vector<int32_t> h_obel;
  int64_t h_boff[5] = {};
  std::vector<double> h_crnr;    // empty: not given
  int32_t crnr_el[4] = {};
  int32_t *d_obel = nullptr;
  double *d_crnr = nullptr, *d_out = nullptr;
};

// ---------------------------------------------------------------- batches
//
// A batch advances many independent patches (each a Ctx with its own device
// tables, state and band storage) through the same launches: every batched
// kernel maps its block to a patch through a segment map (block -> patch,
// first block of each patch) and reads the patch's sizes and buffers from a
// PatchDev descriptor.  Per-patch flags in device memory switch patches on
// and off, so one recorded launch sequence serves every Newton iteration.

struct PatchDev {
  int64_t numel, numnp, nx, nnzb, n_edges, nt;
  int32_t ndf, ne, tb, motion;
  int32_t mm[3];
  int32_t pad0;
  PhysParams ph;
  double enr, nmdf;
  const int32_t *ix, *uel;
  const uint8_t *active;
  const double *ufns, *dbt, *bufns;
  const EdgeItem *edges;
  double *xms, *cps, *kel, *rel;
  const int32_t *rowptr, *colidx, *brow;
  double *bsr;
  const int64_t *gptr, *gsrc, *rptr, *rsrc;
  double *r, *b, *x, *w, *t, *t2;
  void *band, *linv, *uinv;
  int32_t *piv;
  double *V;         // GMRES basis, m + 1 vectors of length nt TILE (batch-owned)
  int32_t prec;      // band storage precision of this patch: 32 or 64
  int32_t sweep_w;   // workgroups of its triangular sweeps
  double *s1, *s2;   // forward/backward outputs of multi-workgroup sweeps (atomic access only)
  int *err;          // count of sweep waits that ran out
  unsigned long long *amax;  // static pivoting: max |K| bits, replaced-pivot count
};

// Block -> patch map of one batched launch.
struct Seg {
  int32_t *map = nullptr;   // [nblk] patch of each block
  int64_t *blk0 = nullptr;  // [P] first block of each patch (unused patches: -1)
  unsigned nblk = 0;
};

// Patches active at one step K of the batched factorization, grouped in
// classes of equal tile half-bandwidth; patches are listed in `order`
// (class by class, each class by decreasing tile-row count), so the active
// members of a class are a prefix of it.
constexpr int MAXC = 40;
struct StepMap {
  int32_t ncls;
  int32_t first[MAXC];   // index into order of the class's first patch
  int32_t count[MAXC];   // active patches of the class at this step
  int32_t tb[MAXC];
  int32_t blk0[MAXC + 1];  // first block of the class in this launch
};

// Per-patch flags and scalars of a batch, all in device memory.  The linear
// solve of each Newton iteration is right-preconditioned GMRES(m) in FP64
// with the kept band LU as preconditioner; H, cs, sn, g, hc hold each
// patch's small Arnoldi/Givens state.
struct BatchFlags {
  uint8_t *on;       // still iterating Newton
  uint8_t *sol;      // linear solve not yet converged in this iteration
  uint8_t *fac;      // needs a factorization before its next solve
  uint8_t *kon;      // Arnoldi process running in the current cycle
  uint8_t *fresh;    // factorized since its last converged solve
  int32_t *iters;    // Newton iterations done in this time step
  int32_t *kits;     // Arnoldi steps of this iteration's solve (all cycles)
  int32_t *kdim;     // basis size of the current cycle
  int32_t *nfac;     // factorizations (running count)
  int32_t *kits_tot; // Arnoldi steps (running count)
  double *time, *dt; // per patch
  double *bn, *relres, *eps, *scale;
  double *relres0;   // true relative residual at the start of the last restart cycle
  double *H;         // [P][m][m+1]: column j of the Hessenberg matrix at H + (p m + j)(m+1)
  double *cs, *sn;   // [P][m] Givens rotations
  double *g;         // [P][m+1] rotated right-hand side
  double *hc;        // [P][m+1] projection coefficients, then the least-squares solution
  double *eps_hist;  // [P][max_iter]
  double *part;      // [vec blocks] per-block partial sums
  double *part2;     // [m+1][vec blocks] per-block partial dot products
  uint8_t *esc;      // FP32 factorization could not precondition the solve: move to FP64
  uint8_t *pert;     // factor with static pivoting (set after a fresh FP64 factorization failed)
  int32_t *count;    // [8] patches still on, solves unconverged, factorizations pending,
                     // hard failures, precision escalations
};

struct BatchPlan {
  int P = 0;
  bool has32 = false, has64 = false;  // precision groups present
  int max_iter = 14;
  int m = 20;                // GMRES basis size
  int refactor_iters = 6;
  int escalate_iters = 12;
  double rtol = 1e-10;
  hipStream_t stream = nullptr;
  PatchDev *d_pd = nullptr;
  int32_t *d_order = nullptr;
  BatchFlags f{};
  std::vector<int> motions;         // motions present
  std::vector<Seg> area, edge;      // per motion present
  Seg bsr, vec, node, zero, tiles, fold;
  Seg mw;                           // blocks of the multi-workgroup sweeps
  std::vector<std::vector<StepMap>> steps32, steps64;  // factorization steps per precision group
  int64_t max_nt = 0;
};

// batched launch sequences (assemble.hip, band.hip, vec.hip)
void b_assemble(BatchPlan &B);
void b_predict(BatchPlan &B);
void b_factor(BatchPlan &B);
void b_band_solve(BatchPlan &B, const uint8_t *mask);  // t = LU^-1 w where mask
void b_solve_begin(BatchPlan &B);  // sol = on; b = -r, x = 0, |b|, kits = 0
void b_gmres_cycle(BatchPlan &B);  // one restart cycle for patches with sol
void b_solve_end(BatchPlan &B);    // unconverged solves: refactor and redo, or give up
void b_update(BatchPlan &B);       // refactor policy, state update, ε, Newton flags, counts
void b_mark_factored(BatchPlan &B);
void b_count(BatchPlan &B);        // count[0] = on, count[2] = on && fac
void band_alloc(Ctx &c, int prec); // band storage of the given precision

// Serialises graph recording against context creation on other host threads
// (creation makes synchronous null-stream copies, which would invalidate a
// capture in progress).
std::mutex &capture_mutex();

// kernels / launchers (assemble.hip)
int launch_assemble(Ctx &c, double time, double dt);
// pulled-element mean position and pull force of the current state into
// out[6], on stream s: el / slot are device copies of c.h_pf_el / c.h_pf_slot,
// rel receives the adjacent elements' residuals (h_pf_el.size() x ne)
void launch_pull_force(Ctx &c, hipStream_t s, const int32_t *el, const int8_t *slot, double *rel,
                       double *out);
void ensure_pull(Ctx &c);  // abi.hip: device copies of the pull-force maps
// generate_output sampling of the device state into c.d_out: xout
// (Z1 x Z2 x 3) then uout (Z1 x Z2 x ndf), column-major, Zα = 3 numαel + 2
void launch_output(Ctx &c);
void launch_predict(Ctx &c, const double *d_u, double dt);
// band LU and solves (band.hip)
void band_fill(Ctx &c, int prec);
// perturb: static pivoting (diagonal-tile pivots below eps max |K| replaced
// by +-eps max |K|), used when a factorization without it failed
void band_factor(Ctx &c, bool perturb = false);
void band_solve(Ctx &c, const double *d_rhs, double *d_x);
// nested-dissection fronts (band.hip; band_fill / band_factor / band_solve
// route to them when Ctx::ordering is 1)
void nd_free(Ctx &c);
int nd_prec(const Ctx &c);       // storage precision of the fronts (0: none)
int nd_fronts(const Ctx &c);     // fronts of the elimination tree (0: not built)
int nd_steps(const Ctx &c);      // dependent tile steps of one factorization
double nd_flops(const Ctx &c);   // algorithmic FLOPs of one factorization
// durations (sum, us) / count / FLOPs / bytes of the last factorization's step launches
struct NdSpans {
  double us = 0, flops = 0, bytes = 0;  // sums over the timed launches
  int count = 0;
};
bool nd_read_spans(Ctx &c, NdSpans &step, NdSpans &schur);
int64_t nd_bytes(const Ctx &c);  // bytes of factor storage
void bsr_spmv(Ctx &c, const double *d_x, double *d_y);
void vec_axpby(Ctx &c, int64_t n, double a, const double *x, double b, double *y);
void vec_zero(Ctx &c, int64_t n, double *y);
void vec_scale(Ctx &c, int64_t n, double a, const double *x, double *y);
double vec_norm2_masked(Ctx &c, const double *x);   // active entries only, blocking
double vec_norm2(Ctx &c, const double *x, int64_t n);  // all entries, blocking
// multi-vector kernels of the Krylov solver (vec.hip)
void vec_mdot(Ctx &c, int64_t n, int nvec, const double *V, int64_t ld, const double *w, double *h_out);
void vec_mcomb(Ctx &c, int64_t n, int nvec, const double *V, int64_t ld, const double *h_coef, double alpha, double *w);
// Krylov solver (krylov.hip): K x = b, x overwritten; returns |b - K x| / |b|
double fgmres(Ctx &c, const double *d_b, double *d_x, int *iters);
void ensure_krylov(Ctx &c);
void apply_update(Ctx &c, const double *d_du, double dt);

}  // namespace mafem

struct mafem_ctx {
  mafem::Ctx c;
};

#define MAFEM_HIP(call)                                                         \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) throw mafem::HipError(e_, #call, __FILE__, __LINE__); \
  } while (0)

namespace mafem {
// invalid or unsupported input: message and MAFEM_E* code for the C ABI
struct ArgError {
  std::string msg;
  int code;
};
struct HipError {
  hipError_t code;
  const char *what, *file;
  int line;
  HipError(hipError_t c, const char *w, const char *f, int l) : code(c), what(w), file(f), line(l) {}
};
}  // namespace mafem
