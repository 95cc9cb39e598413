// Internal declarations shared by the HIP translation units of libmafem.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

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

  // CSC export (1-based) over active DOFs
  int64_t nnz = 0;
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
  bool diag_timed = false;       // ev[10]/ev[11] bracket a k_diag launch not yet read

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
};

// kernels / launchers (assemble.hip)
int launch_assemble(Ctx &c, double time, double dt);
void launch_predict(Ctx &c, const double *d_u, double dt);
// band LU and solves (band.hip)
void band_fill(Ctx &c, int prec);
void band_factor(Ctx &c);
void band_solve(Ctx &c, const double *d_rhs, double *d_x);
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

#define MAFEM_HIP(call)                                                         \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) throw mafem::HipError(e_, #call, __FILE__, __LINE__); \
  } while (0)

namespace mafem {
struct HipError {
  hipError_t code;
  const char *what, *file;
  int line;
  HipError(hipError_t c, const char *w, const char *f, int l) : code(c), what(w), file(f), line(l) {}
};
}  // namespace mafem
