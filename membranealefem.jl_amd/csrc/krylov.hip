// Flexible GMRES(m) in FP64 on the extended tangent, right-preconditioned by
// the band LU (FP64 or FP32 storage).  Arnoldi uses classical Gram--Schmidt
// with one re-orthogonalisation pass (two deterministic multi-dot reductions
// per step); the small Hessenberg least-squares problem is solved on the host
// with Givens rotations.  The preconditioner may be stale (factored from an
// earlier tangent): convergence is measured on the FP64 residual b - K x.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

namespace mafem {

void ensure_krylov(Ctx &c) {
  const int m = std::max(1, std::min(64, (int)c.opts.restart));
  if (c.d_V && c.kry_m == m) return;
  std::lock_guard<std::mutex> no_capture(capture_mutex());  // allocation: not during a capture
  if (c.d_V) {
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    (void)hipFree(c.d_V);
    (void)hipFree(c.d_Z);
    c.d_V = c.d_Z = nullptr;
  }
  const int64_t ld = c.nt * TILE;
  MAFEM_HIP(hipMalloc((void **)&c.d_V, sizeof(double) * ld * (m + 1)));
  MAFEM_HIP(hipMalloc((void **)&c.d_Z, sizeof(double) * ld * m));
  // padding entries (beyond the extended system) must stay zero
  MAFEM_HIP(hipMemset(c.d_V, 0, sizeof(double) * ld * (m + 1)));
  MAFEM_HIP(hipMemset(c.d_Z, 0, sizeof(double) * ld * m));
  if (!c.d_h) {
    MAFEM_HIP(hipMalloc((void **)&c.d_h, sizeof(double) * 72));
    MAFEM_HIP(hipMalloc((void **)&c.d_coef, sizeof(double) * 72));
    MAFEM_HIP(hipMalloc((void **)&c.d_mdot_part, sizeof(double) * 72 * 256));
  }
  c.kry_m = m;
}

double fgmres(Ctx &c, const double *d_b, double *d_x, int *iters_out) {
  ensure_krylov(c);
  const int m = c.kry_m;
  const int64_t n = c.nx, ld = c.nt * TILE;
  const double rtol = c.opts.rtol;
  const int maxit = std::max(1, (int)c.opts.max_iters);
  double *V = c.d_V, *Z = c.d_Z, *w = c.d_w;
  int total = 0;
  const double bn = vec_norm2(c, d_b, n);
  if (bn == 0.0) {
    vec_zero(c, ld, d_x);
    *iters_out = 0;
    return 0.0;
  }
  band_solve(c, d_b, d_x);  // x0 = M^-1 b
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), h(m + 1), h2(m + 1),
      y(m);
  double rel = 0.0;
  for (;;) {
    bsr_spmv(c, d_x, w);
    vec_axpby(c, n, 1.0, d_b, -1.0, w);  // w = b - K x
    const double beta = vec_norm2(c, w, n);
    rel = beta / bn;
    if (!(rel > rtol) || total >= maxit) break;
    vec_scale(c, n, 1.0 / beta, w, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m && total < maxit; ++j) {
      band_solve(c, V + j * ld, Z + j * ld);
      double *vn = V + (j + 1) * ld;
      bsr_spmv(c, Z + j * ld, vn);
      vec_mdot(c, n, j + 1, V, ld, vn, h.data());
      vec_mcomb(c, n, j + 1, V, ld, h.data(), -1.0, vn);
      vec_mdot(c, n, j + 1, V, ld, vn, h2.data());
      vec_mcomb(c, n, j + 1, V, ld, h2.data(), -1.0, vn);
      for (int k = 0; k <= j; ++k) h[k] += h2[k];
      const double hn = vec_norm2(c, vn, n);
      if (hn > 0.0) vec_scale(c, n, 1.0 / hn, vn, vn);
      for (int k = 0; k < j; ++k) {
        const double a = h[k], b2 = h[k + 1];
        h[k] = cs[k] * a + sn[k] * b2;
        h[k + 1] = -sn[k] * a + cs[k] * b2;
      }
      const double r = std::hypot(h[j], hn);
      cs[j] = h[j] / r;
      sn[j] = hn / r;
      h[j] = r;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int k = 0; k <= j; ++k) H[(size_t)k * m + j] = h[k];
      ++total;
      if (std::fabs(g[j + 1]) <= 0.5 * rtol * bn || hn == 0.0) {
        ++j;
        break;
      }
    }
    for (int k = j - 1; k >= 0; --k) {
      double s = g[k];
      for (int l = k + 1; l < j; ++l) s -= H[(size_t)k * m + l] * y[l];
      y[k] = s / H[(size_t)k * m + k];
    }
    if (j > 0) vec_mcomb(c, n, j, Z, ld, y.data(), 1.0, d_x);
  }
  c.n_gmres += total;
  *iters_out = total;
  return rel;
}

}  // namespace mafem
