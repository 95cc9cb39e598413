/*
 * CPU oracle: literal complex-step restatement of the reference's element
 * residual and tangent.  TEST INFRASTRUCTURE ONLY -- the checker for the HIP
 * path and the timed CPU baseline; never linked into the product library.
 *
 * Follows (reference paths):
 *   src/analysis/GeoDynStress.jl:91-205   gds()          (geometry, fields,
 *                                                         stresses, B operators)
 *   src/analysis/FiniteElement.jl:253-330 area_dof_res() (Gauss loop, physics,
 *                                                         Dohrmann--Bochev term)
 *   src/analysis/FiniteElement.jl:208-242 area_res()     (local packing)
 *   src/analysis/FiniteElement.jl:338-400 bdry_res()     (Neumann edge residual)
 *   src/analysis/FiniteElement.jl:431-452 tau_nu()       (edge frame; the
 *                                          normalising dot product conjugates)
 *   src/analysis/FiniteElement.jl:98-126, 156-184  element_tangent()
 *                                          (complex-step columns, including the
 *                                           dt-scaled position column of a
 *                                           motion DOF)
 *
 * Arithmetic is ComplexF64 throughout, as in the reference, with step ek.
 * Pinned by: the closed forms of test/analysis/geo_dyn_stress_tests.jl
 * (paraboloid geometry, H, K; Couette/Poiseuille stress; area), the flat
 * equilibrium invariant r = 0, flat-patch closed forms of the mesh-motion
 * rows (ALEV/ALEVB sigma_m, p^m and rp terms; the EUL rm projection) and of
 * the tension row's surface divergence, central finite differences of the
 * residual, and a Newton-convergence check (tests/test_oracle_*.py).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef double complex cplx;

#define NEN 9
#define REC 55

enum { STATIC = 1, EUL = 2, LAG = 3, ALEV = 4, ALEVB = 5 };
enum { BOTTOM = 1, RIGHT = 2, TOP = 3, LEFT = 4 };
enum { SHEAR = 1, STRETCH = 2, MOMENT = 3 };
enum { F_BEND = 5 };

typedef struct {
  int64_t numel, numnp, ndf;
  const int64_t *IX;     /* 9 x numel, column-major, 1-based node ids        */
  const int64_t *ID;     /* ndf x numnp, column-major, 0 = prescribed        */
  int32_t v_order[3], m_order[3], lam_order, p_order, mm_order[3];
  int32_t motion, scenario;
  double kb, kg, zv, pn, adb, am, ek;
} orc_mesh;

/* Gauss-point quantities (GeoDynStress.jl:42-67), complex */
typedef struct {
  cplx x[3], a1[3], a2[3], xab[3][3]; /* xab[comp][voigt]                   */
  cplx aab[2][2], aup[2][2], aupv[3][2], J, G[3][2], n[3], b[2][2], bup[2][2];
  cplx H, K, v[3], dv[3][2], lam, pm, vm[3], dvm[3][2];
  cplx sig[3], sigm[3], M[3];
  cplx Ba[3][27], Bb[3][27];
} gds_t;

static void gds(const cplx *X, const cplx *C, int64_t ndf, const orc_mesh *m,
                const double *rec, gds_t *g) {
  const double *N = rec + 1, *dN1 = rec + 10, *dN2 = rec + 19;
  const double *dd[3] = {rec + 28, rec + 37, rec + 46};
  /* get_dof_cps: absent DOFs read as zero (GeoDynStress.jl:190-205) */
  cplx V[NEN][3], VM[NEN][3], L[NEN], P[NEN];
  for (int a = 0; a < NEN; ++a) {
    for (int i = 0; i < 3; ++i) {
      V[a][i] = m->v_order[i] ? C[a + NEN * (m->v_order[i] - 1)] : 0;
      VM[a][i] = m->m_order[i] ? C[a + NEN * (m->m_order[i] - 1)] : 0;
    }
    L[a] = m->lam_order ? C[a + NEN * (m->lam_order - 1)] : 0;
    P[a] = m->p_order ? C[a + NEN * (m->p_order - 1)] : 0;
  }
  for (int i = 0; i < 3; ++i) {
    cplx s = 0, s1 = 0, s2 = 0, t[3] = {0, 0, 0};
    for (int a = 0; a < NEN; ++a) {
      cplx xi = X[a + NEN * i];
      s += xi * N[a];
      s1 += xi * dN1[a];
      s2 += xi * dN2[a];
      for (int v = 0; v < 3; ++v) t[v] += xi * dd[v][a];
    }
    g->x[i] = s; g->a1[i] = s1; g->a2[i] = s2;
    for (int v = 0; v < 3; ++v) g->xab[i][v] = t[v];
  }
  cplx a11 = 0, a12 = 0, a22 = 0;
  for (int i = 0; i < 3; ++i) {
    a11 += g->a1[i] * g->a1[i];
    a12 += g->a1[i] * g->a2[i];
    a22 += g->a2[i] * g->a2[i];
  }
  g->aab[0][0] = a11; g->aab[0][1] = a12; g->aab[1][0] = a12; g->aab[1][1] = a22;
  cplx deta = a11 * a22 - a12 * a12;
  g->aup[0][0] = a22 / deta; g->aup[0][1] = -a12 / deta;
  g->aup[1][0] = -a12 / deta; g->aup[1][1] = a11 / deta;
  for (int i = 0; i < 3; ++i)
    for (int be = 0; be < 2; ++be)
      g->aupv[i][be] = g->a1[i] * g->aup[0][be] + g->a2[i] * g->aup[1][be];
  g->J = csqrt(deta);
  for (int v = 0; v < 3; ++v)
    for (int mu = 0; mu < 2; ++mu) {
      cplx s = 0;
      for (int i = 0; i < 3; ++i) s += g->xab[i][v] * g->aupv[i][mu];
      g->G[v][mu] = s;
    }
  g->n[0] = (g->a1[1] * g->a2[2] - g->a1[2] * g->a2[1]) / g->J;
  g->n[1] = (g->a1[2] * g->a2[0] - g->a1[0] * g->a2[2]) / g->J;
  g->n[2] = (g->a1[0] * g->a2[1] - g->a1[1] * g->a2[0]) / g->J;
  cplx bf[3];
  for (int v = 0; v < 3; ++v) {
    cplx s = 0;
    for (int i = 0; i < 3; ++i) s += g->xab[i][v] * g->n[i];
    bf[v] = s;
  }
  g->b[0][0] = bf[0]; g->b[0][1] = bf[2]; g->b[1][0] = bf[2]; g->b[1][1] = bf[1];
  cplx tmp[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      tmp[i][j] = g->aup[i][0] * g->b[0][j] + g->aup[i][1] * g->b[1][j];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      g->bup[i][j] = tmp[i][0] * g->aup[0][j] + tmp[i][1] * g->aup[1][j];
  g->H = (g->aup[0][0] * g->b[0][0] + g->aup[0][1] * g->b[0][1] +
          g->aup[1][0] * g->b[1][0] + g->aup[1][1] * g->b[1][1]) / 2;
  g->K = (g->b[0][0] * g->b[1][1] - g->b[0][1] * g->b[1][0]) / deta;

  for (int i = 0; i < 3; ++i) {
    cplx s = 0, s1 = 0, s2 = 0, t = 0, t1 = 0, t2 = 0;
    for (int a = 0; a < NEN; ++a) {
      s += V[a][i] * N[a]; s1 += V[a][i] * dN1[a]; s2 += V[a][i] * dN2[a];
      t += VM[a][i] * N[a]; t1 += VM[a][i] * dN1[a]; t2 += VM[a][i] * dN2[a];
    }
    g->v[i] = s; g->dv[i][0] = s1; g->dv[i][1] = s2;
    g->vm[i] = t; g->dvm[i][0] = t1; g->dvm[i][1] = t2;
  }
  cplx sl = 0, spm = 0;
  for (int a = 0; a < NEN; ++a) { sl += L[a] * N[a]; spm += P[a] * N[a]; }
  g->lam = sl; g->pm = spm;

  /* stresses (GeoDynStress.jl:148-159) */
  cplx pi[2][2], pim[2][2];
  for (int al = 0; al < 2; ++al)
    for (int be = 0; be < 2; ++be) {
      cplx s = 0, t = 0;
      for (int i = 0; i < 3; ++i) {
        cplx w = g->dv[i][0] * g->aup[0][be] + g->dv[i][1] * g->aup[1][be];
        cplx wm = g->dvm[i][0] * g->aup[0][be] + g->dvm[i][1] * g->aup[1][be];
        s += g->aupv[i][al] * w;
        t += g->aupv[i][al] * wm;
      }
      pi[al][be] = s; pim[al][be] = t;
    }
  cplx sg[2][2], sgm[2][2], Mm[2][2];
  cplx c0 = m->kb * g->H * g->H - m->kg * g->K;
  for (int al = 0; al < 2; ++al)
    for (int be = 0; be < 2; ++be) {
      cplx ps = (pi[al][be] + pi[be][al]) * m->zv;
      cplx pms = (pim[al][be] + pim[be][al]) * m->zv;
      sg[al][be] = g->aup[al][be] * (c0 + g->lam) - g->bup[al][be] * 2 * m->kb * g->H + ps;
      sgm[al][be] = g->aup[al][be] * c0 - g->bup[al][be] * 2 * m->kb * g->H + pms;
      Mm[al][be] = g->aup[al][be] * g->H * (m->kb + 2 * m->kg) - g->bup[al][be] * m->kg;
    }
  g->sig[0] = sg[0][0]; g->sig[1] = sg[1][1]; g->sig[2] = 2 * sg[0][1];
  g->sigm[0] = sgm[0][0]; g->sigm[1] = sgm[1][1]; g->sigm[2] = 2 * sgm[0][1];
  g->M[0] = Mm[0][0]; g->M[1] = Mm[1][1]; g->M[2] = Mm[0][1] + Mm[1][0];

  /* shape function operators (GeoDynStress.jl:163-174) */
  for (int a = 0; a < NEN; ++a) {
    cplx nn[3];
    for (int v = 0; v < 3; ++v)
      nn[v] = dd[v][a] - (dN1[a] * g->G[v][0] + dN2[a] * g->G[v][1]);
    for (int i = 0; i < 3; ++i) {
      int c = i + 3 * a;
      g->Ba[0][c] = g->a1[i] * dN1[a];
      g->Ba[1][c] = g->a2[i] * dN2[a];
      g->Ba[2][c] = (g->a1[i] * dN2[a] + g->a2[i] * dN1[a]) * 0.5;
      for (int v = 0; v < 3; ++v) g->Bb[v][c] = g->n[i] * nn[v];
    }
  }
}

static void inv3(const double A[3][3], double R[3][3]) {
  double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  R[0][0] = c00 / det;
  R[1][0] = c01 / det;
  R[2][0] = c02 / det;
  R[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
  R[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
  R[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
  R[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
  R[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
  R[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
}

/* calc_elem_dof_residuals + calc_elem_residual packing (FiniteElement.jl:208-330) */
static void area_res(const orc_mesh *m, const double *recs /* 9 records */,
                     const cplx *X, const cplx *C, cplx *r) {
  const int64_t ndf = m->ndf;
  const int mo = m->motion;
  const int ale = (mo == ALEV || mo == ALEVB);
  cplx rv[27] = {0}, rm[27] = {0}, rl[NEN] = {0}, rp[NEN] = {0};
  double GDB[3][NEN] = {{0}}, HDB[3][3] = {{0}};
  const double xi[3] = {-sqrt(3.0 / 5.0), 0.0, sqrt(3.0 / 5.0)};
  gds_t g;
  for (int gp = 0; gp < 9; ++gp) {
    const double *rec = recs + REC * gp;
    const double *N = rec + 1, w = rec[0];
    gds(X, C, ndf, m, rec, &g);
    cplx Jw = g.J * w;
    for (int c = 0; c < 27; ++c) {
      cplx s = 0;
      for (int v = 0; v < 3; ++v) s += g.Ba[v][c] * g.sig[v] + g.Bb[v][c] * g.M[v];
      rv[c] += s * Jw;
    }
    if (m->pn != 0.0)
      for (int a = 0; a < NEN; ++a)
        for (int i = 0; i < 3; ++i) rv[i + 3 * a] -= g.n[i] * N[a] * m->pn * Jw;
    cplx trv = 0;
    for (int i = 0; i < 3; ++i) trv += g.aupv[i][0] * g.dv[i][0] + g.aupv[i][1] * g.dv[i][1];
    for (int a = 0; a < NEN; ++a)
      rl[a] += N[a] * w * (g.J * trv - m->adb * g.lam / m->zv);
    if (mo == EUL) {
      cplx nv = g.n[0] * g.v[0] + g.n[1] * g.v[1] + g.n[2] * g.v[2];
      for (int a = 0; a < NEN; ++a)
        for (int i = 0; i < 3; ++i)
          rm[i + 3 * a] += (g.vm[i] - g.n[i] * nv) * N[a] * m->am * Jw;
    } else if (ale) {
      for (int c = 0; c < 27; ++c) {
        cplx s = 0;
        for (int v = 0; v < 3; ++v) s += g.Ba[v][c] * g.sigm[v];
        if (mo == ALEVB)
          for (int v = 0; v < 3; ++v) s += g.Bb[v][c] * g.M[v];
        rm[c] += s * Jw;
      }
      for (int a = 0; a < NEN; ++a)
        for (int i = 0; i < 3; ++i) rm[i + 3 * a] -= g.n[i] * N[a] * g.pm * Jw;
      cplx nd = g.n[0] * (g.vm[0] - g.v[0]) + g.n[1] * (g.vm[1] - g.v[1]) +
                g.n[2] * (g.vm[2] - g.v[2]);
      for (int a = 0; a < NEN; ++a) {
        rp[a] -= N[a] * w * g.J * nd;
        rp[a] -= N[a] * w * m->adb * g.pm / m->zv;
      }
    }
    double ndb[3] = {xi[gp % 3], xi[gp / 3], 1.0};
    for (int i = 0; i < 3; ++i) {
      for (int a = 0; a < NEN; ++a) GDB[i][a] += ndb[i] * N[a] * w;
      for (int j = 0; j < 3; ++j) HDB[i][j] += ndb[i] * ndb[j] * w;
    }
  }
  double Hi[3][3], T[NEN][NEN];
  inv3(HDB, Hi);
  for (int a = 0; a < NEN; ++a)
    for (int b = 0; b < NEN; ++b) {
      double s = 0;
      for (int i = 0; i < 3; ++i) {
        double t = 0;
        for (int j = 0; j < 3; ++j) t += Hi[i][j] * GDB[j][b];
        s += GDB[i][a] * t;
      }
      T[a][b] = s;
    }
  for (int a = 0; a < NEN; ++a) {
    cplx sl = 0, spp = 0;
    for (int b = 0; b < NEN; ++b) {
      sl += T[a][b] * (m->lam_order ? C[b + NEN * (m->lam_order - 1)] : 0);
      if (ale) spp += T[a][b] * C[b + NEN * (m->p_order - 1)];
    }
    rl[a] += sl * m->adb / m->zv;
    if (ale) rp[a] += spp * m->adb / m->zv;
  }
  for (int k = 0; k < NEN * ndf; ++k) r[k] = 0;
  for (int a = 0; a < NEN; ++a) {
    for (int i = 0; i < 3; ++i) {
      if (m->v_order[i]) r[m->v_order[i] - 1 + ndf * a] = rv[i + 3 * a];
      if (m->m_order[i]) r[m->m_order[i] - 1 + ndf * a] = rm[i + 3 * a];
    }
    r[m->lam_order - 1 + ndf * a] = rl[a];
    if (m->p_order) r[m->p_order - 1 + ndf * a] = rp[a];
  }
}

/* calc_τ_ν (FiniteElement.jl:431-452).  The normalisation uses Julia's dot,
 * which conjugates its first argument: the factor has no imaginary part. */
static void tau_nu(int bdry, const gds_t *g, cplx tau[3], cplx nu[3]) {
  for (int i = 0; i < 3; ++i) {
    if (bdry == BOTTOM) tau[i] = g->a1[i];
    else if (bdry == RIGHT) tau[i] = g->a2[i];
    else if (bdry == TOP) tau[i] = -g->a1[i];
    else tau[i] = -g->a2[i];
  }
  cplx d = conj(tau[0]) * tau[0] + conj(tau[1]) * tau[1] + conj(tau[2]) * tau[2];
  cplx s = csqrt(d);
  for (int i = 0; i < 3; ++i) tau[i] /= s;
  nu[0] = tau[1] * g->n[2] - tau[2] * g->n[1];
  nu[1] = tau[2] * g->n[0] - tau[0] * g->n[2];
  nu[2] = tau[0] * g->n[1] - tau[1] * g->n[0];
}

/* calc_bdry_element_residual (FiniteElement.jl:338-400) */
static int bdry_res(const orc_mesh *m, const double *recs /* 3 records */,
                    int bdry, int ntype, double nval, double time, double bend_tm,
                    const cplx *X, const cplx *C, cplx *r) {
  const int64_t ndf = m->ndf;
  cplx rv[27] = {0};
  gds_t g;
  for (int gp = 0; gp < 3; ++gp) {
    const double *rec = recs + REC * gp;
    const double *N = rec + 1, *dN1 = rec + 10, *dN2 = rec + 19, w = rec[0];
    gds(X, C, ndf, m, rec, &g);
    cplx tau[3], nu[3];
    tau_nu(bdry, &g, tau, nu);
    cplx q0 = g.aupv[0][0] * tau[0] + g.aupv[1][0] * tau[1] + g.aupv[2][0] * tau[2];
    cplx q1 = g.aupv[0][1] * tau[0] + g.aupv[1][1] * tau[1] + g.aupv[2][1] * tau[2];
    cplx JG = 1.0 / csqrt(q0 * q0 + q1 * q1);
    if (ntype == STRETCH || ntype == SHEAR) {
      for (int a = 0; a < NEN; ++a)
        for (int i = 0; i < 3; ++i) {
          cplx f = nval * (ntype == STRETCH ? nu[i] : tau[i]);
          rv[i + 3 * a] -= f * N[a] * JG * w;
        }
    } else if (ntype == MOMENT && m->scenario == F_BEND) {
      cplx nup[2];
      for (int al = 0; al < 2; ++al)
        nup[al] = g.aupv[0][al] * nu[0] + g.aupv[1][al] * nu[1] + g.aupv[2][al] * nu[2];
      double Mval = nval * fmin(time / bend_tm, 1.0);
      for (int a = 0; a < NEN; ++a) {
        cplx dn = dN1[a] * nup[0] + dN2[a] * nup[1];
        for (int i = 0; i < 3; ++i) rv[i + 3 * a] -= g.n[i] * dn * Mval * JG * w;
      }
    } else {
      return -1;
    }
  }
  for (int k = 0; k < NEN * ndf; ++k) r[k] = 0;
  for (int a = 0; a < NEN; ++a)
    for (int i = 0; i < 3; ++i)
      if (m->v_order[i]) r[m->v_order[i] - 1 + ndf * a] = rv[i + 3 * a];
  return 0;
}

/* one element: r_el (real part) and the complex-step K_el, column-major
 * (9 ndf) x (9 ndf), following FiniteElement.jl:100-126 / 156-184. */
static int element_tangent(const orc_mesh *m, int64_t el, const double *recs,
                           int is_bdry, int bdry, int ntype, double nval,
                           double time, double bend_tm, const double *xms,
                           const double *cps, double dt, double *r_out,
                           double *K_out) {
  const int64_t ndf = m->ndf, nn = m->numnp, ne = NEN * ndf;
  cplx X[NEN * 3], C[NEN * 8], Xp[NEN * 3], Cp[NEN * 8], r[NEN * 8];
  const int64_t *ix = m->IX + NEN * el;
  for (int a = 0; a < NEN; ++a) {
    int64_t node = ix[a] - 1;
    for (int i = 0; i < 3; ++i) X[a + NEN * i] = xms[node + nn * i];
    for (int d = 0; d < ndf; ++d) C[a + NEN * d] = cps[node + nn * d];
  }
#define RES(XX, CC) (is_bdry ? bdry_res(m, recs, bdry, ntype, nval, time, bend_tm, XX, CC, r) \
                             : (area_res(m, recs, XX, CC, r), 0))
  if (RES(X, C)) return -1;
  for (int k = 0; k < ne; ++k) r_out[k] = creal(r[k]);
  for (int k = 0; k < ne * ne; ++k) K_out[k] = 0.0;
  const double ek = m->ek;
  for (int a = 0; a < NEN; ++a) {
    int64_t node = ix[a] - 1;
    for (int d = 0; d < ndf; ++d) {
      if (m->ID[d + ndf * node] == 0) continue;
      int64_t col = d + ndf * a;
      memcpy(Cp, C, sizeof(cplx) * NEN * ndf);
      Cp[a + NEN * d] += ek * I;
      RES(X, Cp);
      for (int k = 0; k < ne; ++k) K_out[k + ne * col] += cimag(r[k]) / ek;
      int comp = -1;
      for (int i = 0; i < 3; ++i)
        if (m->mm_order[i] == d + 1) { comp = i; break; }
      if (comp >= 0) {
        memcpy(Xp, X, sizeof(cplx) * NEN * 3);
        Xp[a + NEN * comp] += ek * I;
        RES(Xp, C);
        for (int k = 0; k < ne; ++k) K_out[k + ne * col] += (cimag(r[k]) / ek) * dt;
      }
    }
  }
#undef RES
  return 0;
}

/* Area elements [e0, e1) (0-based).  ufns: n_uel x 9 records, record (u, g)
 * at ((u * 9) + g) * 55.  Outputs per element: r (9 ndf) and K ((9 ndf)^2,
 * column-major).  OpenMP over elements when compiled with it. */
int orc_area_elements(const orc_mesh *m, const int64_t *uel_ids, const double *ufns,
                      const double *xms, const double *cps, double dt,
                      int64_t e0, int64_t e1, double *r_out, double *K_out) {
  const int64_t ne = NEN * m->ndf;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(|: err)
  for (int64_t e = e0; e < e1; ++e) {
    const double *recs = ufns + (uel_ids[e] - 1) * 9 * REC;
    err |= element_tangent(m, e, recs, 0, 0, 0, 0.0, 0.0, 1.0, xms, cps, dt,
                           r_out + ne * (e - e0), K_out + ne * ne * (e - e0));
  }
  return err;
}

/* Boundary elements: els[k] is the 0-based area element id, buel[k] the
 * 1-based unique boundary element id; ufns: n_uel x 3 records. */
int orc_bdry_elements(const orc_mesh *m, int bdry, int ntype, double nval,
                      double time, double bend_tm, int64_t nel, const int64_t *els,
                      const int64_t *buel, const double *ufns, const double *xms,
                      const double *cps, double dt, double *r_out, double *K_out) {
  const int64_t ne = NEN * m->ndf;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(|: err)
  for (int64_t k = 0; k < nel; ++k) {
    const double *recs = ufns + (buel[k] - 1) * 3 * REC;
    err |= element_tangent(m, els[k], recs, 1, bdry, ntype, nval, time, bend_tm,
                           xms, cps, dt, r_out + ne * k, K_out + ne * ne * k);
  }
  return err;
}

/* Gauss-point quantities for the closed-form tests: the real parts of
 * x, a_1, a_2, a^{ab}, J, H, K, sigma (Voigt) packed into out[22]. */
void orc_gds_point(const orc_mesh *m, const double *rec, const double *X9,
                   const double *C9, double *out) {
  cplx X[NEN * 3], C[NEN * 8];
  for (int k = 0; k < NEN * 3; ++k) X[k] = X9[k];
  for (int k = 0; k < NEN * m->ndf; ++k) C[k] = C9[k];
  gds_t g;
  gds(X, C, m->ndf, m, rec, &g);
  int k = 0;
  for (int i = 0; i < 3; ++i) out[k++] = creal(g.x[i]);
  for (int i = 0; i < 3; ++i) out[k++] = creal(g.a1[i]);
  for (int i = 0; i < 3; ++i) out[k++] = creal(g.a2[i]);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) out[k++] = creal(g.aup[i][j]);
  out[k++] = creal(g.J);
  out[k++] = creal(g.H);
  out[k++] = creal(g.K);
  for (int v = 0; v < 3; ++v) out[k++] = creal(g.sig[v]);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) out[k++] = creal(g.b[i][j]);
}

/* Residual only (no tangent) of one area element at real state, used by the
 * finite-difference checks. */
void orc_area_residual(const orc_mesh *m, const int64_t *uel_ids, const double *ufns,
                       int64_t el, const double *xms, const double *cps, double *r_out) {
  const int64_t ndf = m->ndf, nn = m->numnp;
  cplx X[NEN * 3], C[NEN * 8], r[NEN * 8];
  const int64_t *ix = m->IX + NEN * el;
  for (int a = 0; a < NEN; ++a) {
    int64_t node = ix[a] - 1;
    for (int i = 0; i < 3; ++i) X[a + NEN * i] = xms[node + nn * i];
    for (int d = 0; d < ndf; ++d) C[a + NEN * d] = cps[node + nn * d];
  }
  area_res(m, ufns + (uel_ids[el] - 1) * 9 * REC, X, C, r);
  for (int k = 0; k < NEN * ndf; ++k) r_out[k] = creal(r[k]);
}
