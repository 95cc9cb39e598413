// Element residual/tangent assembly for gfx950.
//
// One lane per (element, tangent column): lanes carry forward-mode dual
// numbers seeded with that column's direction, so each lane produces one
// column of K_el (all 9 ndf rows) and the residual entry of its own row.
// Element state and the tabulated basis records of the elements a workgroup
// touches are staged in LDS once; K_el is written row-major so a store
// instruction writes one contiguous row across the lanes.
//
// Physics follows src/analysis/GeoDynStress.jl:91-178 and
// src/analysis/FiniteElement.jl:253-330 (area), :338-400 and :431-452 (edges).
// The tangent columns follow FiniteElement.jl:110-126: a column of a motion
// DOF is d r/d cps + Δt d r/d xms, i.e. one directional derivative with seed
// (e_col, Δt e_(node,comp)).  Columns of prescribed DOFs (ID == 0) are zero.
#include "internal.h"

namespace mafem {

struct D {
  double v, t;
};
__device__ __forceinline__ D mk(double v, double t) { return D{v, t}; }
__device__ __forceinline__ D operator+(D a, D b) { return D{a.v + b.v, a.t + b.t}; }
__device__ __forceinline__ D operator-(D a, D b) { return D{a.v - b.v, a.t - b.t}; }
__device__ __forceinline__ D operator-(D a) { return D{-a.v, -a.t}; }
__device__ __forceinline__ D operator*(D a, D b) { return D{a.v * b.v, fma(a.v, b.t, a.t * b.v)}; }
__device__ __forceinline__ D operator*(double s, D a) { return D{s * a.v, s * a.t}; }
__device__ __forceinline__ D operator*(D a, double s) { return D{s * a.v, s * a.t}; }
__device__ __forceinline__ D recip(D a) {
  double r = 1.0 / a.v;
  return D{r, -a.t * r * r};
}
__device__ __forceinline__ D dsqrt(D a) {
  double s = sqrt(a.v);
  return D{s, a.t / (2.0 * s)};
}

// Real numbers with the dual numbers' interface: the derivative part reads as
// zero and is never computed (the area kernel's primal pass).
struct R {
  double v;
  static constexpr double t = 0.0;
};
__device__ __forceinline__ R operator+(R a, R b) { return R{a.v + b.v}; }
__device__ __forceinline__ R operator-(R a, R b) { return R{a.v - b.v}; }
__device__ __forceinline__ R operator-(R a) { return R{-a.v}; }
__device__ __forceinline__ R operator*(R a, R b) { return R{a.v * b.v}; }
__device__ __forceinline__ R operator*(double s, R a) { return R{s * a.v}; }
__device__ __forceinline__ R operator*(R a, double s) { return R{s * a.v}; }
__device__ __forceinline__ R recip(R a) { return R{1.0 / a.v}; }
__device__ __forceinline__ R dsqrt(R a) { return R{sqrt(a.v)}; }

template <class N>
__device__ __forceinline__ N mkN(double v, double t);
template <>
__device__ __forceinline__ D mkN<D>(double v, double t) { return D{v, t}; }
template <>
__device__ __forceinline__ R mkN<R>(double v, double) { return R{v}; }

// DOF layout of mesh.dofs for each motion (get_dofs, src/input/Bc.jl:414-431)
template <int MO>
struct Lay {
  // STATIC: the flat 2-D scenarios (F_CAVI, F_COUE, F_POIS; Bc.jl:58-157,
  // 286-316) with unknowns vx, vy, λ and no vz
  static constexpr int NDF = MO == STATIC ? 3 : (MO == LAG ? 4 : (MO == EUL ? 7 : 8));
  static constexpr int NE = 9 * NDF;
  static constexpr int NV = MO == STATIC ? 2 : 3;  // velocity components with unknowns
  static constexpr bool HAS_VM = (MO == EUL || MO == ALEV || MO == ALEVB);
  static constexpr bool HAS_PM = (MO == ALEV || MO == ALEVB);
  static constexpr int LAMI = MO == STATIC ? 2 : (MO == LAG ? 3 : 6);
  static constexpr int PMI = 7;
  // first column of the motion DOFs (get_m_motion_order, Mesh.jl:529-542)
  static constexpr int MOT0 = (MO == LAG) ? 0 : (MO == STATIC ? -1 : 3);
  // lanes per tangent column of the area kernel: each takes the rows of
  // ceil(9 / PARTS) nodes, so its accumulators fit the register file
  static constexpr int PARTS = AREA_PARTS[MO];
  static constexpr int MAXE = 256 / NE + 2;  // elements a 256-lane block can touch
};

// Geometry of one Gauss point with the lane's seed (GeoDynStress.jl:111-130).
template <class N = D>
struct Geo {
  N a1[3], a2[3], n[3], au1[3], au2[3], J;
  N g00, g01, g11;  // a^{αβ}
  N idet;           // 1 / det a_{αβ}
};

// Passes of the area kernel.  PLAIN: dual numbers throughout.  PRIMAL: one
// lane per (element, Gauss point) evaluates the point's values in real
// arithmetic and keeps them in LDS (slots below).  TANGENT: the lanes of the
// tangent columns take those values as the value parts of their dual numbers,
// so only the derivative parts are computed per column (sums over the nine
// nodes, quotients and the square root of the metric are not repeated).
enum { PLAIN = 0, PRIMAL = 1, TANGENT = 2 };
enum {
  S_A1 = 0, S_A2 = 3, S_N = 6, S_AU1 = 9, S_AU2 = 12, S_J = 15, S_IJ = 16, S_G00 = 17,
  S_G01 = 18, S_G11 = 19, S_IDET = 20, S_X11 = 21, S_X22 = 24, S_X12 = 27, S_GA0 = 30,
  S_GA1 = 33, S_BB = 36, S_H = 39, S_KC = 40, S_B = 41, S_V = 45, S_DV0 = 48, S_DV1 = 51,
  S_LAM = 54, S_SG = 55, S_MV = 58, S_VM = 61, S_DVM0 = 64, S_DVM1 = 67, S_PM = 70,
  S_SGM = 71, N_SLOTS = 74
};

template <int PS, class N>
__device__ __forceinline__ void fix(N &x, double *pr, int k) {
  if constexpr (PS == PRIMAL) pr[k] = x.v;
  else if constexpr (PS == TANGENT) x.v = pr[k];
}
template <int PS, class N, int M>
__device__ __forceinline__ void fixa(N (&x)[M], double *pr, int k) {
#pragma unroll
  for (int i = 0; i < M; ++i) fix<PS>(x[i], pr, k + i);
}

template <class N = D, int PS = PLAIN>
__device__ __forceinline__ void geometry(const double *X /*[3][9]*/, const double *bN1,
                                         const double *bN2, int a0, int i0, double sX,
                                         Geo<N> &g, double *pr = nullptr) {
  const double t1 = sX * bN1[a0], t2 = sX * bN2[a0];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s1 = 0, s2 = 0;
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      s1 = fma(X[i * 9 + a], bN1[a], s1);
      s2 = fma(X[i * 9 + a], bN2[a], s2);
    }
    g.a1[i] = mkN<N>(s1, i == i0 ? t1 : 0.0);
    g.a2[i] = mkN<N>(s2, i == i0 ? t2 : 0.0);
  }
  fixa<PS>(g.a1, pr, S_A1);
  fixa<PS>(g.a2, pr, S_A2);
  N a11 = g.a1[0] * g.a1[0] + g.a1[1] * g.a1[1] + g.a1[2] * g.a1[2];
  N a12 = g.a1[0] * g.a2[0] + g.a1[1] * g.a2[1] + g.a1[2] * g.a2[2];
  N a22 = g.a2[0] * g.a2[0] + g.a2[1] * g.a2[1] + g.a2[2] * g.a2[2];
  N det = a11 * a22 - a12 * a12;
  N iJ;
  if constexpr (PS == TANGENT) {
    // d(1/det) = -d det / det^2, d sqrt(det) = d det / (2 sqrt(det)), d(1/J) = -dJ / J^2
    const double idv = pr[S_IDET], jv = pr[S_J], ijv = pr[S_IJ];
    g.idet = N{idv, -det.t * idv * idv};
    g.J = N{jv, det.t * (0.5 * ijv)};
    iJ = N{ijv, -g.J.t * ijv * ijv};
  } else {
    g.idet = recip(det);
    g.J = dsqrt(det);
    iJ = recip(g.J);
    fix<PS>(g.idet, pr, S_IDET);
    fix<PS>(g.J, pr, S_J);
    fix<PS>(iJ, pr, S_IJ);
  }
  g.g00 = a22 * g.idet;
  g.g01 = -(a12 * g.idet);
  g.g11 = a11 * g.idet;
  fix<PS>(g.g00, pr, S_G00);
  fix<PS>(g.g01, pr, S_G01);
  fix<PS>(g.g11, pr, S_G11);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    g.au1[i] = g.a1[i] * g.g00 + g.a2[i] * g.g01;
    g.au2[i] = g.a1[i] * g.g01 + g.a2[i] * g.g11;
  }
  fixa<PS>(g.au1, pr, S_AU1);
  fixa<PS>(g.au2, pr, S_AU2);
  g.n[0] = (g.a1[1] * g.a2[2] - g.a1[2] * g.a2[1]) * iJ;
  g.n[1] = (g.a1[2] * g.a2[0] - g.a1[0] * g.a2[2]) * iJ;
  g.n[2] = (g.a1[0] * g.a2[1] - g.a1[1] * g.a2[0]) * iJ;
  fixa<PS>(g.n, pr, S_N);
}

// One Gauss point of the area element (GeoDynStress.jl:111-159,
// FiniteElement.jl:284-313).  PRIMAL ends after the stresses, with the point's
// values in pr; PLAIN and TANGENT add the lane's rows (nodes a_lo .. a_lo +
// NPP - 1) of the tangent column to acc and its own row's residual to racc.
template <int MO, class N, int PS, int P, int NR>
__device__ __forceinline__ void gp_eval(const double *X, const double *C, const double *rec,
                                        int a0, int d0, int i0, int a_lo, int col, double sX0,
                                        double sC0, const PhysParams &ph, double azv, double *pr,
                                        double (&acc)[NR], double &racc) {
  using L = Lay<MO>;
  constexpr int NDF = L::NDF, NPP = (9 + P - 1) / P;
  const double w = rec[0];
  const double *bN = rec + 1, *bN1 = rec + 10, *bN2 = rec + 19;
  const double *b11 = rec + 28, *b22 = rec + 37, *b12 = rec + 46;

  Geo<N> g;
  geometry<N, PS>(X, bN1, bN2, a0, i0, sX0, g, pr);
  // second derivatives of position: x_{,11}, x_{,22}, x_{,12}
  N x11[3], x22[3], x12[3];
  {
    const double t11 = sX0 * b11[a0], t22 = sX0 * b22[a0], t12 = sX0 * b12[a0];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double s11 = 0, s22 = 0, s12 = 0;
#pragma unroll
      for (int a = 0; a < 9; ++a) {
        s11 = fma(X[i * 9 + a], b11[a], s11);
        s22 = fma(X[i * 9 + a], b22[a], s22);
        s12 = fma(X[i * 9 + a], b12[a], s12);
      }
      x11[i] = mkN<N>(s11, i == i0 ? t11 : 0.0);
      x22[i] = mkN<N>(s22, i == i0 ? t22 : 0.0);
      x12[i] = mkN<N>(s12, i == i0 ? t12 : 0.0);
    }
  }
  fixa<PS>(x11, pr, S_X11);
  fixa<PS>(x22, pr, S_X22);
  fixa<PS>(x12, pr, S_X12);
  // Christoffel symbols Γ[v][μ] = x_{,v} · a^μ (v = 11, 22, 12)
  N G0[3], G1[3];
  G0[0] = x11[0] * g.au1[0] + x11[1] * g.au1[1] + x11[2] * g.au1[2];
  G0[1] = x22[0] * g.au1[0] + x22[1] * g.au1[1] + x22[2] * g.au1[2];
  G0[2] = x12[0] * g.au1[0] + x12[1] * g.au1[1] + x12[2] * g.au1[2];
  G1[0] = x11[0] * g.au2[0] + x11[1] * g.au2[1] + x11[2] * g.au2[2];
  G1[1] = x22[0] * g.au2[0] + x22[1] * g.au2[1] + x22[2] * g.au2[2];
  G1[2] = x12[0] * g.au2[0] + x12[1] * g.au2[1] + x12[2] * g.au2[2];
  fixa<PS>(G0, pr, S_GA0);
  fixa<PS>(G1, pr, S_GA1);
  // curvature b_{αβ}, b^{αβ}, H, K
  N bb[3];
  bb[0] = x11[0] * g.n[0] + x11[1] * g.n[1] + x11[2] * g.n[2];
  bb[1] = x22[0] * g.n[0] + x22[1] * g.n[1] + x22[2] * g.n[2];
  bb[2] = x12[0] * g.n[0] + x12[1] * g.n[1] + x12[2] * g.n[2];
  fixa<PS>(bb, pr, S_BB);
  const N bb11 = bb[0], bb22 = bb[1], bb12 = bb[2];
  N tb00 = g.g00 * bb11 + g.g01 * bb12, tb01 = g.g00 * bb12 + g.g01 * bb22;
  N tb10 = g.g01 * bb11 + g.g11 * bb12, tb11 = g.g01 * bb12 + g.g11 * bb22;
  N Bm[4];
  Bm[0] = tb00 * g.g00 + tb01 * g.g01;
  Bm[1] = tb00 * g.g01 + tb01 * g.g11;
  Bm[2] = tb10 * g.g00 + tb11 * g.g01;
  Bm[3] = tb10 * g.g01 + tb11 * g.g11;
  fixa<PS>(Bm, pr, S_B);
  const N B00 = Bm[0], B01 = Bm[1], B10 = Bm[2], B11 = Bm[3];
  N H = 0.5 * (g.g00 * bb11 + g.g01 * bb12 + g.g01 * bb12 + g.g11 * bb22);
  N Kc = (bb11 * bb22 - bb12 * bb12) * g.idet;
  fix<PS>(H, pr, S_H);
  fix<PS>(Kc, pr, S_KC);

  // fields: v, v_{,α}, λ, and (EUL/ALE) v^m, v^m_{,α}, p^m
  N v[3], dv0[3], dv1[3];
  const double tN = sC0 * bN[a0], tN1 = sC0 * bN1[a0], tN2 = sC0 * bN2[a0];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s = 0, s1 = 0, s2 = 0;
    if (i < L::NV) {  // an absent vz is zero (get_dof_cps, GeoDynStress.jl:190-205)
#pragma unroll
      for (int a = 0; a < 9; ++a) {
        const double c = C[i * 9 + a];
        s = fma(c, bN[a], s);
        s1 = fma(c, bN1[a], s1);
        s2 = fma(c, bN2[a], s2);
      }
    }
    const bool sd = (d0 == i) && i < L::NV;
    v[i] = mkN<N>(s, sd ? tN : 0.0);
    dv0[i] = mkN<N>(s1, sd ? tN1 : 0.0);
    dv1[i] = mkN<N>(s2, sd ? tN2 : 0.0);
  }
  fixa<PS>(v, pr, S_V);
  fixa<PS>(dv0, pr, S_DV0);
  fixa<PS>(dv1, pr, S_DV1);
  N lam;
  {
    double s = 0;
#pragma unroll
    for (int a = 0; a < 9; ++a) s = fma(C[L::LAMI * 9 + a], bN[a], s);
    lam = mkN<N>(s, d0 == L::LAMI ? tN : 0.0);
  }
  fix<PS>(lam, pr, S_LAM);
  N vm[3], dvm0[3], dvm1[3], pm = mkN<N>(0, 0);
  if constexpr (L::HAS_VM) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double s = 0, s1 = 0, s2 = 0;
#pragma unroll
      for (int a = 0; a < 9; ++a) {
        const double c = C[(3 + i) * 9 + a];
        s = fma(c, bN[a], s);
        s1 = fma(c, bN1[a], s1);
        s2 = fma(c, bN2[a], s2);
      }
      const bool sd = (d0 == 3 + i);
      vm[i] = mkN<N>(s, sd ? tN : 0.0);
      dvm0[i] = mkN<N>(s1, sd ? tN1 : 0.0);
      dvm1[i] = mkN<N>(s2, sd ? tN2 : 0.0);
    }
    fixa<PS>(vm, pr, S_VM);
    fixa<PS>(dvm0, pr, S_DVM0);
    fixa<PS>(dvm1, pr, S_DVM1);
  }
  if constexpr (L::HAS_PM) {
    double s = 0;
#pragma unroll
    for (int a = 0; a < 9; ++a) s = fma(C[L::PMI * 9 + a], bN[a], s);
    pm = mkN<N>(s, d0 == L::PMI ? tN : 0.0);
    fix<PS>(pm, pr, S_PM);
  }

  // stresses (GeoDynStress.jl:148-159), Voigt [11, 22, 2*12] / [11, 22, 12+21]
  N sg[3], Mv[3], sgm[3];
  {
    N w0[3], w1[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      w0[i] = dv0[i] * g.g00 + dv1[i] * g.g01;
      w1[i] = dv0[i] * g.g01 + dv1[i] * g.g11;
    }
    N p00 = g.au1[0] * w0[0] + g.au1[1] * w0[1] + g.au1[2] * w0[2];
    N p01 = g.au1[0] * w1[0] + g.au1[1] * w1[1] + g.au1[2] * w1[2];
    N p10 = g.au2[0] * w0[0] + g.au2[1] * w0[1] + g.au2[2] * w0[2];
    N p11 = g.au2[0] * w1[0] + g.au2[1] * w1[1] + g.au2[2] * w1[2];
    N c0 = ph.kb * (H * H) - ph.kg * Kc;
    N lc = c0 + lam;
    N h2 = (2.0 * ph.kb) * H;
    sg[0] = g.g00 * lc - B00 * h2 + (2.0 * ph.zv) * p00;
    sg[1] = g.g11 * lc - B11 * h2 + (2.0 * ph.zv) * p11;
    sg[2] = 2.0 * (g.g01 * lc - B01 * h2 + ph.zv * (p01 + p10));
    N kk = H * (ph.kb + 2.0 * ph.kg);
    Mv[0] = g.g00 * kk - B00 * ph.kg;
    Mv[1] = g.g11 * kk - B11 * ph.kg;
    Mv[2] = (g.g01 * kk - B01 * ph.kg) + (g.g01 * kk - B10 * ph.kg);
    if constexpr (L::HAS_PM) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        w0[i] = dvm0[i] * g.g00 + dvm1[i] * g.g01;
        w1[i] = dvm0[i] * g.g01 + dvm1[i] * g.g11;
      }
      N q00 = g.au1[0] * w0[0] + g.au1[1] * w0[1] + g.au1[2] * w0[2];
      N q01 = g.au1[0] * w1[0] + g.au1[1] * w1[1] + g.au1[2] * w1[2];
      N q10 = g.au2[0] * w0[0] + g.au2[1] * w0[1] + g.au2[2] * w0[2];
      N q11 = g.au2[0] * w1[0] + g.au2[1] * w1[1] + g.au2[2] * w1[2];
      sgm[0] = g.g00 * c0 - B00 * h2 + (2.0 * ph.zv) * q00;
      sgm[1] = g.g11 * c0 - B11 * h2 + (2.0 * ph.zv) * q11;
      sgm[2] = 2.0 * (g.g01 * c0 - B01 * h2 + ph.zv * (q01 + q10));
      fixa<PS>(sgm, pr, S_SGM);
    }
  }
  fixa<PS>(sg, pr, S_SG);
  fixa<PS>(Mv, pr, S_MV);
  if constexpr (PS == PRIMAL) return;
  else {
    // per-Gauss-point factors of the residual rows
    const N wJ = w * g.J;
    N A1[3], A2[3], Nn[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      A1[i] = wJ * g.a1[i];
      A2[i] = wJ * g.a2[i];
      Nn[i] = wJ * g.n[i];
    }
    N tr = g.au1[0] * dv0[0] + g.au1[1] * dv0[1] + g.au1[2] * dv0[2] +
           g.au2[0] * dv1[0] + g.au2[1] * dv1[1] + g.au2[2] * dv1[2];
    const N ql = w * (g.J * tr - azv * lam);
    N gm[3], qp = mkN<N>(0, 0), pmJ = mkN<N>(0, 0);
    if constexpr (MO == EUL) {
      N nv = g.n[0] * v[0] + g.n[1] * v[1] + g.n[2] * v[2];
#pragma unroll
      for (int i = 0; i < 3; ++i) gm[i] = (ph.am * wJ) * (vm[i] - g.n[i] * nv);
    }
    if constexpr (L::HAS_PM) {
      N nd = g.n[0] * (vm[0] - v[0]) + g.n[1] * (vm[1] - v[1]) + g.n[2] * (vm[2] - v[2]);
      qp = -(wJ * nd) - (w * azv) * pm;
      pmJ = pm * wJ;
    }

// local row j NDF + r of this lane is row a NDF + r of the element
#define ADD(r, val)                                  \
  do {                                               \
    const N v_ = (val);                              \
    acc[j * NDF + (r)] += v_.t;                      \
    racc += (a * NDF + (r) == col) ? v_.v : 0.0;     \
  } while (0)

#pragma unroll
    for (int j = 0; j < NPP; ++j) {
      const int a = a_lo + j;
      if (9 % P != 0 && a >= 9) break;
      const double n1 = bN1[a], n2 = bN2[a], na = bN[a];
      N s1 = n1 * sg[0] + (0.5 * n2) * sg[2];
      N s2 = n2 * sg[1] + (0.5 * n1) * sg[2];
      N nn0 = mkN<N>(b11[a], 0.0) - (n1 * G0[0] + n2 * G1[0]);
      N nn1 = mkN<N>(b22[a], 0.0) - (n1 * G0[1] + n2 * G1[1]);
      N nn2 = mkN<N>(b12[a], 0.0) - (n1 * G0[2] + n2 * G1[2]);
      N m = nn0 * Mv[0] + nn1 * Mv[1] + nn2 * Mv[2];
#pragma unroll
      for (int i = 0; i < L::NV; ++i) {
        N rv = A1[i] * s1 + A2[i] * s2 + Nn[i] * m;
        if (ph.pn != 0.0) rv = rv - (ph.pn * na) * Nn[i];
        ADD(i, rv);
      }
      ADD(L::LAMI, na * ql);
      if constexpr (MO == EUL) {
#pragma unroll
        for (int i = 0; i < 3; ++i) ADD(3 + i, na * gm[i]);
      }
      if constexpr (L::HAS_PM) {
        N t1 = n1 * sgm[0] + (0.5 * n2) * sgm[2];
        N t2 = n2 * sgm[1] + (0.5 * n1) * sgm[2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          N rm = A1[i] * t1 + A2[i] * t2;
          if constexpr (MO == ALEVB) rm = rm + Nn[i] * m;
          rm = rm - na * (pmJ * g.n[i]);
          ADD(3 + i, rm);
        }
        ADD(L::PMI, na * qp);
      }
    }
#undef ADD
  }
}

// LIST = false: elements 0..numel-1, outputs K_el and r_el per element, in
// two passes (PRIMAL, then TANGENT).  LIST = true: the numel elements
// elist[0..numel-1], output r_el only, at the element's position in the list
// (PLAIN: the tangents are dead code there).  P lanes share each tangent
// column (lane part p takes the rows of nodes p NPP .. p NPP + NPP - 1, NPP =
// ceil(9 / P)); the part holding the column's own row writes its r_el entry.
template <int MO, bool LIST = false, int P = LIST ? 1 : Lay<MO>::PARTS>
__device__ __forceinline__ void area_body(
    int64_t bid, int64_t numel, int64_t numnp, const int32_t *__restrict__ ix,
    const int32_t *__restrict__ uel, const uint8_t *__restrict__ active,
    const double *__restrict__ xms, const double *__restrict__ cps,
    const double *__restrict__ ufns, const double *__restrict__ dbt, PhysParams ph,
    double dt, double *__restrict__ kel, double *__restrict__ rel,
    const int32_t *__restrict__ elist = nullptr) {
  using L = Lay<MO>;
  auto EL = [&](int64_t i) -> int64_t { return LIST ? (int64_t)elist[i] : i; };
  constexpr int NDF = L::NDF, NE = L::NE;
  constexpr int NPP = (9 + P - 1) / P, NR = NPP * NDF, LPE = NE * P;
  constexpr int MAXE = 256 / LPE + 2;  // elements a 256-lane block can touch
  constexpr bool STAGED = !LIST;
  constexpr int NS = L::HAS_VM ? N_SLOTS : S_VM;  // slots this motion uses
  __shared__ double sX[MAXE][3 * 9];
  __shared__ double sC[MAXE][NDF * 9];
  __shared__ double sB[MAXE][9 * RECP];
  __shared__ double sP[STAGED ? MAXE * 9 : 1][STAGED ? NS : 1];

  const int64_t f0 = bid * 256;
  const int64_t e_lo = f0 / LPE;
  int64_t e_hi = (f0 + 255) / LPE;
  if (e_hi > numel - 1) e_hi = numel - 1;
  const int nle = (int)(e_hi - e_lo + 1);
  // stage element state: positions and control points of the 9 nodes
  for (int k = threadIdx.x; k < nle * 9 * (3 + NDF); k += 256) {
    const int le = k / (9 * (3 + NDF)), rem = k % (9 * (3 + NDF));
    const int comp = rem / 9, a = rem % 9;
    const int64_t node = ix[EL(e_lo + le) * 9 + a];
    if (comp < 3)
      sX[le][comp * 9 + a] = xms[node + numnp * comp];
    else
      sC[le][(comp - 3) * 9 + a] = cps[node + numnp * (comp - 3)];
  }
  // stage the 9 Gauss-point basis records of each element
  for (int k = threadIdx.x; k < nle * 9 * RECP; k += 256) {
    const int le = k / (9 * RECP), rem = k % (9 * RECP);
    sB[le][rem] = ufns[(int64_t)uel[EL(e_lo + le)] * 9 * RECP + rem];
  }
  __syncthreads();

  const double azv = ph.adb / ph.zv;
  if constexpr (STAGED) {
    // primal pass: one lane per (element, Gauss point) of the block
    for (int k = threadIdx.x; k < nle * 9; k += 256) {
      const int le = k / 9, gp = k % 9;
      double none[1];
      double rnone = 0.0;
      gp_eval<MO, R, PRIMAL, 1, 1>(sX[le], sC[le], &sB[le][gp * RECP], 0, -1, -1, 0, -1, 0.0,
                                   0.0, ph, azv, sP[k], none, rnone);
    }
    __syncthreads();
  }

  const int64_t f = f0 + threadIdx.x;
  const bool valid = f < numel * LPE;
  const int64_t e = valid ? f / LPE : e_hi;
  const int64_t eg = EL(e);
  const int part = (int)(f % LPE) / NE;
  const int col = (int)(f % NE);
  const int a_lo = part * NPP;  // first node of this lane's rows
  const int le = (int)(e - e_lo);
  const int a0 = col / NDF, d0 = col % NDF;
  const bool own_row = a0 >= a_lo && a0 < a_lo + NPP;
  const int64_t node0 = ix[eg * 9 + a0];
  const bool act = active[d0 + NDF * node0] != 0;
  const double sC0 = act ? 1.0 : 0.0;
  const bool mot = (L::MOT0 >= 0) && d0 >= L::MOT0 && d0 < L::MOT0 + 3;
  const int i0 = mot ? d0 - L::MOT0 : -1;
  const double sX0 = (act && mot) ? dt : 0.0;
  const double *C = sC[le];

  double acc[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) acc[k] = 0.0;
  double racc = 0.0;

  for (int gp = 0; gp < 9; ++gp) {
    if constexpr (STAGED)
      gp_eval<MO, D, TANGENT, P>(sX[le], C, &sB[le][gp * RECP], a0, d0, i0, a_lo, col, sX0,
                                   sC0, ph, azv, sP[le * 9 + gp], acc, racc);
    else
      gp_eval<MO, D, PLAIN, P>(sX[le], C, &sB[le][gp * RECP], a0, d0, i0, a_lo, col, sX0, sC0,
                                 ph, azv, nullptr, acc, racc);
  }

  // Dohrmann--Bochev projection (FiniteElement.jl:315-327): linear in λ, p^m
  {
    const double *T = dbt + (int64_t)uel[eg] * 81;
    if (d0 == L::LAMI && act) {
#pragma unroll
      for (int j = 0; j < NPP; ++j)
        if (9 % P == 0 || a_lo + j < 9) acc[j * NDF + L::LAMI] += T[(a_lo + j) * 9 + a0] * azv;
    }
    if constexpr (L::HAS_PM) {
      if (d0 == L::PMI && act) {
#pragma unroll
        for (int j = 0; j < NPP; ++j)
          if (9 % P == 0 || a_lo + j < 9) acc[j * NDF + L::PMI] += T[(a_lo + j) * 9 + a0] * azv;
      }
    }
    if (own_row && (d0 == L::LAMI || (L::HAS_PM && d0 == L::PMI))) {
      double s = 0;
      for (int b = 0; b < 9; ++b) s = fma(T[a0 * 9 + b], C[d0 * 9 + b], s);
      racc += s * azv;
    }
  }

  if (valid) {
    if constexpr (!LIST) {
      double *K = kel + e * (int64_t)NE * NE;
#pragma unroll
      for (int j = 0; j < NPP; ++j)
        if (9 % P == 0 || a_lo + j < 9) {
#pragma unroll
          for (int r = 0; r < NDF; ++r) K[((a_lo + j) * NDF + r) * NE + col] = acc[j * NDF + r];
        }
    }
    if (own_row) rel[e * NE + col] = racc;
  }
}

// Neumann edge elements (FiniteElement.jl:338-400): 3 Gauss points, v rows only.
template <int MO>
__device__ __forceinline__ void edge_body(
    int64_t bid, int64_t n_items, int64_t numnp, const EdgeItem *__restrict__ items,
    const int32_t *__restrict__ ix, const uint8_t *__restrict__ active,
    const double *__restrict__ xms, const double *__restrict__ cps,
    const double *__restrict__ bufns, PhysParams ph, double time, double dt,
    double *__restrict__ kel, double *__restrict__ rel) {
  using L = Lay<MO>;
  constexpr int NDF = L::NDF, NE = L::NE, MAXE = L::MAXE;
  __shared__ double sX[MAXE][3 * 9];
  __shared__ double sB[MAXE][3 * RECP];

  const int64_t f0 = bid * 256;
  const int64_t k_lo = f0 / NE;
  int64_t k_hi = (f0 + 255) / NE;
  if (k_hi > n_items - 1) k_hi = n_items - 1;
  const int nle = (int)(k_hi - k_lo + 1);
  for (int k = threadIdx.x; k < nle * 27; k += 256) {
    const int le = k / 27, rem = k % 27, comp = rem / 9, a = rem % 9;
    const int64_t node = ix[(int64_t)items[k_lo + le].el * 9 + a];
    sX[le][comp * 9 + a] = xms[node + numnp * comp];
  }
  for (int k = threadIdx.x; k < nle * 3 * RECP; k += 256) {
    const int le = k / (3 * RECP), rem = k % (3 * RECP);
    sB[le][rem] = bufns[(int64_t)items[k_lo + le].rec * RECP + rem];
  }
  __syncthreads();

  const int64_t f = f0 + threadIdx.x;
  const bool valid = f < n_items * NE;
  const int64_t k = valid ? f / NE : k_hi;
  const int col = (int)(f % NE);
  const int le = (int)(k - k_lo);
  const EdgeItem it = items[k];
  const int a0 = col / NDF, d0 = col % NDF;
  const int64_t node0 = ix[(int64_t)it.el * 9 + a0];
  const bool act = active[d0 + NDF * node0] != 0;
  const bool mot = (L::MOT0 >= 0) && d0 >= L::MOT0 && d0 < L::MOT0 + 3;
  const int i0 = mot ? d0 - L::MOT0 : -1;
  const double sX0 = (act && mot) ? dt : 0.0;
  const double *X = sX[le];

  double acc[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) acc[q] = 0.0;
  double racc = 0.0;
  const double mval = (it.ntype == MOMENT) ? it.nval * fmin(time / ph.bend_tm, 1.0) : 0.0;

  for (int gp = 0; gp < 3; ++gp) {
    const double *rec = &sB[le][gp * RECP];
    const double w = rec[0];
    const double *bN = rec + 1, *bN1 = rec + 10, *bN2 = rec + 19;
    Geo<> g;
    geometry(X, bN1, bN2, a0, i0, sX0, g);
    // edge frame (calc_τ_ν): τ = ±a_1 / ±a_2, normalised by the conjugating
    // dot product, whose complex-step derivative is zero: only τ's own
    // variation is carried, divided by the primal norm.
    D tau[3];
    const double sgn = (it.bdry == TOP || it.bdry == LEFT) ? -1.0 : 1.0;
    const bool along1 = (it.bdry == BOTTOM || it.bdry == TOP);
#pragma unroll
    for (int i = 0; i < 3; ++i) tau[i] = sgn * (along1 ? g.a1[i] : g.a2[i]);
    const double tn = sqrt(tau[0].v * tau[0].v + tau[1].v * tau[1].v + tau[2].v * tau[2].v);
    const double itn = 1.0 / tn;
#pragma unroll
    for (int i = 0; i < 3; ++i) tau[i] = itn * tau[i];
    D nu[3];
    nu[0] = tau[1] * g.n[2] - tau[2] * g.n[1];
    nu[1] = tau[2] * g.n[0] - tau[0] * g.n[2];
    nu[2] = tau[0] * g.n[1] - tau[1] * g.n[0];
    D q0 = g.au1[0] * tau[0] + g.au1[1] * tau[1] + g.au1[2] * tau[2];
    D q1 = g.au2[0] * tau[0] + g.au2[1] * tau[1] + g.au2[2] * tau[2];
    const D JGw = w * recip(dsqrt(q0 * q0 + q1 * q1));
    if (it.ntype == STRETCH || it.ntype == SHEAR) {
      D fv[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) fv[i] = (it.nval * JGw) * (it.ntype == STRETCH ? nu[i] : tau[i]);
#pragma unroll
      for (int a = 0; a < 9; ++a)
#pragma unroll
        for (int i = 0; i < L::NV; ++i) {
          const D val = -(bN[a] * fv[i]);
          acc[a * 3 + i] += val.t;
          racc += (a * NDF + i == col) ? val.v : 0.0;
        }
    } else {  // MOMENT (validated on the host: F_BEND only)
      D nu0 = g.au1[0] * nu[0] + g.au1[1] * nu[1] + g.au1[2] * nu[2];
      D nu1 = g.au2[0] * nu[0] + g.au2[1] * nu[1] + g.au2[2] * nu[2];
      const D s = mval * JGw;
#pragma unroll
      for (int a = 0; a < 9; ++a) {
        const D dn = bN1[a] * nu0 + bN2[a] * nu1;
#pragma unroll
        for (int i = 0; i < L::NV; ++i) {
          const D val = -((g.n[i] * dn) * s);
          acc[a * 3 + i] += val.t;
          racc += (a * NDF + i == col) ? val.v : 0.0;
        }
      }
    }
  }
  if (valid) {
    double *K = kel + k * (int64_t)NE * NE;
#pragma unroll
    for (int row = 0; row < NE; ++row) {
      const int a = row / NDF, d = row % NDF;
      K[row * NE + col] = (d < L::NV) ? acc[a * 3 + d] : 0.0;
    }
    rel[k * NE + col] = racc;
  }
}

template <int MO, int P = Lay<MO>::PARTS>
__global__ __launch_bounds__(256) void k_area(
    int64_t numel, int64_t numnp, const int32_t *__restrict__ ix,
    const int32_t *__restrict__ uel, const uint8_t *__restrict__ active,
    const double *__restrict__ xms, const double *__restrict__ cps,
    const double *__restrict__ ufns, const double *__restrict__ dbt, PhysParams ph,
    double dt, double *__restrict__ kel, double *__restrict__ rel) {
  area_body<MO, false, P>(blockIdx.x, numel, numnp, ix, uel, active, xms, cps, ufns, dbt, ph, dt,
                          kel, rel);
}

template <int MO>
__global__ __launch_bounds__(256) void k_edge(
    int64_t n_items, int64_t numnp, const EdgeItem *__restrict__ items,
    const int32_t *__restrict__ ix, const uint8_t *__restrict__ active,
    const double *__restrict__ xms, const double *__restrict__ cps,
    const double *__restrict__ bufns, PhysParams ph, double time, double dt,
    double *__restrict__ kel, double *__restrict__ rel) {
  edge_body<MO>(blockIdx.x, n_items, numnp, items, ix, active, xms, cps, bufns, ph, time, dt,
                kel, rel);
}

// Residuals of a list of area elements (the pull force's adjacent elements).
template <int MO>
__global__ __launch_bounds__(256) void k_area_list(
    int64_t n, int64_t numnp, const int32_t *__restrict__ ix, const int32_t *__restrict__ uel,
    const uint8_t *__restrict__ active, const double *__restrict__ xms,
    const double *__restrict__ cps, const double *__restrict__ ufns,
    const double *__restrict__ dbt, PhysParams ph, const int32_t *__restrict__ elist,
    double *__restrict__ rel) {
  area_body<MO, true>(blockIdx.x, n, numnp, ix, uel, active, xms, cps, ufns, dbt, ph, 0.0, nullptr,
                      rel, elist);
}

// Pull force and pulled-element position (PullForce.jl:61-80,
// Analysis.jl:81-85).  Threads 0-2: f[c] = sum over pulled nodes k of
// (sum over adjacent elements, in order, of the v_c row of the adjacent
// element's local node that is pulled node k).  Threads 3-5: mean of xms over
// the pulled element's 9 nodes.
__global__ void k_pull_reduce(int n, int ne, int ndf, const int8_t *__restrict__ slot,
                              const double *__restrict__ rel, const int32_t *__restrict__ pull_ix,
                              const double *__restrict__ xms, int64_t numnp,
                              double *__restrict__ out) {
  const int t = threadIdx.x;
  if (t < 3) {
    double f = 0.0;
    for (int k = 0; k < 9; ++k) {
      double s = 0.0;
      for (int q = 0; q < n; ++q)
        for (int a = 0; a < 9; ++a)
          if (slot[q * 9 + a] == k) s += rel[(int64_t)q * ne + a * ndf + t];
      f += s;
    }
    out[3 + t] = f;
  } else if (t < 6) {
    const int c = t - 3;
    double s = 0.0;
    for (int a = 0; a < 9; ++a) s += xms[pull_ix[a] + numnp * c];
    out[c] = s / 9.0;
  }
}

// generate_output (Output.jl:30-118): one thread per sample (i1, i2) of the
// Z1 x Z2 grid (0-based).  Interior samples are the area Gauss points
// (i = 3 (el - 1) + gp, shifted by one), the outer rows/columns the boundary
// Gauss points of BOTTOM (i2 = 0), RIGHT (i1 = Z1 - 1), TOP (i2 = Z2 - 1) and
// LEFT (i1 = 0) in boundary-element order, the four corners the corner
// records.  Each sample is Σ_a state[node_a] N_a over the element's 9 nodes.
struct OutGrid {
  int64_t Z1, Z2, num1el, numnp;
  int32_t ndf;
  int32_t crnr_el[4];
  int64_t boff[4];
};

__global__ __launch_bounds__(256) void k_generate_output(
    OutGrid g, const int32_t *__restrict__ ix, const int32_t *__restrict__ uel,
    const double *__restrict__ ufns, const double *__restrict__ bufns,
    const int32_t *__restrict__ obel, const double *__restrict__ crnr,
    const double *__restrict__ xms, const double *__restrict__ cps, double *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t npts = g.Z1 * g.Z2;
  if (t >= npts) return;
  const int64_t i1 = t % g.Z1, i2 = t / g.Z1;
  const bool e1 = (i1 == 0 || i1 == g.Z1 - 1), e2 = (i2 == 0 || i2 == g.Z2 - 1);
  int64_t el;
  const double *N;
  if (!e1 && !e2) {
    const int64_t el1 = (i1 - 1) / 3, el2 = (i2 - 1) / 3;
    const int gp = (int)((i1 - 1) % 3 + 3 * ((i2 - 1) % 3));
    el = el1 + g.num1el * el2;
    N = ufns + ((int64_t)uel[el] * 9 + gp) * RECP + 1;
  } else if (e1 && e2) {
    const int k = (i1 == 0 ? 0 : 1) + (i2 == 0 ? 0 : 2);
    el = g.crnr_el[k];
    N = crnr + k * RECP + 1;
  } else {
    int b;
    int64_t k;
    if (i2 == 0) { b = 0; k = i1 - 1; }
    else if (i1 == g.Z1 - 1) { b = 1; k = i2 - 1; }
    else if (i2 == g.Z2 - 1) { b = 2; k = i1 - 1; }
    else { b = 3; k = i2 - 1; }
    const int32_t *pr = obel + 2 * (g.boff[b] + k / 3);
    el = pr[0];
    N = bufns + ((int64_t)pr[1] + k % 3) * RECP + 1;
  }
  const int32_t *nodes = ix + el * 9;
  for (int c = 0; c < 3 + g.ndf; ++c) {
    const double *src = c < 3 ? xms + g.numnp * c : cps + g.numnp * (c - 3);
    double s = 0.0;
    for (int a = 0; a < 9; ++a) s = fma(src[nodes[a]], N[a], s);
    out[t + npts * c] = s;
  }
}

// Batched forms: block -> patch through the segment map; patches that are
// off skip the launch.
template <int MO>
__global__ __launch_bounds__(256) void kb_area(Seg sg, const PatchDev *__restrict__ pd,
                                               const uint8_t *__restrict__ on,
                                               const double *__restrict__ dts) {
  const int p = sg.map[blockIdx.x];
  if (!on[p]) return;
  const PatchDev &d = pd[p];
  area_body<MO>(blockIdx.x - sg.blk0[p], d.numel, d.numnp, d.ix, d.uel, d.active, d.xms, d.cps,
                d.ufns, d.dbt, d.ph, dts[p], d.kel, d.rel);
}

template <int MO>
__global__ __launch_bounds__(256) void kb_edge(Seg sg, const PatchDev *__restrict__ pd,
                                               const uint8_t *__restrict__ on,
                                               const double *__restrict__ times,
                                               const double *__restrict__ dts) {
  const int p = sg.map[blockIdx.x];
  if (!on[p]) return;
  const PatchDev &d = pd[p];
  edge_body<MO>(blockIdx.x - sg.blk0[p], d.n_edges, d.numnp, d.edges, d.ix, d.active, d.xms,
                d.cps, d.bufns, d.ph, times[p], dts[p], d.kel + d.numel * (int64_t)d.ne * d.ne,
                d.rel + d.numel * (int64_t)d.ne);
}

// Deterministic gather of element blocks into the node-block matrix of the
// extended system: entries of prescribed DOFs become identity rows/columns.
__device__ __forceinline__ void gather_bsr_at(
    int64_t t, int64_t nnzb, int ndf, int ne, const int32_t *__restrict__ brow,
    const int32_t *__restrict__ colidx, const uint8_t *__restrict__ active,
    const int64_t *__restrict__ gptr, const int64_t *__restrict__ gsrc,
    const double *__restrict__ kel, double *__restrict__ bsr) {
  const int bb = ndf * ndf;
  if (t >= nnzb * bb) return;
  const int64_t blk = t / bb;
  const int rc = (int)(t % bb), r = rc / ndf, c = rc % ndf;
  const int64_t p = brow[blk], q = colidx[blk];
  const bool ar = active[r + (int64_t)ndf * p] != 0, ac = active[c + (int64_t)ndf * q] != 0;
  double s = 0.0;
  if (ar && ac) {
    const int64_t off = (int64_t)r * ne + c;
    for (int64_t k = gptr[blk]; k < gptr[blk + 1]; ++k) s += kel[gsrc[k] + off];
  } else if (!ar && p == q && r == c) {
    s = 1.0;
  }
  bsr[t] = s;
}

__device__ __forceinline__ void gather_r_at(int64_t t, int64_t nx, int ndf,
                                            const uint8_t *__restrict__ active,
                                            const int64_t *__restrict__ rptr,
                                            const int64_t *__restrict__ rsrc,
                                            const double *__restrict__ rel, double *__restrict__ r) {
  if (t >= nx) return;
  const int64_t p = t / ndf;
  const int d = (int)(t % ndf);
  double s = 0.0;
  if (active[t]) {
    for (int64_t k = rptr[p]; k < rptr[p + 1]; ++k) s += rel[rsrc[k] + d];
  }
  r[t] = s;
}

// x += Δt u[:, motion] at node p (update_xms!, FiniteElement.jl:408-423); u is
// column-major numnp x ndf (the predictor passes cps).
__device__ __forceinline__ void predict_at(int64_t p, int64_t numnp, const int32_t *mm,
                                           const double *__restrict__ u, double dt,
                                           double *__restrict__ xms) {
  if (p >= numnp) return;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (mm[i] > 0) xms[p + numnp * i] += dt * u[p + numnp * (mm[i] - 1)];
}

__global__ __launch_bounds__(256) void k_gather_bsr(
    int64_t nnzb, int ndf, int ne, const int32_t *__restrict__ brow,
    const int32_t *__restrict__ colidx, const uint8_t *__restrict__ active,
    const int64_t *__restrict__ gptr, const int64_t *__restrict__ gsrc,
    const double *__restrict__ kel, double *__restrict__ bsr) {
  gather_bsr_at((int64_t)blockIdx.x * 256 + threadIdx.x, nnzb, ndf, ne, brow, colidx, active, gptr,
                gsrc, kel, bsr);
}

__global__ __launch_bounds__(256) void k_gather_r(
    int64_t nx, int ndf, const uint8_t *__restrict__ active,
    const int64_t *__restrict__ rptr, const int64_t *__restrict__ rsrc,
    const double *__restrict__ rel, double *__restrict__ r) {
  gather_r_at((int64_t)blockIdx.x * 256 + threadIdx.x, nx, ndf, active, rptr, rsrc, rel, r);
}

__global__ void k_predict(int64_t numnp, int m0, int m1, int m2, const double *__restrict__ u,
                          double dt, double *__restrict__ xms) {
  const int32_t mm[3] = {m0, m1, m2};
  predict_at((int64_t)blockIdx.x * 256 + threadIdx.x, numnp, mm, u, dt, xms);
}

__global__ __launch_bounds__(256) void kb_gather_bsr(Seg sg, const PatchDev *__restrict__ pd,
                                                     const uint8_t *__restrict__ on) {
  const int p = sg.map[blockIdx.x];
  if (!on[p]) return;
  const PatchDev &d = pd[p];
  gather_bsr_at((blockIdx.x - sg.blk0[p]) * 256 + threadIdx.x, d.nnzb, d.ndf, d.ne, d.brow,
                d.colidx, d.active, d.gptr, d.gsrc, d.kel, d.bsr);
}

__global__ __launch_bounds__(256) void kb_gather_r(Seg sg, const PatchDev *__restrict__ pd,
                                                   const uint8_t *__restrict__ on) {
  const int p = sg.map[blockIdx.x];
  if (!on[p]) return;
  const PatchDev &d = pd[p];
  gather_r_at((blockIdx.x - sg.blk0[p]) * 256 + threadIdx.x, d.nx, d.ndf, d.active, d.rptr,
              d.rsrc, d.rel, d.r);
}

__global__ __launch_bounds__(256) void kb_predict(Seg sg, const PatchDev *__restrict__ pd,
                                                  const uint8_t *__restrict__ on,
                                                  const double *__restrict__ dts) {
  const int p = sg.map[blockIdx.x];
  if (!on[p]) return;
  const PatchDev &d = pd[p];
  predict_at((blockIdx.x - sg.blk0[p]) * 256 + threadIdx.x, d.numnp, d.mm, d.cps, dts[p], d.xms);
}

static hipError_t launch_area(Ctx &c, double dt) {
  const int64_t lanes = c.numel * c.ne * area_parts(c.motion);
  const dim3 grid((unsigned)((lanes + 255) / 256));
#define AREA(MO)                                                                        \
  hipLaunchKernelGGL(k_area<MO>, grid, dim3(256), 0, c.stream, c.numel, c.numnp, c.d_ix, \
                     c.d_uel, c.d_active, c.d_xms, c.d_cps, c.d_ufns, c.d_dbt, c.phys, dt, \
                     c.d_kel, c.d_rel)
  switch (c.motion) {
    case STATIC: AREA(STATIC); break;
    case LAG: AREA(LAG); break;
    case EUL: AREA(EUL); break;
    case ALEV: AREA(ALEV); break;
    case ALEVB: AREA(ALEVB); break;
    default: return hipErrorInvalidValue;
  }
#undef AREA
  return hipGetLastError();
}

static hipError_t launch_edge(Ctx &c, double time, double dt) {
  if (c.n_edges == 0) return hipSuccess;
  const int64_t lanes = c.n_edges * c.ne;
  const dim3 grid((unsigned)((lanes + 255) / 256));
  double *kel = c.d_kel + c.numel * (int64_t)c.ne * c.ne;
  double *rel = c.d_rel + c.numel * (int64_t)c.ne;
#define EDGE(MO)                                                                            \
  hipLaunchKernelGGL(k_edge<MO>, grid, dim3(256), 0, c.stream, c.n_edges, c.numnp, c.d_edges, \
                     c.d_ix, c.d_active, c.d_xms, c.d_cps, c.d_bufns, c.phys, time, dt, kel, rel)
  switch (c.motion) {
    case STATIC: EDGE(STATIC); break;
    case LAG: EDGE(LAG); break;
    case EUL: EDGE(EUL); break;
    case ALEV: EDGE(ALEV); break;
    case ALEVB: EDGE(ALEVB); break;
    default: return hipErrorInvalidValue;
  }
#undef EDGE
  return hipGetLastError();
}

int launch_assemble(Ctx &c, double time, double dt) {
  MAFEM_HIP(hipEventRecord(c.ev[8], c.stream));
  MAFEM_HIP(launch_area(c, dt));
  MAFEM_HIP(hipEventRecord(c.ev[9], c.stream));
  MAFEM_HIP(launch_edge(c, time, dt));
  const int64_t nb = c.nnzb * c.ndf * c.ndf;
  hipLaunchKernelGGL(k_gather_bsr, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, c.stream,
                     c.nnzb, c.ndf, c.ne, c.d_brow, c.d_colidx, c.d_active, c.d_gptr, c.d_gsrc,
                     c.d_kel, c.d_bsr);
  MAFEM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_gather_r, dim3((unsigned)((c.nx + 255) / 256)), dim3(256), 0, c.stream,
                     c.nx, c.ndf, c.d_active, c.d_rptr, c.d_rsrc, c.d_rel, c.d_r);
  MAFEM_HIP(hipGetLastError());
  return 0;
}

void launch_pull_force(Ctx &c, hipStream_t s, const int32_t *el, const int8_t *slot, double *rel,
                       double *out) {
  const int64_t n = (int64_t)c.h_pf_el.size();
  const dim3 grid((unsigned)((n * c.ne + 255) / 256));
#define PLIST(MO)                                                                            \
  hipLaunchKernelGGL(k_area_list<MO>, grid, dim3(256), 0, s, n, c.numnp, c.d_ix, c.d_uel,    \
                     c.d_active, c.d_xms, c.d_cps, c.d_ufns, c.d_dbt, c.phys, el, rel)
  switch (c.motion) {
    case LAG: PLIST(LAG); break;
    case EUL: PLIST(EUL); break;
    case ALEV: PLIST(ALEV); break;
    case ALEVB: PLIST(ALEVB); break;
    default: MAFEM_HIP(hipErrorInvalidValue);
  }
#undef PLIST
  MAFEM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_pull_reduce, dim3(1), dim3(64), 0, s, (int)n, c.ne, c.ndf, slot,
                     (const double *)rel, c.d_ix + 9 * (int64_t)c.pf_pull_el, c.d_xms, c.numnp,
                     out);
  MAFEM_HIP(hipGetLastError());
}

void launch_output(Ctx &c) {
  OutGrid g;
  g.Z1 = 3 * c.num1el + 2;
  g.Z2 = 3 * c.num2el + 2;
  g.num1el = c.num1el;
  g.numnp = c.numnp;
  g.ndf = c.ndf;
  for (int k = 0; k < 4; ++k) {
    g.crnr_el[k] = c.crnr_el[k];
    g.boff[k] = c.h_boff[k];
  }
  const int64_t npts = g.Z1 * g.Z2;
  hipLaunchKernelGGL(k_generate_output, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0,
                     c.stream, g, c.d_ix, c.d_uel, c.d_ufns, c.d_bufns, c.d_obel, c.d_crnr,
                     c.d_xms, c.d_cps, c.d_out);
  MAFEM_HIP(hipGetLastError());
}

void launch_predict(Ctx &c, const double *d_u, double dt) {
  hipLaunchKernelGGL(k_predict, dim3((unsigned)((c.numnp + 255) / 256)), dim3(256), 0, c.stream,
                     c.numnp, c.mm_order[0], c.mm_order[1], c.mm_order[2], d_u, dt, c.d_xms);
  MAFEM_HIP(hipGetLastError());
}

void b_assemble(BatchPlan &B) {
  for (size_t k = 0; k < B.motions.size(); ++k) {
    const Seg &sa = B.area[k], &se = B.edge[k];
#define BAREA(MO)                                                                              \
  hipLaunchKernelGGL(kb_area<MO>, dim3(sa.nblk), dim3(256), 0, B.stream, sa, B.d_pd, B.f.on,   \
                     B.f.dt);                                                                  \
  if (se.nblk)                                                                                 \
    hipLaunchKernelGGL(kb_edge<MO>, dim3(se.nblk), dim3(256), 0, B.stream, se, B.d_pd, B.f.on, \
                       B.f.time, B.f.dt)
    switch (B.motions[k]) {
      case STATIC: BAREA(STATIC); break;
      case LAG: BAREA(LAG); break;
      case EUL: BAREA(EUL); break;
      case ALEV: BAREA(ALEV); break;
      case ALEVB: BAREA(ALEVB); break;
      default: throw HipError(hipErrorInvalidValue, "batch motion", __FILE__, __LINE__);
    }
#undef BAREA
    MAFEM_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(kb_gather_bsr, dim3(B.bsr.nblk), dim3(256), 0, B.stream, B.bsr, B.d_pd, B.f.on);
  hipLaunchKernelGGL(kb_gather_r, dim3(B.vec.nblk), dim3(256), 0, B.stream, B.vec, B.d_pd, B.f.on);
  MAFEM_HIP(hipGetLastError());
}

void b_predict(BatchPlan &B) {
  hipLaunchKernelGGL(kb_predict, dim3(B.node.nblk), dim3(256), 0, B.stream, B.node, B.d_pd, B.f.on,
                     B.f.dt);
  MAFEM_HIP(hipGetLastError());
}

}  // namespace mafem
