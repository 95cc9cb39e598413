// CPU check of the nested-dissection front plan (csrc/nd_plan.h): a dense
// emulation of the multifrontal factorization and solve that the device runs
// over the plan, with the same front layout (identity padding of the
// fully-summed part, zero padding of the boundary), the same entry ownership
// (an entry belongs to the front that eliminates the earlier of its two
// nodes), the same children-in-order gather of Schur complements and the same
// tile algorithm (64 x 64 tiles, partial pivoting inside each diagonal tile,
// U panels from the permuted rows, L panels by column solves, the stored L
// tiles of a tile row permuted by that row's pivots).  The matrix is random
// on the node-block pattern (5 x 5 node stencil), with identity rows and
// columns for some prescribed unknowns and a zero diagonal entry per node (a
// constraint unknown, so pivoting is needed).  The solution is compared with
// a dense partial-pivoting LU solve of the same matrix.
//
// usage: nd_check n1 n2 ndf leaf seed   -> prints "relres_nd relres_dense
// max_rel_diff fronts levels steps" and exits 0, or prints an error and exits 1
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "nd_plan.h"

using namespace mafem::nd;
typedef std::vector<double> Vec;

struct Dense {
  int64_t n;
  Vec a;
  double &at(int64_t i, int64_t j) { return a[i * n + j]; }
};

static bool adjacent(int64_t p, int64_t q, int64_t n1) {
  return std::llabs(p % n1 - q % n1) <= 2 && std::llabs(p / n1 - q / n1) <= 2;
}

// tile-local LU of the first fT tile rows of a front stored densely (S x S)
static void front_lu(Dense &M, int fT, int FT, std::vector<int> &piv) {
  const int T = NT;
  piv.assign((size_t)fT * T, 0);
  for (int K = 0; K < fT; ++K) {
    const int64_t r0 = (int64_t)K * T;
    // diagonal tile: partial pivoting within its rows, rows stored in pivot order
    std::vector<int> rows(T);
    for (int i = 0; i < T; ++i) rows[i] = i;
    Dense D{T, Vec((size_t)T * T)};
    for (int i = 0; i < T; ++i)
      for (int j = 0; j < T; ++j) D.at(i, j) = M.at(r0 + i, r0 + j);
    std::vector<int> chosen(T, 0), order(T);
    for (int j = 0; j < T; ++j) {
      int p = -1;
      double best = -1.0;
      for (int i = 0; i < T; ++i)
        if (!chosen[i] && std::fabs(D.at(i, j)) > best) best = std::fabs(D.at(i, j)), p = i;
      chosen[p] = 1;
      order[j] = p;
      const double pv = D.at(p, j);
      for (int i = 0; i < T; ++i) {
        if (chosen[i]) continue;
        const double l = pv != 0.0 ? D.at(i, j) / pv : 0.0;
        D.at(i, j) = l;
        for (int c = j + 1; c < T; ++c) D.at(i, c) -= l * D.at(p, c);
      }
    }
    for (int j = 0; j < T; ++j) piv[(size_t)K * T + j] = order[j];
    for (int i = 0; i < T; ++i)
      for (int j = 0; j < T; ++j) M.at(r0 + i, r0 + j) = D.at(order[i], j);
    // U panels: U_KJ = L^-1 P A_KJ (unit lower L)
    for (int J = K + 1; J < FT; ++J) {
      const int64_t c0 = (int64_t)J * T;
      Dense X{T, Vec((size_t)T * T)};
      for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) X.at(i, j) = M.at(r0 + order[i], c0 + j);
      for (int i = 0; i < T; ++i)
        for (int k = 0; k < i; ++k) {
          const double l = M.at(r0 + i, r0 + k);
          for (int j = 0; j < T; ++j) X.at(i, j) -= l * X.at(k, j);
        }
      for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) M.at(r0 + i, c0 + j) = X.at(i, j);
    }
    // L panels: L_IK = A_IK U^-1
    for (int I = K + 1; I < FT; ++I) {
      const int64_t q0 = (int64_t)I * T;
      for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) {
          double s = M.at(q0 + i, r0 + j);
          for (int k = 0; k < j; ++k) s -= M.at(q0 + i, r0 + k) * M.at(r0 + k, r0 + j);
          M.at(q0 + i, r0 + j) = s / M.at(r0 + j, r0 + j);
        }
    }
    // stored L tiles of tile row K take its row order
    for (int J = 0; J < K; ++J) {
      const int64_t c0 = (int64_t)J * T;
      Dense X{T, Vec((size_t)T * T)};
      for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) X.at(i, j) = M.at(r0 + order[i], c0 + j);
      for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) M.at(r0 + i, c0 + j) = X.at(i, j);
    }
    // trailing update
    const int64_t e = (int64_t)FT * T;
    for (int64_t i = r0 + T; i < e; ++i)
      for (int k = 0; k < T; ++k) {
        const double l = M.at(i, r0 + k);
        if (l == 0.0) continue;
        for (int64_t j = r0 + T; j < e; ++j) M.at(i, j) -= l * M.at(r0 + k, j);
      }
  }
}

int main(int argc, char **argv) {
  if (argc < 6) return 2;
  const int64_t n1 = atoll(argv[1]), n2 = atoll(argv[2]);
  const int ndf = atoi(argv[3]), leaf = atoi(argv[4]);
  std::mt19937_64 rng(strtoull(argv[5], nullptr, 10));
  std::normal_distribution<double> g(0.0, 1.0);
  const Plan P = build_plan(n1, n2, ndf, leaf);
  const int64_t numnp = n1 * n2, N = numnp * ndf;

  // matrix on the node-block pattern
  Dense A{N, Vec((size_t)(N * N), 0.0)};
  std::vector<char> presc(N, 0);
  for (int64_t p = 0; p < numnp; ++p)
    if (p % 7 == 3) presc[p * ndf] = 1;  // a prescribed unknown on some nodes
  for (int64_t p = 0; p < numnp; ++p)
    for (int64_t q = 0; q < numnp; ++q) {
      if (!adjacent(p, q, n1)) continue;
      for (int r = 0; r < ndf; ++r)
        for (int c = 0; c < ndf; ++c) {
          const int64_t i = p * ndf + r, j = q * ndf + c;
          if (presc[i] || presc[j]) {
            A.at(i, j) = i == j ? 1.0 : 0.0;
            continue;
          }
          double v = g(rng);
          if (p == q && r == c) v = (r == ndf - 1) ? 0.0 : v + 4.0;
          A.at(i, j) = v;
        }
    }
  Vec rhs(N);
  for (auto &v : rhs) v = g(rng);

  // structural checks: every coupled entry owned by exactly one front that holds both
  std::vector<int64_t> own_count((size_t)(N * N), 0);
  for (int32_t t = 0; t < (int32_t)P.fr.size(); ++t) {
    const Front &F = P.fr[t];
    if (F.level != 0 && (F.child[0] < 0 || F.child[1] < 0)) { puts("separator without two children"); return 1; }
    for (int s = 0; s < 2; ++s)
      if (F.child[s] >= 0 && P.fr[F.child[s]].level >= F.level) { puts("level order"); return 1; }
    for (int64_t k = 0; k < F.f; ++k)
      for (int64_t l = 0; l < F.f + F.b; ++l) {
        const int64_t d1 = F.dofs[k], d2 = F.dofs[l];
        if (!adjacent(d1 / ndf, d2 / ndf, n1)) continue;
        own_count[d1 * N + d2]++;
        if (l >= F.f) own_count[d2 * N + d1]++;
      }
  }
  for (int64_t p = 0; p < numnp; ++p)
    for (int64_t q = 0; q < numnp; ++q) {
      if (!adjacent(p, q, n1)) continue;
      const int64_t d1 = p * ndf, d2 = q * ndf;
      if (own_count[d1 * N + d2] != 1) {
        printf("entry (%lld, %lld) in %lld fronts\n", (long long)d1, (long long)d2,
               (long long)own_count[d1 * N + d2]);
        return 1;
      }
    }

  // numeric: fronts level by level
  const size_t nf = P.fr.size();
  std::vector<Dense> Fm(nf);
  std::vector<std::vector<int>> piv(nf);
  for (int lev = 0; lev < P.levels; ++lev)
    for (int32_t t : P.by_level[lev]) {
      const Front &F = P.fr[t];
      const int64_t S = F.size();
      Dense M{S, Vec((size_t)(S * S), 0.0)};
      std::vector<int64_t> dof_at(S, -1);
      for (int64_t k = 0; k < F.f + F.b; ++k) dof_at[F.pos_of(k)] = F.dofs[k];
      for (int64_t q1 = 0; q1 < S; ++q1)
        for (int64_t q2 = 0; q2 < S; ++q2) {
          double v = 0.0;
          const int64_t d1 = dof_at[q1], d2 = dof_at[q2];
          if (d1 < 0 || d2 < 0) {
            v = (q1 == q2 && q1 >= F.f && q1 < (int64_t)F.fT * NT) ? 1.0 : 0.0;
          } else {
            const int32_t o = std::min(P.node_owner[d1 / ndf], P.node_owner[d2 / ndf]);
            if (o == t && adjacent(d1 / ndf, d2 / ndf, n1)) v = A.at(d1, d2);
          }
          for (int s = 0; s < 2; ++s) {
            if (F.child[s] < 0) continue;
            const int32_t c1 = P.cmap[t][s][q1], c2 = P.cmap[t][s][q2];
            if (c1 >= 0 && c2 >= 0) v += Fm[F.child[s]].at(c1, c2);
          }
          M.at(q1, q2) = v;
        }
      front_lu(M, F.fT, F.FT, piv[t]);
      Fm[t] = std::move(M);
    }

  // solve: forward by levels (children's boundary parts gathered in order)
  std::vector<Vec> y(nf), x(nf);
  for (int lev = 0; lev < P.levels; ++lev)
    for (int32_t t : P.by_level[lev]) {
      const Front &F = P.fr[t];
      const int64_t S = F.size(), fs = (int64_t)F.fT * NT;
      Vec v(S, 0.0);
      for (int64_t k = 0; k < F.f; ++k) v[k] = rhs[F.dofs[k]];
      for (int s = 0; s < 2; ++s) {
        if (F.child[s] < 0) continue;
        for (int64_t q = 0; q < S; ++q)
          if (P.cmap[t][s][q] >= 0) v[q] += y[F.child[s]][P.cmap[t][s][q]];
      }
      Dense &M = Fm[t];
      // F rows: tile pivots, unit lower solve
      for (int K = 0; K < F.fT; ++K) {
        Vec tmp(NT);
        for (int i = 0; i < NT; ++i) tmp[i] = v[(int64_t)K * NT + piv[t][(size_t)K * NT + i]];
        for (int i = 0; i < NT; ++i) v[(int64_t)K * NT + i] = tmp[i];
        for (int i = 0; i < NT; ++i) {
          const int64_t r = (int64_t)K * NT + i;
          double s = v[r];
          for (int64_t c = 0; c < r; ++c) s -= M.at(r, c) * v[c];
          v[r] = s;
        }
      }
      for (int64_t r = fs; r < S; ++r) {
        double s = v[r];
        for (int64_t c = 0; c < fs; ++c) s -= M.at(r, c) * v[c];
        v[r] = s;
      }
      y[t] = v;
    }
  Vec sol(N, 0.0);
  for (int lev = P.levels - 1; lev >= 0; --lev)
    for (int32_t t : P.by_level[lev]) {
      const Front &F = P.fr[t];
      const int64_t S = F.size(), fs = (int64_t)F.fT * NT;
      Vec v = y[t];
      for (int64_t q = fs; q < S; ++q) v[q] = 0.0;
      for (int64_t k = F.f; k < F.f + F.b; ++k) v[F.pos_of(k)] = sol[F.dofs[k]];
      Dense &M = Fm[t];
      for (int64_t r = fs - 1; r >= 0; --r) {
        double s = v[r];
        for (int64_t c = r + 1; c < S; ++c) s -= M.at(r, c) * v[c];
        v[r] = s / M.at(r, r);
      }
      for (int64_t k = 0; k < F.f; ++k) sol[F.dofs[k]] = v[k];
    }

  // dense reference
  Dense D = A;
  Vec b = rhs;
  for (int64_t j = 0; j < N; ++j) {
    int64_t p = j;
    for (int64_t i = j + 1; i < N; ++i)
      if (std::fabs(D.at(i, j)) > std::fabs(D.at(p, j))) p = i;
    if (p != j) {
      for (int64_t c = 0; c < N; ++c) std::swap(D.at(j, c), D.at(p, c));
      std::swap(b[j], b[p]);
    }
    for (int64_t i = j + 1; i < N; ++i) {
      const double l = D.at(i, j) / D.at(j, j);
      if (l == 0.0) continue;
      for (int64_t c = j; c < N; ++c) D.at(i, c) -= l * D.at(j, c);
      b[i] -= l * b[j];
    }
  }
  Vec ref(N);
  for (int64_t i = N - 1; i >= 0; --i) {
    double s = b[i];
    for (int64_t c = i + 1; c < N; ++c) s -= D.at(i, c) * ref[c];
    ref[i] = s / D.at(i, i);
  }
  auto relres = [&](const Vec &z) {
    double rn = 0, bn = 0;
    for (int64_t i = 0; i < N; ++i) {
      double s = rhs[i];
      for (int64_t j = 0; j < N; ++j) s -= A.at(i, j) * z[j];
      rn += s * s;
      bn += rhs[i] * rhs[i];
    }
    return std::sqrt(rn / bn);
  };
  double dmax = 0, rmax = 0;
  for (int64_t i = 0; i < N; ++i) {
    dmax = std::max(dmax, std::fabs(sol[i] - ref[i]));
    rmax = std::max(rmax, std::fabs(ref[i]));
  }
  printf("%.3e %.3e %.3e %zu %d %lld\n", relres(sol), relres(ref), dmax / rmax, nf, P.levels,
         (long long)P.critical_steps());
  return 0;
}
