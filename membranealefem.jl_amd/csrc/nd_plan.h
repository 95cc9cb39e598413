// Nested-dissection elimination tree of the node grid, for the multifrontal
// factorization of the extended node-block system (host code only).
//
// Nodes i1 + n1 i2 of the (n1 x n2) control-point grid couple when they share
// a biquadratic element, |Δi1| <= 2 and |Δi2| <= 2 (the block pattern the
// tangent is assembled into), so two node lines separate a box.  The box is
// bisected along its longer side by the two middle lines until both sides are
// at most `leaf` nodes; each separator (or leaf box) is a front whose
// fully-summed unknowns are its nodes' ndf unknowns, and whose boundary is the
// set of later-eliminated nodes coupled to its subtree.  Fronts are numbered
// in postorder (children before parents, elimination order), a front's level
// is its height (leaves 0), so every level depends only on lower ones.
//
// Front storage (TILE-row tiles): positions 0..f-1 are the fully-summed
// unknowns, f..fT TILE - 1 identity padding, fT TILE .. fT TILE + b - 1 the
// boundary unknowns (ordered by eliminating front, then node), then zero
// padding to FT TILE.  The partial LU eliminates the first fT tile rows; the
// trailing (FT - fT) tile block is the Schur complement passed to the parent.
//
// This replaces the fill-reducing ordering UMFPACK applies inside
// `K_gl \ r_gl` (src/analysis/FiniteElement.jl:38).
#pragma once
#include <algorithm>
#include <array>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <vector>

namespace mafem {
namespace nd {

constexpr int NT = 64;  // tile edge of the front storage

struct Front {
  int64_t f = 0, b = 0;        // fully-summed / boundary unknowns
  int32_t fT = 0, FT = 0;      // tile rows of the fully-summed part / of the front
  int32_t level = 0;
  int32_t child[2] = {-1, -1};
  int32_t parent = -1;
  int64_t box[4] = {0, 0, 0, 0};  // [a1, b1) x [a2, b2) of its subtree
  std::vector<int64_t> nodes;  // separator (or leaf) nodes, increasing
  std::vector<int64_t> bnodes; // boundary nodes, by eliminating front then node
  std::vector<int64_t> dofs;   // extended indices: f fully-summed, then b boundary
  // storage position of entry k of dofs
  int64_t pos_of(int64_t k) const { return k < f ? k : (int64_t)fT * NT + (k - f); }
  int64_t size() const { return (int64_t)FT * NT; }
};

struct Plan {
  int64_t n1 = 0, n2 = 0;
  int ndf = 0;
  std::vector<Front> fr;              // postorder
  std::vector<int32_t> node_owner;    // front whose fully-summed set holds each node
  int levels = 0;
  std::vector<std::vector<int32_t>> by_level;  // fronts of each level, fT decreasing
  // cmap[t][s][q]: position in child s of front t of the unknown at storage
  // position q of t (-1: not in that child's boundary); empty without child s
  std::vector<std::array<std::vector<int32_t>, 2>> cmap;

  // dependent tile steps of one factorization: sum over levels of the largest fT
  int64_t critical_steps() const {
    int64_t s = 0;
    for (const auto &l : by_level) s += l.empty() ? 0 : fr[l.front()].fT;
    return s;
  }
};

inline Plan build_plan(int64_t n1, int64_t n2, int ndf, int leaf = 4) {
  if (n1 < 1 || n2 < 1 || ndf < 1 || leaf < 4) throw std::invalid_argument("nd plan sizes");
  Plan P;
  P.n1 = n1;
  P.n2 = n2;
  P.ndf = ndf;
  const int64_t numnp = n1 * n2;
  P.node_owner.assign(numnp, -1);
  std::function<int32_t(int64_t, int64_t, int64_t, int64_t)> rec =
      [&](int64_t a1, int64_t b1, int64_t a2, int64_t b2) -> int32_t {
    const int64_t w1 = b1 - a1, w2 = b2 - a2;
    Front F;
    F.box[0] = a1, F.box[1] = b1, F.box[2] = a2, F.box[3] = b2;
    if (w1 <= leaf && w2 <= leaf) {
      for (int64_t i2 = a2; i2 < b2; ++i2)
        for (int64_t i1 = a1; i1 < b1; ++i1) F.nodes.push_back(i1 + n1 * i2);
    } else if (w1 >= w2) {
      const int64_t m = a1 + (w1 - 2) / 2;  // lines m, m + 1; both sides non-empty (w1 >= 5)
      F.child[0] = rec(a1, m, a2, b2);
      F.child[1] = rec(m + 2, b1, a2, b2);
      for (int64_t i2 = a2; i2 < b2; ++i2)
        for (int64_t i1 = m; i1 < m + 2; ++i1) F.nodes.push_back(i1 + n1 * i2);
    } else {
      const int64_t m = a2 + (w2 - 2) / 2;
      F.child[0] = rec(a1, b1, a2, m);
      F.child[1] = rec(a1, b1, m + 2, b2);
      for (int64_t i2 = m; i2 < m + 2; ++i2)
        for (int64_t i1 = a1; i1 < b1; ++i1) F.nodes.push_back(i1 + n1 * i2);
    }
    const int32_t id = (int32_t)P.fr.size();
    for (int s = 0; s < 2; ++s)
      if (F.child[s] >= 0) P.fr[F.child[s]].parent = id;
    for (int64_t p : F.nodes) P.node_owner[p] = id;
    P.fr.push_back(std::move(F));
    return id;
  };
  rec(0, n1, 0, n2);

  // boundary sets, levels, unknown lists
  std::vector<char> mark(numnp, 0);
  for (int32_t t = 0; t < (int32_t)P.fr.size(); ++t) {
    Front &F = P.fr[t];
    std::vector<int64_t> B;
    auto add = [&](int64_t q) {
      if (P.node_owner[q] > t && !mark[q]) {
        mark[q] = 1;
        B.push_back(q);
      }
    };
    int lev = 0;
    for (int s = 0; s < 2; ++s)
      if (F.child[s] >= 0) {
        const Front &C = P.fr[F.child[s]];
        lev = std::max(lev, C.level + 1);
        for (int64_t q : C.bnodes) add(q);
      }
    for (int64_t p : F.nodes) {
      const int64_t i1 = p % n1, i2 = p / n1;
      for (int64_t j2 = std::max<int64_t>(0, i2 - 2); j2 <= std::min(n2 - 1, i2 + 2); ++j2)
        for (int64_t j1 = std::max<int64_t>(0, i1 - 2); j1 <= std::min(n1 - 1, i1 + 2); ++j1)
          add(j1 + n1 * j2);
    }
    for (int64_t q : B) mark[q] = 0;
    std::sort(B.begin(), B.end(), [&](int64_t x, int64_t y) {
      const int32_t ox = P.node_owner[x], oy = P.node_owner[y];
      return ox != oy ? ox < oy : x < y;
    });
    F.bnodes = std::move(B);
    F.level = lev;
    F.f = (int64_t)F.nodes.size() * ndf;
    F.b = (int64_t)F.bnodes.size() * ndf;
    F.fT = (int32_t)((F.f + NT - 1) / NT);
    F.FT = F.fT + (int32_t)((F.b + NT - 1) / NT);
    F.dofs.reserve(F.f + F.b);
    for (int64_t p : F.nodes)
      for (int d = 0; d < ndf; ++d) F.dofs.push_back(p * ndf + d);
    for (int64_t p : F.bnodes)
      for (int d = 0; d < ndf; ++d) F.dofs.push_back(p * ndf + d);
    P.levels = std::max(P.levels, lev + 1);
  }
  P.by_level.assign(P.levels, {});
  for (int32_t t = 0; t < (int32_t)P.fr.size(); ++t) P.by_level[P.fr[t].level].push_back(t);
  for (auto &l : P.by_level)
    std::stable_sort(l.begin(), l.end(),
                     [&](int32_t x, int32_t y) { return P.fr[x].fT > P.fr[y].fT; });

  // child maps, through a dense unknown -> position table of the parent
  P.cmap.resize(P.fr.size());
  std::vector<int32_t> pos(numnp * ndf, -1);
  for (int32_t t = 0; t < (int32_t)P.fr.size(); ++t) {
    const Front &F = P.fr[t];
    if (F.child[0] < 0 && F.child[1] < 0) continue;
    for (int64_t k = 0; k < (int64_t)F.dofs.size(); ++k) pos[F.dofs[k]] = (int32_t)F.pos_of(k);
    for (int s = 0; s < 2; ++s) {
      if (F.child[s] < 0) continue;
      const Front &C = P.fr[F.child[s]];
      std::vector<int32_t> &m = P.cmap[t][s];
      m.assign(F.size(), -1);
      for (int64_t k = C.f; k < C.f + C.b; ++k) {
        const int32_t q = pos[C.dofs[k]];
        if (q < 0) throw std::logic_error("nd plan: child boundary outside its parent front");
        m[q] = (int32_t)C.pos_of(k);
      }
    }
    for (int64_t d : F.dofs) pos[d] = -1;
  }
  return P;
}

}  // namespace nd
}  // namespace mafem
