// C ABI of libmafem (declared in include/mafem.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

using namespace mafem;


static thread_local std::string g_create_err;

std::mutex &mafem::capture_mutex() {
  static std::mutex m;
  return m;
}

namespace {

template <class T>
T *dalloc(Ctx &c, int64_t n) {
  void *p = nullptr;
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc(&p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw ArgError{"device allocation of " + std::to_string(sizeof(T) * (size_t)n) +
                       " bytes failed",
                   MAFEM_ENOMEM};
  }
  return (T *)p;
}

template <class T>
T *dupload(Ctx &c, const std::vector<T> &v) {
  T *p = dalloc<T>(c, (int64_t)v.size());
  if (!v.empty())
    MAFEM_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return p;
}

void free_all(Ctx &c) {
  void *ptrs[] = {c.d_xms,  c.d_cps,    c.d_ix,      c.d_uel,     c.d_active, c.d_ufns,
                  c.d_dbt,  c.d_bufns,  c.d_edges,   c.d_kel,     c.d_rel,    c.d_rowptr,
                  c.d_colidx, c.d_brow, c.d_bsr,     c.d_gptr,    c.d_gsrc,   c.d_rptr,
                  c.d_rsrc, c.d_r,      c.d_x,       c.d_w,       c.d_t,      c.d_b,
                  c.d_red,  c.d_csc_src, c.d_csc_val, c.d_eq,     c.d_band,   c.d_linv,
                  c.d_uinv, c.d_piv,  c.d_V,       c.d_Z,       c.d_h,      c.d_coef,
                  c.d_mdot_part, c.d_t2, c.d_presc_idx, c.d_presc_val, c.d_s1, c.d_s2,
                  c.d_sweep_err, c.d_pf_el, c.d_pf_slot, c.d_pf_rel, c.d_pf_out,
                  c.d_obel, c.d_crnr, c.d_out, c.d_amax, c.d_tw_z, c.d_tw_br,
                  c.d_tw_cr, c.d_tw_dr, c.d_tw_s1, c.d_tw_s2, c.d_pub, c.d_span};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  nd_free(c);
  if (c.h_red) (void)hipHostFree(c.h_red);
  for (auto &e : c.ev)
    if (e) (void)hipEventDestroy(e);
  if (c.stream) (void)hipStreamDestroy(c.stream);
}

}  // namespace

// Device copies of the pull-force maps and its output buffers (first use).
// Callers check c.pf_err first; HIP failures throw HipError.
void mafem::ensure_pull(Ctx &c) {
  if (c.d_pf_out || !c.pf_err.empty()) return;
  std::lock_guard<std::mutex> no_capture(capture_mutex());
  const size_t n = c.h_pf_el.size();
  MAFEM_HIP(hipMalloc((void **)&c.d_pf_el, sizeof(int32_t) * n));
  MAFEM_HIP(hipMalloc((void **)&c.d_pf_slot, n * 9));
  MAFEM_HIP(hipMalloc((void **)&c.d_pf_rel, sizeof(double) * n * c.ne));
  MAFEM_HIP(hipMemcpy(c.d_pf_el, c.h_pf_el.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  MAFEM_HIP(hipMemcpy(c.d_pf_slot, c.h_pf_slot.data(), n * 9, hipMemcpyHostToDevice));
  MAFEM_HIP(hipMalloc((void **)&c.d_pf_out, sizeof(double) * 6));
}

namespace {

int fail(Ctx &c, const std::string &m, int code) {
  c.err = m;
  return code;
}

// running count of replaced pivots (blocking read of the device counter)
int32_t read_pert(Ctx &c) {
  unsigned long long v[2] = {0, 0};
  MAFEM_HIP(hipMemcpyAsync(v, c.d_amax, sizeof(v), hipMemcpyDeviceToHost, c.stream));
  MAFEM_HIP(hipStreamSynchronize(c.stream));
  return (int32_t)v[1];
}

int from_hip(Ctx &c, const HipError &h) {
  c.err = std::string("HIP error ") + hipGetErrorString(h.code) + " in " + h.what + " (" + h.file +
          ":" + std::to_string(h.line) + ")";
  return MAFEM_EHIP;
}

// The supported DOF layouts are those of get_dofs (src/input/Bc.jl:414-431).
bool layout_ok(const mafem_mesh_desc *m, int motion, int &ndf) {
  auto eq3 = [](const int32_t *a, int x, int y, int z) { return a[0] == x && a[1] == y && a[2] == z; };
  if (motion == STATIC) {  // flat 2-D scenarios: vx, vy, λ (Bc.jl:58-157, 286-316)
    ndf = 3;
    return eq3(m->v_order, 1, 2, 0) && eq3(m->m_order, 0, 0, 0) && m->lam_order == 3 &&
           m->p_order == 0;
  }
  if (!eq3(m->v_order, 1, 2, 3)) return false;
  if (motion == LAG) {
    ndf = 4;
    return eq3(m->m_order, 0, 0, 0) && m->lam_order == 4 && m->p_order == 0;
  }
  if (motion == EUL) {
    ndf = 7;
    return eq3(m->m_order, 4, 5, 6) && m->lam_order == 7 && m->p_order == 0;
  }
  if (motion == ALEV || motion == ALEVB) {
    ndf = 8;
    return eq3(m->m_order, 4, 5, 6) && m->lam_order == 7 && m->p_order == 8;
  }
  return false;
}

void build(Ctx &c, const mafem_mesh_desc *m, const mafem_params *p) {
  if (!m || !p) throw ArgError{"null mesh or params", MAFEM_EINVAL};
  if (m->numel != m->num1el * m->num2el || m->numnp != m->num1np * m->num2np ||
      m->num1np != m->num1el + 2 || m->num2np != m->num2el + 2 || m->numel <= 0)
    throw ArgError{"inconsistent flat-patch mesh sizes", MAFEM_EINVAL};
  int ndf = 0;
  if (!layout_ok(m, p->motion, ndf))
    throw ArgError{"unsupported motion / DOF layout (supported: LAG, EUL, ALEV, ALEVB, STATIC 2-D with "
                   "the get_dofs ordering)",
                   MAFEM_EUNSUPPORTED};
  if (m->ndf != ndf) throw ArgError{"ndf does not match the DOF layout", MAFEM_EINVAL};
  if (!m->IX || !m->ID || !m->area_uel_ids || !m->area_ufns || m->area_n_uel <= 0)
    throw ArgError{"missing mesh tables", MAFEM_EINVAL};
  for (int64_t k = 0; k < m->n_neu; ++k) {
    const int b = m->neu_bdry[k], t = m->neu_type[k];
    if (b < 1 || b > 4 || t < 1 || t > 3) throw ArgError{"bad Neumann entry", MAFEM_EINVAL};
    if (t == MOMENT && p->scenario != F_BEND)
      throw ArgError{"Neumann boundary condition not implemented", MAFEM_EUNSUPPORTED};
    if (m->bdry_n_el[b - 1] > 0 && (!m->bdry_elems[b - 1] || !m->bdry_uel_ids[b - 1] ||
                                    !m->bdry_ufns[b - 1] || m->bdry_n_uel[b - 1] <= 0))
      throw ArgError{"missing boundary tables", MAFEM_EINVAL};
  }
  if (p->zv == 0.0) throw ArgError{"zv must be nonzero", MAFEM_EINVAL};

  c.num1el = m->num1el; c.num2el = m->num2el; c.numel = m->numel;
  c.num1np = m->num1np; c.num2np = m->num2np; c.numnp = m->numnp;
  c.ndf = ndf; c.ne = 9 * ndf; c.motion = p->motion; c.enr = p->enr;
  c.phys = PhysParams{p->kb, p->kg, p->zv, p->pn, p->adb, p->am, p->bend_tm, p->scenario};
  for (int i = 0; i < 3; ++i) { c.v_order[i] = m->v_order[i]; c.m_order[i] = m->m_order[i]; }
  c.lam_order = m->lam_order; c.p_order = m->p_order;
  for (int i = 0; i < 3; ++i) c.mm_order[i] = (p->motion == LAG) ? m->v_order[i] : m->m_order[i];
  const int64_t numel = c.numel, numnp = c.numnp, ne = c.ne;

  // IX, ID
  std::vector<int32_t> ix(9 * numel);
  for (int64_t k = 0; k < 9 * numel; ++k) {
    const int64_t v = m->IX[k];
    if (v < 1 || v > numnp) throw ArgError{"IX entry out of range", MAFEM_EINVAL};
    ix[k] = (int32_t)(v - 1);
  }
  // pull-force maps (get_pull_el_id and get_adj_maps, PullForce.jl:8-53, with
  // poly = 2): the 5 x 5 element ids around the pulled one taken by id
  // arithmetic (rows wrap as in the reference), and for each the pairing of
  // the pulled element's shared local nodes with its own, both in increasing
  // local order
  {
    const int64_t pull = (numel + 1) / 2;  // ceil(numel / 2), 1-based
    const int poly = 2;
    c.pf_pull_el = (int32_t)(pull - 1);
    const int32_t *pn = &ix[9 * (pull - 1)];
    for (int i = -poly; i <= poly && c.pf_err.empty(); ++i)
      for (int64_t e = pull + i * m->num1el - poly; e <= pull + i * m->num1el + poly; ++e) {
        if (e < 1 || e > numel) {
          c.pf_err = "pull force: adjacent element id " + std::to_string(e) +
                     " is outside 1.." + std::to_string(numel);
          break;
        }
        const int32_t *an = &ix[9 * (e - 1)];
        int l1[9], l2[9], n1 = 0, n2 = 0;
        for (int a = 0; a < 9; ++a)
          for (int b = 0; b < 9; ++b)
            if (pn[a] == an[b]) { l1[n1++] = a; break; }
        for (int b = 0; b < 9; ++b)
          for (int a = 0; a < 9; ++a)
            if (an[b] == pn[a]) { l2[n2++] = b; break; }
        c.h_pf_el.push_back((int32_t)(e - 1));
        for (int b = 0; b < 9; ++b) c.h_pf_slot.push_back(-1);
        for (int k = 0; k < n1 && k < n2; ++k)
          c.h_pf_slot[c.h_pf_slot.size() - 9 + l2[k]] = (int8_t)l1[k];
      }
    if (!c.pf_err.empty()) {
      c.h_pf_el.clear();
      c.h_pf_slot.clear();
    }
  }

  std::vector<uint8_t> active(ndf * numnp);
  int64_t nmdf = 0;
  std::vector<int64_t> eq_ext;  // extended index of each equation
  for (int64_t k = 0; k < ndf * numnp; ++k) {
    active[k] = m->ID[k] != 0;
    nmdf = std::max<int64_t>(nmdf, m->ID[k]);
  }
  if (nmdf != m->nmdf) throw ArgError{"nmdf does not match max(ID)", MAFEM_EINVAL};
  c.nmdf = nmdf;
  eq_ext.assign(nmdf, -1);
  for (int64_t k = 0; k < ndf * numnp; ++k)
    if (m->ID[k] > 0) eq_ext[m->ID[k] - 1] = k;  // ID is ndf x numnp: k = d + ndf p
  for (auto v : eq_ext)
    if (v < 0) throw ArgError{"ID is not a numbering of 1..nmdf", MAFEM_EINVAL};

  // area basis: Julia Matrix{GpBasisFnsζα}(n_uel, 9) -> [uel][gp][RECP]
  const int64_t nu = m->area_n_uel;
  std::vector<int32_t> uel(numel);
  for (int64_t e = 0; e < numel; ++e) {
    const int64_t u = m->area_uel_ids[e];
    if (u < 1 || u > nu) throw ArgError{"area uel id out of range", MAFEM_EINVAL};
    uel[e] = (int32_t)(u - 1);
  }
  std::vector<double> ufns(nu * 9 * RECP, 0.0), dbt(nu * 81);
  const double xi[3] = {-std::sqrt(3.0 / 5.0), 0.0, std::sqrt(3.0 / 5.0)};
  for (int64_t u = 0; u < nu; ++u) {
    double G[3][9] = {{0}}, H[3][3] = {{0}};
    for (int g = 0; g < 9; ++g) {
      const double *src = m->area_ufns + REC * (u + nu * g);
      std::memcpy(&ufns[(u * 9 + g) * RECP], src, sizeof(double) * REC);
      // Dohrmann--Bochev matrices with N_DB = (ξ1, ξ2, 1) at the raw
      // abscissae (FiniteElement.jl:315-320)
      const double ndb[3] = {xi[g % 3], xi[g / 3], 1.0};
      const double w = src[0];
      for (int i = 0; i < 3; ++i) {
        for (int a = 0; a < 9; ++a) G[i][a] += ndb[i] * src[1 + a] * w;
        for (int j = 0; j < 3; ++j) H[i][j] += ndb[i] * ndb[j] * w;
      }
    }
    // H^-1 by cofactors, then T = G^T H^-1 G (FiniteElement.jl:323)
    const double c00 = H[1][1] * H[2][2] - H[1][2] * H[2][1];
    const double c01 = H[1][2] * H[2][0] - H[1][0] * H[2][2];
    const double c02 = H[1][0] * H[2][1] - H[1][1] * H[2][0];
    const double det = H[0][0] * c00 + H[0][1] * c01 + H[0][2] * c02;
    double Hi[3][3];
    Hi[0][0] = c00 / det; Hi[1][0] = c01 / det; Hi[2][0] = c02 / det;
    Hi[0][1] = (H[0][2] * H[2][1] - H[0][1] * H[2][2]) / det;
    Hi[1][1] = (H[0][0] * H[2][2] - H[0][2] * H[2][0]) / det;
    Hi[2][1] = (H[0][1] * H[2][0] - H[0][0] * H[2][1]) / det;
    Hi[0][2] = (H[0][1] * H[1][2] - H[0][2] * H[1][1]) / det;
    Hi[1][2] = (H[0][2] * H[1][0] - H[0][0] * H[1][2]) / det;
    Hi[2][2] = (H[0][0] * H[1][1] - H[0][1] * H[1][0]) / det;
    for (int a = 0; a < 9; ++a)
      for (int b = 0; b < 9; ++b) {
        double s = 0;
        for (int i = 0; i < 3; ++i) {
          double t = 0;
          for (int j = 0; j < 3; ++j) t += Hi[i][j] * G[j][b];
          s += G[i][a] * t;
        }
        dbt[u * 81 + a * 9 + b] = s;
      }
  }

  // boundary records and edge work items, in the order of mesh.inh_neu_bcs
  std::vector<double> bufns;
  int64_t base[4];
  for (int b = 0; b < 4; ++b) {
    base[b] = (int64_t)bufns.size() / RECP;
    const int64_t nub = m->bdry_n_uel[b];
    for (int64_t u = 0; u < nub; ++u)
      for (int g = 0; g < 3; ++g) {
        const double *src = m->bdry_ufns[b] + REC * (u + nub * g);
        const size_t off = bufns.size();
        bufns.resize(off + RECP, 0.0);
        std::memcpy(&bufns[off], src, sizeof(double) * REC);
      }
  }
  // output sampling tables: boundary elements with their first records, and
  // the corner records (corner elements 1, num1el, numel - num1el + 1, numel,
  // Mesh.jl:134-139)
  for (int b = 0; b < 4; ++b) {
    c.h_boff[b] = (int64_t)c.h_obel.size() / 2;
    for (int64_t j = 0; j < m->bdry_n_el[b]; ++j) {
      const int64_t el = m->bdry_elems[b][j], u = m->bdry_uel_ids[b][j];
      if (el < 1 || el > numel || u < 1 || u > m->bdry_n_uel[b])
        throw ArgError{"boundary table entry out of range", MAFEM_EINVAL};
      c.h_obel.push_back((int32_t)(el - 1));
      c.h_obel.push_back((int32_t)(base[b] + (u - 1) * 3));
    }
  }
  c.h_boff[4] = (int64_t)c.h_obel.size() / 2;
  if (m->crnr_ufns) {
    c.h_crnr.assign(4 * RECP, 0.0);
    for (int k = 0; k < 4; ++k)
      std::memcpy(&c.h_crnr[k * RECP], m->crnr_ufns + REC * k, sizeof(double) * REC);
  }
  c.crnr_el[0] = 0;
  c.crnr_el[1] = (int32_t)(c.num1el - 1);
  c.crnr_el[2] = (int32_t)(numel - c.num1el);
  c.crnr_el[3] = (int32_t)(numel - 1);

  std::vector<EdgeItem> edges;
  for (int64_t k = 0; k < m->n_neu; ++k) {
    const int b = m->neu_bdry[k] - 1;
    for (int64_t j = 0; j < m->bdry_n_el[b]; ++j) {
      const int64_t el = m->bdry_elems[b][j], u = m->bdry_uel_ids[b][j];
      if (el < 1 || el > numel || u < 1 || u > m->bdry_n_uel[b])
        throw ArgError{"boundary table entry out of range", MAFEM_EINVAL};
      edges.push_back(EdgeItem{(int32_t)(el - 1), (int32_t)(base[b] + (u - 1) * 3), b + 1,
                               m->neu_type[k], m->neu_val[k]});
    }
  }
  c.n_edges = (int64_t)edges.size();

  // node-block pattern: nodes sharing an element (|Δi1| <= 2, |Δi2| <= 2)
  const int64_t n1 = c.num1np, n2 = c.num2np;
  std::vector<int32_t> rowptr(numnp + 1), colidx, brow;
  colidx.reserve(numnp * 25);
  for (int64_t p = 0; p < numnp; ++p) {
    rowptr[p] = (int32_t)colidx.size();
    const int64_t i1 = p % n1, i2 = p / n1;
    for (int64_t j2 = std::max<int64_t>(0, i2 - 2); j2 <= std::min(n2 - 1, i2 + 2); ++j2)
      for (int64_t j1 = std::max<int64_t>(0, i1 - 2); j1 <= std::min(n1 - 1, i1 + 2); ++j1) {
        colidx.push_back((int32_t)(j1 + n1 * j2));
        brow.push_back((int32_t)p);
      }
  }
  rowptr[numnp] = (int32_t)colidx.size();
  c.nnzb = (int64_t)colidx.size();
  auto blk_of = [&](int64_t p, int64_t q) -> int64_t {
    const int32_t *b = colidx.data() + rowptr[p], *e = colidx.data() + rowptr[p + 1];
    const int32_t *it = std::lower_bound(b, e, (int32_t)q);
    return rowptr[p] + (it - b);
  };

  // deterministic gather lists (counting sort by destination; sources in
  // element order, edges after area elements)
  const int64_t ntot = numel + c.n_edges;
  auto elem_nodes = [&](int64_t k, int a) -> int64_t {
    const int64_t e = k < numel ? k : edges[k - numel].el;
    return ix[e * 9 + a];
  };
  std::vector<int64_t> gptr(c.nnzb + 1, 0), rptr(numnp + 1, 0);
  for (int64_t k = 0; k < ntot; ++k)
    for (int a = 0; a < 9; ++a) {
      const int64_t pa = elem_nodes(k, a);
      rptr[pa + 1]++;
      for (int b = 0; b < 9; ++b) gptr[blk_of(pa, elem_nodes(k, b)) + 1]++;
    }
  for (int64_t i = 0; i < c.nnzb; ++i) gptr[i + 1] += gptr[i];
  for (int64_t i = 0; i < numnp; ++i) rptr[i + 1] += rptr[i];
  std::vector<int64_t> gsrc(gptr[c.nnzb]), rsrc(rptr[numnp]);
  {
    std::vector<int64_t> gfill(gptr.begin(), gptr.end() - 1), rfill(rptr.begin(), rptr.end() - 1);
    for (int64_t k = 0; k < ntot; ++k)
      for (int a = 0; a < 9; ++a) {
        const int64_t pa = elem_nodes(k, a);
        rsrc[rfill[pa]++] = k * ne + a * ndf;
        for (int b = 0; b < 9; ++b)
          gsrc[gfill[blk_of(pa, elem_nodes(k, b))]++] = k * ne * ne + (int64_t)a * ndf * ne + b * ndf;
      }
  }

  // kept for the CSC export, which is built on first use (ensure_csc)
  c.h_ID.assign(m->ID, m->ID + ndf * numnp);
  c.h_rowptr = rowptr;
  c.h_eq = eq_ext;
  c.nnz = -1;

  // band layout
  c.nx = numnp * ndf;
  c.nt = (c.nx + TILE - 1) / TILE;
  const int64_t kl = (2 * n1 + 3) * ndf - 1;
  c.tb = (int)((kl + TILE - 1) / TILE);
  if (c.tb > c.nt - 1) c.tb = (int)std::max<int64_t>(1, c.nt - 1);
  if (c.tb + 1 > SOLVE_RING)
    throw ArgError{"patch too wide for the band solver (num1el <= 262)", MAFEM_EUNSUPPORTED};
  // two-ended factorization (T down, B up, separator last) once the chain of
  // tile rows is long enough for halving it to pay; MAFEM_TWO_ENDED=0 keeps one
  // chain
  {
    const char *e = std::getenv("MAFEM_TWO_ENDED");
    if (c.nt >= 8 * (int64_t)c.tb && !(e && std::atoi(e) == 0)) {
      c.tw_a = (c.nt - c.tb) / 2;
      c.tw_nb = c.nt - c.tb - c.tw_a;
    }
  }
  c.band_bytes = c.band_rows() * (2 * c.tb + 1) * (int64_t)TILE * TILE * (int64_t)sizeof(double);
  // default ordering of the single-context factorization: nested dissection
  // where the band would be long enough to run two-ended (from about 16x16
  // elements up), else the band; MAFEM_ORDERING=band|nd overrides
  c.ordering = c.nt >= 8 * (int64_t)c.tb ? 1 : 0;
  if (const char *e = std::getenv("MAFEM_ORDERING")) {
    if (std::strcmp(e, "band") == 0) c.ordering = 0;
    else if (std::strcmp(e, "nd") == 0) c.ordering = 1;
  }

  // device allocations and uploads
  MAFEM_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  for (auto &e : c.ev) MAFEM_HIP(hipEventCreate(&e));
  c.d_xms = dalloc<double>(c, numnp * 3);
  c.d_cps = dalloc<double>(c, numnp * ndf);
  MAFEM_HIP(hipMemset(c.d_xms, 0, sizeof(double) * numnp * 3));
  MAFEM_HIP(hipMemset(c.d_cps, 0, sizeof(double) * numnp * ndf));
  c.d_ix = dupload(c, ix);
  c.d_uel = dupload(c, uel);
  c.d_active = dupload(c, active);
  c.d_ufns = dupload(c, ufns);
  c.d_dbt = dupload(c, dbt);
  c.d_bufns = dupload(c, bufns);
  c.d_edges = dupload(c, edges);
  c.d_kel = dalloc<double>(c, ntot * ne * ne);
  c.d_rel = dalloc<double>(c, ntot * ne);
  c.d_rowptr = dupload(c, rowptr);
  c.d_colidx = dupload(c, colidx);
  c.d_brow = dupload(c, brow);
  c.d_bsr = dalloc<double>(c, c.nnzb * ndf * ndf);
  c.d_gptr = dupload(c, gptr);
  c.d_gsrc = dupload(c, gsrc);
  c.d_rptr = dupload(c, rptr);
  c.d_rsrc = dupload(c, rsrc);
  const int64_t nv = c.nt * TILE;
  c.d_r = dalloc<double>(c, nv);
  c.d_x = dalloc<double>(c, nv);
  c.d_w = dalloc<double>(c, nv);
  c.d_t = dalloc<double>(c, nv);
  c.d_b = dalloc<double>(c, nv);
  c.d_t2 = dalloc<double>(c, nv);
  for (double *v : {c.d_r, c.d_x, c.d_w, c.d_t, c.d_b, c.d_t2})
    MAFEM_HIP(hipMemset(v, 0, sizeof(double) * nv));
  c.d_red = dalloc<double>(c, 512);
  MAFEM_HIP(hipHostMalloc((void **)&c.h_red, sizeof(double) * 16, hipHostMallocDefault));
  c.d_eq = dupload(c, eq_ext);
  c.opts = mafem_solver_opts{MAFEM_SOLVER_DIRECT, 64, 30, 200, 20, 3, 1e-13};
  c.d_piv = dalloc<int32_t>(c, c.band_rows() * TILE);
  // sweeps over several workgroups once the chain of tile rows is long
  const int64_t chain = c.tw_a > 0 ? std::min(c.tw_a, c.tw_nb) : c.nt;
  c.sweep_groups = (int)std::min<int64_t>(16, std::max<int64_t>(1, chain / 32));
  if (c.tw_a > 0) {
    const int64_t nr = (c.tw_nb + c.tb) * TILE;
    c.d_tw_z = dalloc<double>(c, nv);
    c.d_tw_br = dalloc<double>(c, nr);
    c.d_tw_cr = dalloc<double>(c, nr);
    c.d_tw_dr = dalloc<double>(c, nr);
    c.d_tw_s1 = dalloc<double>(c, nr);
    c.d_tw_s2 = dalloc<double>(c, nr);
  }
  if (const char *e = std::getenv("MAFEM_SWEEP_GROUPS")) {
    const int g = std::atoi(e);
    if (g >= 1 && g <= 64) c.sweep_groups = g;
  }
  c.d_s1 = dalloc<double>(c, nv);
  c.d_s2 = dalloc<double>(c, nv);
  c.d_sweep_err = dalloc<int>(c, 1);
  c.d_pub = dalloc<unsigned long long>(c, 4 * ((int64_t)TILE * TILE + TILE));
  c.d_span = dalloc<unsigned long long>(c, 2);
  {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c.dev) == hipSuccess && khz > 0)
      c.wall_khz = khz;
  }
  c.d_amax = dalloc<unsigned long long>(c, 2);
  MAFEM_HIP(hipMemset(c.d_amax, 0, 2 * sizeof(unsigned long long)));
  MAFEM_HIP(hipMemset(c.d_sweep_err, 0, sizeof(int)));
}

// A wait between workgroups of one launch (multi-workgroup sweeps, the
// factorization's look-ahead panels) that ran out leaves an error count (the
// results are invalid): checked at the end of every solve and time step.
void check_sweeps(Ctx &c) {
  int e = 0;
  MAFEM_HIP(hipMemcpyAsync(&e, c.d_sweep_err, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  MAFEM_HIP(hipStreamSynchronize(c.stream));
  if (e != 0) {
    MAFEM_HIP(hipMemsetAsync(c.d_sweep_err, 0, sizeof(int), c.stream));
    throw ArgError{"a wait inside a factorization or triangular sweep ran out: the solves "
                   "since the last check are invalid and so is the context state they "
                   "updated (set it again before continuing)", MAFEM_EHIP};
  }
}

// CSC export over active DOFs, columns in equation order (the pattern of
// K_gl, FiniteElement.jl:129-136, as a structural superset)
void ensure_csc(Ctx &c) {
  if (c.nnz >= 0) return;
  const int64_t n1 = c.num1np, n2 = c.num2np, nmdf = c.nmdf;
  const int ndf = c.ndf;
  auto blk_of = [&](int64_t p, int64_t q) -> int64_t {
    const int64_t i1 = p % n1, i2 = p / n1, j1 = q % n1, j2 = q / n1;
    const int64_t lo1 = std::max<int64_t>(0, i1 - 2), hi1 = std::min(n1 - 1, i1 + 2);
    const int64_t lo2 = std::max<int64_t>(0, i2 - 2);
    return c.h_rowptr[p] + (j2 - lo2) * (hi1 - lo1 + 1) + (j1 - lo1);
  };
  c.h_colptr.assign(nmdf + 1, 1);
  c.h_rowval.clear();
  std::vector<int64_t> csc_src;
  for (int64_t j = 0; j < nmdf; ++j) {
    c.h_colptr[j] = (int64_t)c.h_rowval.size() + 1;
    const int64_t kx = c.h_eq[j], q = kx / ndf, cc = kx % ndf;
    const int64_t q1 = q % n1, q2 = q / n1;
    for (int64_t p2 = std::max<int64_t>(0, q2 - 2); p2 <= std::min(n2 - 1, q2 + 2); ++p2)
      for (int64_t p1 = std::max<int64_t>(0, q1 - 2); p1 <= std::min(n1 - 1, q1 + 2); ++p1) {
        const int64_t pp = p1 + n1 * p2;
        const int64_t bk = blk_of(pp, q);
        for (int r = 0; r < ndf; ++r) {
          const int64_t id = c.h_ID[r + ndf * pp];
          if (id == 0) continue;
          c.h_rowval.push_back(id);
          csc_src.push_back(bk * ndf * ndf + r * ndf + cc);
        }
      }
  }
  c.h_colptr[nmdf] = (int64_t)c.h_rowval.size() + 1;
  const int64_t nnz = (int64_t)c.h_rowval.size();
  c.d_csc_src = dupload(c, csc_src);
  c.d_csc_val = dalloc<double>(c, nnz);
  c.nnz = nnz;
}

__global__ void k_gather_idx(int64_t n, const int64_t *__restrict__ idx,
                             const double *__restrict__ src, double *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) dst[t] = src[idx[t]];
}

__global__ void k_scatter_idx(int64_t n, const int64_t *__restrict__ idx,
                              const double *__restrict__ src, double *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) dst[idx[t]] = src[t];
}

inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

float elapsed(Ctx &c, int a, int b) {
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c.ev[a], c.ev[b]);
  return ms;
}

// Solves K x = b (extended vectors) with the factored band and FP64
// iterative refinement against the block-sparse K; returns |b - K x| / |b|.
// Refinement stops at tol, after max_sweeps, or once a sweep no longer halves
// the residual (it has reached the rounding floor of K x for this b, as
// LAPACK's and UMFPACK's refinement stop).
double solve_refined(Ctx &c, const double *d_b, double *d_x, int max_sweeps, double tol) {
  const double bn = vec_norm2_masked(c, d_b);
  band_solve(c, d_b, d_x);
  double rel = 0.0, prev = 0.0;
  int sweeps = 0;
  if (bn == 0.0) {
    c.stats.last_relres = 0.0;
    c.stats.refine_sweeps = 0;
    return 0.0;
  }
  for (;;) {
    bsr_spmv(c, d_x, c.d_w);
    vec_axpby(c, c.nx, 1.0, d_b, -1.0, c.d_w);  // w = b - K x
    rel = vec_norm2_masked(c, c.d_w) / bn;
    if (rel < tol || sweeps >= max_sweeps || (sweeps > 0 && rel > 0.5 * prev)) break;
    prev = rel;
    band_solve(c, c.d_w, c.d_t);
    vec_axpby(c, c.nx, 1.0, c.d_t, 1.0, d_x);
    ++sweeps;
  }
  c.stats.last_relres = rel;
  c.stats.refine_sweeps = sweeps;
  return rel;
}

void fill_ordering_stats(const Ctx &c, mafem_stats &st) {
  const bool nd = c.ordering == 1;
  st.factor_chains = nd ? 0 : (c.tw_a > 0 ? 2 : 1);
  st.separator_rows = !nd && c.tw_a > 0 ? c.tb : 0;
  st.fronts = nd ? nd_fronts(c) : 0;
  st.front_steps = nd ? nd_steps(c) : 0;
  st.front_flops = nd ? nd_flops(c) : 0.0;
}

// algorithmic FLOPs of the band's timed step launch (a middle step: per chain
// tb^2 tile products, the next diagonal LU and 2 tb panel solves)
double band_step_flops(const Ctx &c) {
  const double t3 = (double)TILE * TILE * TILE, tb = c.tb;
  return (c.tw_a > 0 ? 2.0 : 1.0) * (tb * tb * 2.0 * t3 + 2.0 / 3.0 * t3 + 2.0 * tb * t3);
}

// storage precision of the factorization the context holds
int factor_prec(const Ctx &c) { return c.ordering == 1 ? nd_prec(c) : c.band_prec; }

// Autotuner of the sweep split: on the first factorization of a context whose
// sweeps run over several workgroups, time one band solve for each candidate
// count (2, 4, 8, 16; at most one per four tile rows) and keep the fastest.
// All candidates give bit-identical results (the per-row summation order does
// not depend on the count), so the choice affects time only.  An
// MAFEM_SWEEP_GROUPS setting disables the tuning.
void tune_sweeps(Ctx &c, const double *d_b) {
  if (c.ordering == 1) return;  // the fronts' sweeps have no split to tune
  c.sweep_tuned = true;
  if (c.sweep_groups <= 1 || std::getenv("MAFEM_SWEEP_GROUPS")) return;
  float best = 1e30f;
  int best_w = c.sweep_groups;
  const int64_t chain = c.tw_a > 0 ? std::min(c.tw_a, c.tw_nb) : c.nt;
  for (int w : {2, 4, 8, 16}) {
    if (w > chain / 4) break;
    c.sweep_groups = w;
    band_solve(c, d_b, c.d_t);  // warm
    MAFEM_HIP(hipEventRecord(c.ev[6], c.stream));
    band_solve(c, d_b, c.d_t);
    MAFEM_HIP(hipEventRecord(c.ev[7], c.stream));
    MAFEM_HIP(hipEventSynchronize(c.ev[7]));
    float ms = 0.f;
    MAFEM_HIP(hipEventElapsedTime(&ms, c.ev[6], c.ev[7]));
    if (ms < best) {
      best = ms;
      best_w = w;
    }
  }
  c.sweep_groups = best_w;
}

void refactor(Ctx &c) {
  band_fill(c, c.opts.factor_precision);
  band_factor(c, c.perturb_pivots);
  c.have_factor = true;
}

// Solves K x = b with the configured solver; events ev[2] (start) and ev[3]
// (after any up-front factorization) bracket the factorization phase.
double linear_solve(Ctx &c, const double *d_b, double *d_x) {
  const mafem_solver_opts &o = c.opts;
  MAFEM_HIP(hipEventRecord(c.ev[2], c.stream));
  if (o.mode == MAFEM_SOLVER_DIRECT) {
    refactor(c);
    if (!c.sweep_tuned) tune_sweeps(c, d_b);
    MAFEM_HIP(hipEventRecord(c.ev[3], c.stream));
    double rel = solve_refined(c, d_b, d_x, o.refine_sweeps, o.rtol);
    if (rel <= 1e3 * o.rtol) return rel;
    if (!c.perturb_pivots && !(rel <= 1e-2)) {
      // a diagonal tile met a (numerically) zero pivot: from now on factor
      // with static pivoting and solve again
      c.perturb_pivots = true;
      refactor(c);
      rel = solve_refined(c, d_b, d_x, o.refine_sweeps, o.rtol);
      if (rel <= 1e3 * o.rtol) return rel;
    }
    // refinement stalls when the tile-local pivoting grew the factors: finish
    // with FGMRES preconditioned by the same fresh factorization
    ensure_krylov(c);
    int its = 0;
    const double rel2 = fgmres(c, d_b, d_x, &its);
    c.stats.last_relres = rel2;
    return rel2;
  }
  bool fresh = false;
  if (!c.have_factor || factor_prec(c) != o.factor_precision) {
    refactor(c);
    fresh = true;
  }
  if (!c.sweep_tuned) tune_sweeps(c, d_b);
  MAFEM_HIP(hipEventRecord(c.ev[3], c.stream));
  int its = 0;
  double rel = fgmres(c, d_b, d_x, &its);
  if (!(rel <= o.rtol) && !fresh) {
    // the kept preconditioner no longer suffices: refactor from this tangent
    refactor(c);
    int its2 = 0;
    rel = fgmres(c, d_b, d_x, &its2);
    its += its2;
  } else if (its > o.refactor_iters) {
    c.have_factor = false;  // refactor before the next solve
  }
  c.stats.last_relres = rel;
  return rel;
}

}  // namespace

// ------------------------------------------------ allocation-free pull force
//
// Scratch of mafem_pull_force_noalloc: the adjacent-element list, the node
// slot maps, the adjacent elements' residuals and the 6 outputs.  Each device
// the library runs on holds one copy until the process ends, shared by every
// call on that device (calls that overlap on one device race).  Host code
// never passes `scratch` itself to a kernel or copy: its name is a host-side
// stand-in; the device address comes from hipGetSymbolAddress, taken once
// per device with that device current.
constexpr size_t kScratchBytes = 64 * 1024;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxAdj = 25;  // (2 poly + 1)^2 adjacent elements for poly = 2

struct PullMaps {
  int32_t n;
  int32_t el[kMaxAdj];
  int8_t slot[9 * kMaxAdj];
};

__global__ void k_stage_pull_maps(PullMaps m, int32_t *__restrict__ el, int8_t *__restrict__ slot) {
  for (int t = threadIdx.x; t < 9 * m.n; t += blockDim.x) {
    slot[t] = m.slot[t];
    if (t < m.n) el[t] = m.el[t];
  }
}

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_addr[kMaxDevices] = {};

// device address of `scratch` on device dev (current), looked up once per device
unsigned char *scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices)
    throw ArgError{"device id beyond the scratch address table", MAFEM_EUNSUPPORTED};
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    MAFEM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw ArgError{"no device address for the scratch array", MAFEM_EHIP};
    g_scratch_addr[dev] = p;
  }
  return (unsigned char *)g_scratch_addr[dev];
}

}  // namespace

#define ABI_GUARD(ctx, ...)                                    \
  do {                                                         \
    if (!(ctx)) return MAFEM_EINVAL;                           \
    Ctx &c = (ctx)->c;                                         \
    try {                                                      \
      __VA_ARGS__                                              \
    } catch (const HipError &h) {                              \
      return from_hip(c, h);                                   \
    } catch (const ArgError &a) {                              \
      return fail(c, a.msg, a.code);                           \
    } catch (const std::bad_alloc &) {                         \
      return fail(c, "host allocation failed", MAFEM_ENOMEM);  \
    }                                                          \
  } while (0)

extern "C" {

int mafem_create(mafem_ctx **out, const mafem_mesh_desc *mesh, const mafem_params *p) {
  if (!out) return MAFEM_EINVAL;
  *out = nullptr;
  mafem_ctx *ctx = new (std::nothrow) mafem_ctx();
  if (!ctx) {
    g_create_err = "host allocation failed";
    return MAFEM_ENOMEM;
  }
  int rc = MAFEM_OK;
  std::lock_guard<std::mutex> no_capture(capture_mutex());
  try {
    int ndev = 0;
    MAFEM_HIP(hipGetDeviceCount(&ndev));
    if (ndev < 1) throw ArgError{"no HIP device", MAFEM_EHIP};
    MAFEM_HIP(hipGetDevice(&ctx->c.dev));
    build(ctx->c, mesh, p);
  } catch (const HipError &h) {
    rc = from_hip(ctx->c, h);
  } catch (const ArgError &a) {
    rc = fail(ctx->c, a.msg, a.code);
  } catch (const std::bad_alloc &) {
    rc = fail(ctx->c, "host allocation failed", MAFEM_ENOMEM);
  }
  if (rc != MAFEM_OK) {
    g_create_err = ctx->c.err;
    free_all(ctx->c);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return MAFEM_OK;
}

void mafem_destroy(mafem_ctx *ctx) {
  if (!ctx) return;
  std::lock_guard<std::mutex> no_capture(capture_mutex());
  (void)hipSetDevice(ctx->c.dev);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  free_all(ctx->c);
  delete ctx;
}

const char *mafem_last_error(const mafem_ctx *ctx) {
  return ctx ? ctx->c.err.c_str() : g_create_err.c_str();
}

int mafem_set_state(mafem_ctx *ctx, const double *xms, const double *cps) {
  ABI_GUARD(ctx, {
    std::lock_guard<std::mutex> no_capture(capture_mutex());
    if (!xms || !cps) return fail(c, "null state", MAFEM_EINVAL);
    MAFEM_HIP(hipSetDevice(c.dev));
    MAFEM_HIP(hipMemcpyAsync(c.d_xms, xms, sizeof(double) * c.numnp * 3, hipMemcpyHostToDevice, c.stream));
    MAFEM_HIP(hipMemcpyAsync(c.d_cps, cps, sizeof(double) * c.numnp * c.ndf, hipMemcpyHostToDevice, c.stream));
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    return MAFEM_OK;
  });
}

int mafem_get_state(mafem_ctx *ctx, double *xms, double *cps) {
  ABI_GUARD(ctx, {
    std::lock_guard<std::mutex> no_capture(capture_mutex());
    MAFEM_HIP(hipSetDevice(c.dev));
    if (xms)
      MAFEM_HIP(hipMemcpyAsync(xms, c.d_xms, sizeof(double) * c.numnp * 3, hipMemcpyDeviceToHost, c.stream));
    if (cps)
      MAFEM_HIP(hipMemcpyAsync(cps, c.d_cps, sizeof(double) * c.numnp * c.ndf, hipMemcpyDeviceToHost, c.stream));
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    return MAFEM_OK;
  });
}

int mafem_set_prescribed(mafem_ctx *ctx, int64_t n, const int64_t *node, const int32_t *dof,
                         const double *val) {
  ABI_GUARD(ctx, {
    std::lock_guard<std::mutex> no_capture(capture_mutex());
    if (n < 0 || (n > 0 && (!node || !dof || !val))) return fail(c, "bad prescribed list", MAFEM_EINVAL);
    MAFEM_HIP(hipSetDevice(c.dev));
    std::vector<int64_t> idx(n);
    for (int64_t k = 0; k < n; ++k) {
      if (node[k] < 1 || node[k] > c.numnp || dof[k] < 1 || dof[k] > c.ndf)
        return fail(c, "prescribed entry out of range", MAFEM_EINVAL);
      idx[k] = (node[k] - 1) + c.numnp * (dof[k] - 1);
    }
    if (n == 0) return MAFEM_OK;
    if (n > c.presc_cap) {  // staging buffers grow on demand and are kept
      if (c.d_presc_idx) (void)hipFree(c.d_presc_idx);
      if (c.d_presc_val) (void)hipFree(c.d_presc_val);
      c.d_presc_idx = nullptr;
      c.d_presc_val = nullptr;
      c.presc_cap = 0;
      c.d_presc_idx = dalloc<int64_t>(c, n);
      c.d_presc_val = dalloc<double>(c, n);
      c.presc_cap = n;
    }
    MAFEM_HIP(hipMemcpyAsync(c.d_presc_idx, idx.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
                             c.stream));
    MAFEM_HIP(hipMemcpyAsync(c.d_presc_val, val, sizeof(double) * n, hipMemcpyHostToDevice,
                             c.stream));
    hipLaunchKernelGGL(k_scatter_idx, dim3(nb256(n)), dim3(256), 0, c.stream, n, c.d_presc_idx,
                       c.d_presc_val, c.d_cps);
    MAFEM_HIP(hipGetLastError());
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    return MAFEM_OK;
  });
}

int mafem_set_dof_ids(mafem_ctx *ctx, const int64_t *ID, int64_t nmdf) {
  ABI_GUARD(ctx, {
    std::lock_guard<std::mutex> no_capture(capture_mutex());
    if (!ID || nmdf < 1) return fail(c, "null ID or nmdf < 1", MAFEM_EINVAL);
    const int64_t n = c.ndf * c.numnp;
    std::vector<uint8_t> active(n);
    std::vector<int64_t> eq(nmdf, -1);
    for (int64_t k = 0; k < n; ++k) {
      const int64_t id = ID[k];
      if (id < 0 || id > nmdf || (id > 0 && eq[id - 1] >= 0))
        return fail(c, "ID is not a numbering of 1..nmdf", MAFEM_EINVAL);
      active[k] = id != 0;
      if (id > 0) eq[id - 1] = k;
    }
    for (int64_t v : eq)
      if (v < 0) return fail(c, "ID is not a numbering of 1..nmdf", MAFEM_EINVAL);
    MAFEM_HIP(hipSetDevice(c.dev));
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    MAFEM_HIP(hipMemcpy(c.d_active, active.data(), n, hipMemcpyHostToDevice));
    if (nmdf != c.nmdf) {
      if (c.d_eq) (void)hipFree(c.d_eq);
      c.d_eq = nullptr;
      c.d_eq = dalloc<int64_t>(c, nmdf);
    }
    MAFEM_HIP(hipMemcpy(c.d_eq, eq.data(), sizeof(int64_t) * nmdf, hipMemcpyHostToDevice));
    c.nmdf = nmdf;
    c.h_ID.assign(ID, ID + n);
    c.h_eq = eq;
    // the CSC export pattern follows the numbering: rebuilt on next use
    if (c.d_csc_src) (void)hipFree(c.d_csc_src);
    if (c.d_csc_val) (void)hipFree(c.d_csc_val);
    c.d_csc_src = nullptr;
    c.d_csc_val = nullptr;
    c.nnz = -1;
    c.h_colptr.clear();
    c.h_rowval.clear();
    c.have_factor = false;  // a kept preconditioner belongs to the old Dirichlet set
    return MAFEM_OK;
  });
}

int mafem_predict(mafem_ctx *ctx, double dt) {
  ABI_GUARD(ctx, {
    MAFEM_HIP(hipSetDevice(c.dev));
    launch_predict(c, c.d_cps, dt);
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    return MAFEM_OK;
  });
}

int mafem_nnz(mafem_ctx *ctx, int64_t *nnz) {
  ABI_GUARD(ctx, {
    if (!nnz) return fail(c, "null nnz", MAFEM_EINVAL);
    ensure_csc(c);
    *nnz = c.nnz;
    return MAFEM_OK;
  });
}

int mafem_assemble(mafem_ctx *ctx, double time, double dt, double *r, int64_t *colptr,
                   int64_t *rowval, double *nzval) {
  ABI_GUARD(ctx, {
    MAFEM_HIP(hipSetDevice(c.dev));
    if (colptr || rowval || nzval) ensure_csc(c);
    MAFEM_HIP(hipEventRecord(c.ev[0], c.stream));
    launch_assemble(c, time, dt);
    MAFEM_HIP(hipEventRecord(c.ev[1], c.stream));
    if (r) {
      hipLaunchKernelGGL(k_gather_idx, dim3(nb256(c.nmdf)), dim3(256), 0, c.stream, c.nmdf, c.d_eq,
                         c.d_r, c.d_w);
      MAFEM_HIP(hipGetLastError());
      MAFEM_HIP(hipMemcpyAsync(r, c.d_w, sizeof(double) * c.nmdf, hipMemcpyDeviceToHost, c.stream));
    }
    if (nzval) {
      hipLaunchKernelGGL(k_gather_idx, dim3(nb256(c.nnz)), dim3(256), 0, c.stream, c.nnz,
                         c.d_csc_src, c.d_bsr, c.d_csc_val);
      MAFEM_HIP(hipGetLastError());
      MAFEM_HIP(hipMemcpyAsync(nzval, c.d_csc_val, sizeof(double) * c.nnz, hipMemcpyDeviceToHost, c.stream));
    }
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    if (colptr) std::memcpy(colptr, c.h_colptr.data(), sizeof(int64_t) * (c.nmdf + 1));
    if (rowval) std::memcpy(rowval, c.h_rowval.data(), sizeof(int64_t) * c.nnz);
    c.stats = mafem_stats{};
    c.stats.assemble_us = 1e3 * elapsed(c, 0, 1);
    c.stats.area_kernel_us = 1e3 * elapsed(c, 8, 9);
    c.stats.n_assemble = 1;
    return MAFEM_OK;
  });
}

int mafem_solve(mafem_ctx *ctx, const double *rhs, double *du, double *relres) {
  ABI_GUARD(ctx, {
    if (!rhs || !du) return fail(c, "null vector", MAFEM_EINVAL);
    MAFEM_HIP(hipSetDevice(c.dev));
    MAFEM_HIP(hipMemcpy(c.d_t, rhs, sizeof(double) * c.nmdf, hipMemcpyHostToDevice));
    vec_zero(c, c.nt * TILE, c.d_b);
    hipLaunchKernelGGL(k_scatter_idx, dim3(nb256(c.nmdf)), dim3(256), 0, c.stream, c.nmdf, c.d_eq,
                       c.d_t, c.d_b);
    MAFEM_HIP(hipGetLastError());
    const int32_t nf0 = c.n_factor, ng0 = c.n_gmres;
    const double rel = linear_solve(c, c.d_b, c.d_x);
    MAFEM_HIP(hipEventRecord(c.ev[4], c.stream));
    hipLaunchKernelGGL(k_gather_idx, dim3(nb256(c.nmdf)), dim3(256), 0, c.stream, c.nmdf, c.d_eq,
                       c.d_x, c.d_t);
    MAFEM_HIP(hipGetLastError());
    MAFEM_HIP(hipMemcpyAsync(du, c.d_t, sizeof(double) * c.nmdf, hipMemcpyDeviceToHost, c.stream));
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    check_sweeps(c);
    if (relres) *relres = rel;
    c.stats.factor_us = 1e3 * elapsed(c, 2, 3);
    c.stats.solve_us = 1e3 * elapsed(c, 3, 4);
    c.stats.n_factor = c.n_factor - nf0;
    c.stats.gmres_iters = c.n_gmres - ng0;
    c.stats.sweep_groups = c.sweep_groups;
    fill_ordering_stats(c, c.stats);
    return MAFEM_OK;
  });
}

int mafem_time_step(mafem_ctx *ctx, double time, double dt, int32_t max_iter, int32_t *iters,
                    double *eps_hist) {
  ABI_GUARD(ctx, {
    if (max_iter < 1) return fail(c, "max_iter must be >= 1", MAFEM_EINVAL);
    MAFEM_HIP(hipSetDevice(c.dev));
    mafem_stats st{};
    double eps = INFINITY;
    int it = 0;
    const int32_t nbs0 = c.n_band_solve, nf0 = c.n_factor, ng0 = c.n_gmres;
    const int32_t np0 = read_pert(c);
    int n_step = 0, n_sch = 0;
    double step_fl = 0.0, step_by = 0.0;
    c.step_timed = false;
    MAFEM_HIP(hipEventRecord(c.ev[0], c.stream));
    while (it < max_iter) {
      MAFEM_HIP(hipEventRecord(c.ev[1], c.stream));
      launch_assemble(c, time, dt);
      vec_scale(c, c.nt * TILE, -1.0, c.d_r, c.d_b);  // b = -r
      linear_solve(c, c.d_b, c.d_x);
      MAFEM_HIP(hipEventRecord(c.ev[4], c.stream));
      apply_update(c, c.d_x, dt);
      MAFEM_HIP(hipEventRecord(c.ev[5], c.stream));
      eps = vec_norm2_masked(c, c.d_x) / (double)c.nmdf;  // synchronizes
      st.assemble_us += 1e3 * elapsed(c, 1, 2);
      st.area_kernel_us += 1e3 * elapsed(c, 8, 9);
      st.n_assemble += 1;
      st.factor_us += 1e3 * elapsed(c, 2, 3);
      st.solve_us += 1e3 * elapsed(c, 3, 4);
      st.update_us += 1e3 * elapsed(c, 4, 5);
      st.refine_sweeps += c.stats.refine_sweeps;
      if (c.step_timed) {  // on the context's stream: another thread may be capturing a graph
        unsigned long long sp[2];
        MAFEM_HIP(hipMemcpyAsync(sp, c.d_span, sizeof sp, hipMemcpyDeviceToHost, c.stream));
        MAFEM_HIP(hipStreamSynchronize(c.stream));
        if (sp[1] > sp[0]) {
          st.step_kernel_us += 1e3 * (double)(sp[1] - sp[0]) / c.wall_khz;
          ++n_step;
          step_fl += band_step_flops(c);
        }
        c.step_timed = false;
      }
      if (c.ordering == 1) {
        NdSpans sa, sb;
        if (nd_read_spans(c, sa, sb)) {
          st.step_kernel_us += sa.us;
          n_step += sa.count;
          step_fl += sa.flops;
          step_by += sa.bytes;
          st.schur_kernel_us += sb.us;
          n_sch += sb.count;
          st.schur_flops += sb.flops;
          st.schur_bytes += sb.bytes;
        }
      }
      st.last_relres = c.stats.last_relres;
      if (eps_hist) eps_hist[it] = eps;
      ++it;
      if (!std::isfinite(eps)) break;
      if (eps < c.enr) break;
    }
    MAFEM_HIP(hipEventRecord(c.ev[6], c.stream));
    MAFEM_HIP(hipEventSynchronize(c.ev[6]));
    check_sweeps(c);
    st.total_us = 1e3 * elapsed(c, 0, 6);
    st.newton_iters = it;
    st.n_band_solve = c.n_band_solve - nbs0;
    st.n_factor = c.n_factor - nf0;
    st.gmres_iters = c.n_gmres - ng0;
    if (n_step > 0) {
      st.step_kernel_us /= n_step;
      st.step_flops = step_fl / n_step;
      st.step_bytes = step_by / n_step;
    }
    if (n_sch > 0) {
      st.schur_kernel_us /= n_sch;
      st.schur_flops /= n_sch;
      st.schur_bytes /= n_sch;
    }
    st.sweep_groups = c.sweep_groups;
    st.pivot_perturbations = read_pert(c) - np0;
    fill_ordering_stats(c, st);
    c.stats = st;
    if (iters) *iters = it;
    if (!(eps < c.enr)) {
      c.err = "did not reach Newton--Raphson tolerance";
      return MAFEM_NOT_CONVERGED;
    }
    return MAFEM_OK;
  });
}

int mafem_pull_force(mafem_ctx *ctx, double *x_pull, double *f_pull) {
  ABI_GUARD(ctx, {
    if (!x_pull || !f_pull) return fail(c, "null output", MAFEM_EINVAL);
    if (!c.pf_err.empty()) return fail(c, c.pf_err, MAFEM_EINVAL);
    if (c.motion == STATIC) return fail(c, "f_pull needs 3-D velocity", MAFEM_EINVAL);
    MAFEM_HIP(hipSetDevice(c.dev));
    ensure_pull(c);
    launch_pull_force(c, c.stream, c.d_pf_el, c.d_pf_slot, c.d_pf_rel, c.d_pf_out);
    MAFEM_HIP(hipMemcpyAsync(c.h_red, c.d_pf_out, sizeof(double) * 6, hipMemcpyDeviceToHost,
                             c.stream));
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < 3; ++k) {
      x_pull[k] = c.h_red[k];
      f_pull[k] = c.h_red[3 + k];
    }
    return MAFEM_OK;
  });
}

int mafem_pull_force_noalloc(mafem_ctx *ctx, double *x_pull, double *f_pull) {
  ABI_GUARD(ctx, {
    if (!x_pull || !f_pull) return fail(c, "null output", MAFEM_EINVAL);
    if (!c.pf_err.empty()) return fail(c, c.pf_err, MAFEM_EINVAL);
    if (c.motion == STATIC) return fail(c, "f_pull needs 3-D velocity", MAFEM_EINVAL);
    const int64_t n = (int64_t)c.h_pf_el.size();
    // scratch layout: el[n] at 0, slot[9 n] at 128, rel[n ne] at 384, out[6] after it
    const size_t need = 384 + sizeof(double) * ((size_t)n * c.ne + 6);
    if (n > kMaxAdj || need > kScratchBytes)
      return fail(c,
                  "pull force scratch: " + std::to_string(need) + " bytes needed, " +
                      std::to_string(kScratchBytes) + " available (" + std::to_string(n) +
                      " adjacent elements of " + std::to_string(c.ne) + " residual rows)",
                  MAFEM_EUNSUPPORTED);
    MAFEM_HIP(hipSetDevice(c.dev));
    unsigned char *base = scratch_on(c.dev);
    int32_t *el = (int32_t *)base;
    int8_t *slot = (int8_t *)(base + 128);
    double *rel = (double *)(base + 384);
    double *out = rel + (size_t)n * c.ne;
    PullMaps m{};
    m.n = (int32_t)n;
    for (int64_t k = 0; k < n; ++k) m.el[k] = c.h_pf_el[k];
    for (int64_t k = 0; k < 9 * n; ++k) m.slot[k] = c.h_pf_slot[k];
    hipLaunchKernelGGL(k_stage_pull_maps, dim3(1), dim3(256), 0, c.stream, m, el, slot);
    MAFEM_HIP(hipGetLastError());
    launch_pull_force(c, c.stream, el, slot, rel, out);
    MAFEM_HIP(hipMemcpyAsync(c.h_red, out, sizeof(double) * 6, hipMemcpyDeviceToHost, c.stream));
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < 3; ++k) {
      x_pull[k] = c.h_red[k];
      f_pull[k] = c.h_red[3 + k];
    }
    return MAFEM_OK;
  });
}

int mafem_generate_output(mafem_ctx *ctx, double *xout, double *uout) {
  ABI_GUARD(ctx, {
    if (c.h_crnr.empty()) return fail(c, "corner basis tables (crnr_ufns) not given", MAFEM_EINVAL);
    const int64_t want[4] = {c.num1el, c.num2el, c.num1el, c.num2el};
    for (int b = 0; b < 4; ++b)
      if (c.h_boff[b + 1] - c.h_boff[b] != want[b])
        return fail(c, "boundary element lists do not cover the edges", MAFEM_EINVAL);
    MAFEM_HIP(hipSetDevice(c.dev));
    const int64_t npts = (3 * c.num1el + 2) * (3 * c.num2el + 2);
    if (!c.d_out) {
      std::lock_guard<std::mutex> no_capture(capture_mutex());
      c.d_obel = dupload(c, c.h_obel);
      c.d_crnr = dupload(c, c.h_crnr);
      c.d_out = dalloc<double>(c, npts * (3 + c.ndf));
    }
    launch_output(c);
    if (xout)
      MAFEM_HIP(hipMemcpyAsync(xout, c.d_out, sizeof(double) * npts * 3, hipMemcpyDeviceToHost,
                               c.stream));
    if (uout)
      MAFEM_HIP(hipMemcpyAsync(uout, c.d_out + npts * 3, sizeof(double) * npts * c.ndf,
                               hipMemcpyDeviceToHost, c.stream));
    MAFEM_HIP(hipStreamSynchronize(c.stream));
    return MAFEM_OK;
  });
}

int mafem_get_stats(const mafem_ctx *ctx, mafem_stats *out) {
  if (!ctx || !out) return MAFEM_EINVAL;
  *out = ctx->c.stats;
  return MAFEM_OK;
}

int mafem_set_solver(mafem_ctx *ctx, const mafem_solver_opts *o) {
  ABI_GUARD(ctx, {
    if (!o) return fail(c, "null solver options", MAFEM_EINVAL);
    if ((o->mode != MAFEM_SOLVER_DIRECT && o->mode != MAFEM_SOLVER_KRYLOV) ||
        (o->factor_precision != 64 && o->factor_precision != 32) || o->restart < 1 ||
        o->restart > 64 || o->max_iters < 1 || o->refine_sweeps < 0 || !(o->rtol > 0.0))
      return fail(c, "invalid solver options", MAFEM_EINVAL);
    if (o->factor_precision != c.opts.factor_precision) c.have_factor = false;
    c.opts = *o;
    return MAFEM_OK;
  });
}

int mafem_set_ordering(mafem_ctx *ctx, int32_t ordering) {
  ABI_GUARD(ctx, {
    if (ordering != MAFEM_ORDER_BAND && ordering != MAFEM_ORDER_ND)
      return fail(c, "invalid ordering", MAFEM_EINVAL);
    if (ordering != c.ordering) c.have_factor = false;
    c.ordering = ordering;
    return MAFEM_OK;
  });
}

int mafem_get_solver(const mafem_ctx *ctx, mafem_solver_opts *o) {
  if (!ctx || !o) return MAFEM_EINVAL;
  *o = ctx->c.opts;
  return MAFEM_OK;
}

}  // extern "C"
