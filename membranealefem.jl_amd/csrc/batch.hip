// Batches: many independent patches advanced through one time step together
// (declared in include/mafem.h).  Each member is an ordinary context; the
// batch builds descriptors of their device buffers, the block -> patch maps
// of every batched launch and the per-step class tables of the batched band
// factorization, then drives the Newton iterations of all members with one
// launch per phase.  The launch sequences are recorded once as HIP graphs
// (assembly, factorization, refinement, update) and replayed; per-patch
// flags in device memory decide which patches each replay touches, so the
// graphs stay valid for every iteration and step.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "internal.h"

using namespace mafem;

struct mafem_batch {
  BatchPlan B;
  std::vector<mafem_ctx *> members;
  std::vector<PatchDev> h_pd;
  mafem_batch_opts opts{};
  std::string err;
  int dev = 0;
  hipEvent_t ev[8] = {};
  int32_t *h_count = nullptr;  // pinned mirror of B.f.count
  hipGraphExec_t g_asm = nullptr, g_fac = nullptr, g_cyc = nullptr, g_end = nullptr,
                 g_upd = nullptr;
  std::vector<void *> owned;   // device allocations of the plan (rebuilt with it)
  std::vector<void *> fowned;  // device allocations of the flags (live with the batch)
  std::vector<int> prec;       // band storage precision of each member
  mafem_batch_stats stats{};
};

namespace {

struct BatchError {
  std::string msg;
  int code;
};

template <class T>
T *balloc(mafem_batch &b, int64_t n) {
  void *p = nullptr;
  if (n <= 0) n = 1;
  if (hipMalloc(&p, sizeof(T) * (size_t)n) != hipSuccess) {
    (void)hipGetLastError();
    throw BatchError{"device allocation failed", MAFEM_ENOMEM};
  }
  b.owned.push_back(p);
  return (T *)p;
}

template <class T>
T *falloc(mafem_batch &b, int64_t n) {
  void *p = nullptr;
  if (n <= 0) n = 1;
  if (hipMalloc(&p, sizeof(T) * (size_t)n) != hipSuccess) {
    (void)hipGetLastError();
    throw BatchError{"device allocation failed", MAFEM_ENOMEM};
  }
  b.fowned.push_back(p);
  return (T *)p;
}

template <class T>
T *bupload(mafem_batch &b, const std::vector<T> &v) {
  T *p = balloc<T>(b, (int64_t)v.size());
  if (!v.empty()) MAFEM_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return p;
}

void destroy_graphs(mafem_batch &b) {
  for (hipGraphExec_t *g : {&b.g_asm, &b.g_fac, &b.g_cyc, &b.g_end, &b.g_upd})
    if (*g) {
      (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
}

void free_plan(mafem_batch &b) {
  destroy_graphs(b);
  for (void *p : b.owned) (void)hipFree(p);
  b.owned.clear();
}

// Segment map from per-patch block counts (0 = patch not in this launch);
// `threads` is the workgroup size of every kernel launched over the segment.
Seg make_seg(mafem_batch &b, const std::vector<int64_t> &blocks, unsigned threads) {
  uint64_t total = 0;
  for (int64_t n : blocks) total += n > 0 ? (uint64_t)n : 0;
  if (!launch_fits(total, threads))
    throw BatchError{"batch too large for one launch: " + std::to_string(total) + " workgroups of " +
                         std::to_string(threads) + " work-items exceed a dispatch's 2^32 - 1",
                     MAFEM_EUNSUPPORTED};
  std::vector<int32_t> map;
  std::vector<int64_t> blk0(blocks.size(), -1);
  for (size_t p = 0; p < blocks.size(); ++p) {
    if (blocks[p] <= 0) continue;
    blk0[p] = (int64_t)map.size();
    map.insert(map.end(), (size_t)blocks[p], (int32_t)p);
  }
  Seg s;
  s.nblk = (unsigned)map.size();
  s.map = bupload(b, map);
  s.blk0 = bupload(b, blk0);
  return s;
}

PatchDev describe(const Ctx &c) {
  PatchDev d{};
  d.numel = c.numel; d.numnp = c.numnp; d.nx = c.nx; d.nnzb = c.nnzb;
  d.n_edges = c.n_edges; d.nt = c.nt;
  d.ndf = c.ndf; d.ne = c.ne; d.tb = c.tb; d.motion = c.motion;
  for (int i = 0; i < 3; ++i) d.mm[i] = c.mm_order[i];
  d.ph = c.phys; d.enr = c.enr; d.nmdf = (double)c.nmdf;
  d.ix = c.d_ix; d.uel = c.d_uel; d.active = c.d_active; d.ufns = c.d_ufns; d.dbt = c.d_dbt;
  d.bufns = c.d_bufns; d.edges = c.d_edges;
  d.xms = c.d_xms; d.cps = c.d_cps; d.kel = c.d_kel; d.rel = c.d_rel;
  d.rowptr = c.d_rowptr; d.colidx = c.d_colidx; d.brow = c.d_brow; d.bsr = c.d_bsr;
  d.gptr = c.d_gptr; d.gsrc = c.d_gsrc; d.rptr = c.d_rptr; d.rsrc = c.d_rsrc;
  d.r = c.d_r; d.b = c.d_b; d.x = c.d_x; d.w = c.d_w; d.t = c.d_t; d.t2 = c.d_t2;
  d.band = c.d_band; d.linv = c.d_linv; d.uinv = c.d_uinv; d.piv = c.d_piv;
  d.prec = c.band_prec;
  d.sweep_w = 1;
  d.s1 = c.d_s1;
  d.s2 = c.d_s2;
  d.err = c.d_sweep_err;
  d.amax = c.d_amax;
  return d;
}

// Workgroups of each patch's sweeps.  A split sweep is bound by the per-row
// hand-off (~6 us) once each workgroup streams its rows faster than that:
// 3 workgroups for FP32 storage, 4 for FP64 are enough.  The longest chains
// (at least 96 tile rows) get a split first, while the total stays within
// what the device keeps resident at once (one 1024-thread sweep workgroup per
// CU), so every workgroup a sweep waits on is running.
void assign_sweep_groups(std::vector<PatchDev> &pd, int budget) {
  std::vector<int> order(pd.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pd[x].nt > pd[y].nt; });
  for (int p : order) {
    if (pd[p].nt < 96) break;
    const int w = pd[p].prec == 64 ? 4 : 3;
    if (w > budget) break;
    pd[p].sweep_w = w;
    budget -= w;
  }
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Factorization step tables of the patches of one precision group.
std::vector<std::vector<StepMap>> step_tables(const mafem_batch &b, const std::vector<int32_t> &order,
                                              int prec) {
  struct Cls { int first, n, tb; };
  std::vector<Cls> cls;
  int64_t max_nt = 0;
  for (int i = 0; i < (int)order.size(); ++i) {
    const PatchDev &d = b.h_pd[order[i]];
    if (d.prec != prec) continue;
    if (cls.empty() || cls.back().tb != d.tb || cls.back().first + cls.back().n != i)
      cls.push_back({i, 0, d.tb});
    cls.back().n++;
    max_nt = std::max(max_nt, d.nt);
  }
  std::vector<std::vector<StepMap>> steps(max_nt);
  for (int64_t K = 0; K < max_nt; ++K) {
    StepMap sm{};
    int32_t blk = 0;
    for (const Cls &c : cls) {
      int cnt = 0;
      while (cnt < c.n && b.h_pd[order[c.first + cnt]].nt > K) ++cnt;
      if (cnt == 0) continue;
      if (sm.ncls == MAXC) {
        sm.blk0[sm.ncls] = blk;
        steps[K].push_back(sm);
        sm = StepMap{};
        blk = 0;
      }
      sm.first[sm.ncls] = c.first;
      sm.count[sm.ncls] = cnt;
      sm.tb[sm.ncls] = c.tb;
      sm.blk0[sm.ncls] = blk;
      blk += cnt;
      sm.ncls++;
    }
    if (sm.ncls > 0) {
      sm.blk0[sm.ncls] = blk;
      steps[K].push_back(sm);
    }
  }
  return steps;
}

// Builds descriptors, segment maps and factorization step tables (the flags
// are allocated once per batch and keep their contents).
void build_plan(mafem_batch &b) {
  free_plan(b);
  BatchPlan &B = b.B;
  const int P = (int)b.members.size();
  B.P = P;
  B.m = b.opts.restart;
  B.refactor_iters = b.opts.refactor_iters;
  B.escalate_iters = b.opts.escalate_iters;
  B.rtol = b.opts.rtol;
  b.h_pd.resize(P);
  B.has32 = B.has64 = false;
  for (int p = 0; p < P; ++p) {
    Ctx &c = b.members[p]->c;
    band_alloc(c, b.prec[p]);
    b.h_pd[p] = describe(c);
    b.h_pd[p].V = balloc<double>(b, (B.m + 1) * c.nt * (int64_t)TILE);
    (b.prec[p] == 32 ? B.has32 : B.has64) = true;
  }
  int dev_cus = 0;
  MAFEM_HIP(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, b.dev));
  int mw_budget = (3 * dev_cus) / 4;
  if (const char *e = std::getenv("MAFEM_BATCH_SWEEP_BUDGET")) mw_budget = std::atoi(e);
  assign_sweep_groups(b.h_pd, mw_budget);
  B.d_pd = bupload(b, b.h_pd);

  // segment maps
  B.motions.clear();
  for (const PatchDev &d : b.h_pd)
    if (std::find(B.motions.begin(), B.motions.end(), d.motion) == B.motions.end())
      B.motions.push_back(d.motion);
  B.area.clear();
  B.edge.clear();
  for (int mo : B.motions) {
    std::vector<int64_t> ba(P, 0), be(P, 0);
    for (int p = 0; p < P; ++p) {
      const PatchDev &d = b.h_pd[p];
      if (d.motion != mo) continue;
      ba[p] = cdiv(d.numel * d.ne * area_parts(d.motion), 256);
      be[p] = cdiv(d.n_edges * d.ne, 256);
    }
    B.area.push_back(make_seg(b, ba, 256));
    B.edge.push_back(make_seg(b, be, 256));
  }
  std::vector<int64_t> bb(P), bv(P), bn(P), bz(P), bt(P), bf(P);
  for (int p = 0; p < P; ++p) {
    const PatchDev &d = b.h_pd[p];
    bb[p] = cdiv(d.nnzb * d.ndf * d.ndf, 256);
    bv[p] = cdiv(d.nt * TILE, 256);
    bn[p] = cdiv(d.numnp, 256);
    bz[p] = cdiv(d.nt * (2 * d.tb + 1) * (int64_t)TILE * TILE * (d.prec / 8), 65536);
    bt[p] = d.nt;
    bf[p] = d.nt * (2 * d.tb + 1);
  }
  B.bsr = make_seg(b, bb, 256);
  B.vec = make_seg(b, bv, 256);
  B.node = make_seg(b, bn, 256);
  B.zero = make_seg(b, bz, 256);
  B.tiles = make_seg(b, bt, 256);  // tile inversion (64), diagonal products (256)
  B.fold = make_seg(b, bf, 256);
  std::vector<int64_t> bw(P);
  for (int p = 0; p < P; ++p) bw[p] = b.h_pd[p].sweep_w > 1 ? b.h_pd[p].sweep_w : 0;
  B.mw = make_seg(b, bw, 1024);  // multi-workgroup sweeps

  // factorization order: by precision, then classes of equal tb, each by
  // decreasing nt
  std::vector<int32_t> order(P);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
    const PatchDev &a = b.h_pd[x], &c = b.h_pd[y];
    if (a.prec != c.prec) return a.prec < c.prec;
    return a.tb != c.tb ? a.tb < c.tb : a.nt > c.nt;
  });
  B.d_order = bupload(b, order);
  B.steps32 = step_tables(b, order, 32);
  B.steps64 = step_tables(b, order, 64);
  B.max_nt = 0;
  for (const PatchDev &d : b.h_pd) B.max_nt = std::max(B.max_nt, d.nt);
  B.f.part = balloc<double>(b, std::max<int64_t>(B.vec.nblk, 1));
  B.f.part2 = balloc<double>(b, std::max<int64_t>((int64_t)B.vec.nblk * (B.m + 1), 1));
}

// Per-patch flags and Arnoldi state, once per batch.
void alloc_flags(mafem_batch &b) {
  BatchFlags &f = b.B.f;
  const int P = (int)b.members.size(), m = b.opts.restart;
  f.on = falloc<uint8_t>(b, P);
  f.sol = falloc<uint8_t>(b, P);
  f.fac = falloc<uint8_t>(b, P);
  f.kon = falloc<uint8_t>(b, P);
  f.fresh = falloc<uint8_t>(b, P);
  f.esc = falloc<uint8_t>(b, P);
  f.pert = falloc<uint8_t>(b, P);
  f.iters = falloc<int32_t>(b, P);
  f.kits = falloc<int32_t>(b, P);
  f.kdim = falloc<int32_t>(b, P);
  f.nfac = falloc<int32_t>(b, P);
  f.kits_tot = falloc<int32_t>(b, P);
  f.time = falloc<double>(b, P);
  f.dt = falloc<double>(b, P);
  f.bn = falloc<double>(b, P);
  f.relres = falloc<double>(b, P);
  f.relres0 = falloc<double>(b, P);
  f.eps = falloc<double>(b, P);
  f.scale = falloc<double>(b, P);
  f.H = falloc<double>(b, (int64_t)P * m * (m + 1));
  f.cs = falloc<double>(b, (int64_t)P * m);
  f.sn = falloc<double>(b, (int64_t)P * m);
  f.g = falloc<double>(b, (int64_t)P * (m + 1));
  f.hc = falloc<double>(b, (int64_t)P * (m + 1));
  f.eps_hist = nullptr;
  f.count = falloc<int32_t>(b, 8);
  const std::vector<uint8_t> zeros8(P, 0), ones8(P, 1);
  const std::vector<int32_t> zeros32(std::max(P, 8), 0);
  for (uint8_t *a : {f.on, f.sol, f.kon, f.fresh, f.esc, f.pert})
    MAFEM_HIP(hipMemcpy(a, zeros8.data(), P, hipMemcpyHostToDevice));
  MAFEM_HIP(hipMemcpy(f.fac, ones8.data(), P, hipMemcpyHostToDevice));  // no factor yet
  for (int32_t *a : {f.iters, f.kits, f.kdim, f.nfac, f.kits_tot})
    MAFEM_HIP(hipMemcpy(a, zeros32.data(), 4 * P, hipMemcpyHostToDevice));
  MAFEM_HIP(hipMemcpy(f.count, zeros32.data(), 32, hipMemcpyHostToDevice));
}

// eps_hist depends on max_iter: (re)allocated when it changes
void ensure_hist(mafem_batch &b, int max_iter) {
  BatchPlan &B = b.B;
  if (B.f.eps_hist && B.max_iter == max_iter) return;
  destroy_graphs(b);  // the update graph holds the old pointer
  B.max_iter = max_iter;
  B.f.eps_hist = falloc<double>(b, (int64_t)B.P * max_iter);
}

// The plan is stale when a member's band storage moved (a precision change
// through its own solver entry points).
bool plan_stale(const mafem_batch &b) {
  for (size_t p = 0; p < b.members.size(); ++p) {
    const Ctx &c = b.members[p]->c;
    if (c.d_band != b.h_pd[p].band || c.band_prec != b.prec[p]) return true;
  }
  return false;
}

// Recording holds the capture mutex, so no context is being created on
// another host thread meanwhile (its synchronous null-stream copies would
// invalidate the capture).
template <class F>
hipGraphExec_t record(mafem_batch &b, F &&issue) {
  std::lock_guard<std::mutex> no_create(capture_mutex());
  hipGraph_t g = nullptr;
  MAFEM_HIP(hipStreamBeginCapture(b.B.stream, hipStreamCaptureModeRelaxed));
  try {
    issue();
  } catch (...) {
    (void)hipStreamEndCapture(b.B.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  MAFEM_HIP(hipStreamEndCapture(b.B.stream, &g));
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MAFEM_HIP(e);
  return ex;
}

void issue_asm(mafem_batch &b) {
  b_assemble(b.B);
  b_solve_begin(b.B);
}
void issue_fac(mafem_batch &b) {
  b_factor(b.B);
  b_mark_factored(b.B);
}
void issue_cyc(mafem_batch &b) { b_gmres_cycle(b.B); }
void issue_end(mafem_batch &b) { b_solve_end(b.B); }
void issue_upd(mafem_batch &b) { b_update(b.B); }

void run(mafem_batch &b, hipGraphExec_t &g, void (*issue)(mafem_batch &)) {
  if (!b.opts.use_graphs) {
    issue(b);
    return;
  }
  if (!g) g = record(b, [&] { issue(b); });
  MAFEM_HIP(hipGraphLaunch(g, b.B.stream));
}

// Patches whose fresh FP32 factorization could not precondition the solve
// move to FP64 band storage: the plan is rebuilt around them (flags and
// iterates are kept) and they are refactored before the solve continues.
void escalate(mafem_batch &b) {
  const int P = b.B.P;
  std::vector<uint8_t> esc(P);
  MAFEM_HIP(hipMemcpy(esc.data(), b.B.f.esc, P, hipMemcpyDeviceToHost));
  for (int p = 0; p < P; ++p)
    if (esc[p]) b.prec[p] = 64;
  const std::vector<uint8_t> zeros8(P, 0);
  MAFEM_HIP(hipMemcpy(b.B.f.esc, zeros8.data(), P, hipMemcpyHostToDevice));
  build_plan(b);
}

void read_counts(mafem_batch &b) {
  MAFEM_HIP(hipMemcpyAsync(b.h_count, b.B.f.count, 32, hipMemcpyDeviceToHost, b.B.stream));
  MAFEM_HIP(hipStreamSynchronize(b.B.stream));
}

float ms_between(hipEvent_t a, hipEvent_t c) {
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, c);
  return ms;
}

}  // namespace

#define BATCH_GUARD(bp, ...)                                   \
  do {                                                         \
    if (!(bp)) return MAFEM_EINVAL;                            \
    mafem_batch &b = *(bp);                                    \
    try {                                                      \
      __VA_ARGS__                                              \
    } catch (const HipError &h) {                              \
      b.err = std::string("HIP error ") + hipGetErrorString(h.code) + " in " + h.what + " (" + \
              h.file + ":" + std::to_string(h.line) + ")";     \
      return MAFEM_EHIP;                                       \
    } catch (const BatchError &e) {                            \
      b.err = e.msg;                                           \
      return e.code;                                           \
    } catch (const std::bad_alloc &) {                         \
      b.err = "host allocation failed";                        \
      return MAFEM_ENOMEM;                                     \
    }                                                          \
  } while (0)

static thread_local std::string g_batch_err;

extern "C" {

int mafem_device_memory(size_t *free_bytes, size_t *total_bytes) {
  if (!free_bytes || !total_bytes) return MAFEM_EINVAL;
  return hipMemGetInfo(free_bytes, total_bytes) == hipSuccess ? MAFEM_OK : MAFEM_EHIP;
}

int mafem_batch_create(mafem_batch **out, int32_t n, mafem_ctx *const *members,
                       const mafem_batch_opts *opts) {
  if (!out) return MAFEM_EINVAL;
  *out = nullptr;
  if (n < 1 || !members) {
    g_batch_err = "a batch needs at least one member";
    return MAFEM_EINVAL;
  }
  mafem_batch_opts o{32, 20, 1, 6, 1, 12, 1e-10};
  if (opts) o = *opts;
  if ((o.factor_precision != 32 && o.factor_precision != 64) || o.restart < 1 ||
      o.restart > 64 || o.max_cycles < 1 || o.refactor_iters < -1 || o.escalate_iters < 1 ||
      !(o.rtol > 0.0)) {
    g_batch_err = "invalid batch options";
    return MAFEM_EINVAL;
  }
  mafem_batch *bp = new (std::nothrow) mafem_batch();
  if (!bp) {
    g_batch_err = "host allocation failed";
    return MAFEM_ENOMEM;
  }
  bp->opts = o;
  int rc = MAFEM_OK;
  {
    mafem_batch &b = *bp;
    try {
      MAFEM_HIP(hipGetDevice(&b.dev));
      for (int32_t i = 0; i < n; ++i) {
        if (!members[i]) throw BatchError{"null member", MAFEM_EINVAL};
        if (members[i]->c.dev != b.dev) throw BatchError{"member on another device", MAFEM_EINVAL};
        for (int32_t j = 0; j < i; ++j)
          if (members[j] == members[i]) throw BatchError{"member listed twice", MAFEM_EINVAL};
        b.members.push_back(members[i]);
      }
      MAFEM_HIP(hipStreamCreateWithFlags(&b.B.stream, hipStreamNonBlocking));
      for (auto &e : b.ev) MAFEM_HIP(hipEventCreate(&e));
      MAFEM_HIP(hipHostMalloc((void **)&b.h_count, 32, hipHostMallocDefault));
      // a member keeps the precision a previous batch chose for it and, when
      // still valid, the factorization that batch left in its band
      b.prec.resize(n);
      for (int32_t i = 0; i < n; ++i) {
        const Ctx &c = members[i]->c;
        b.prec[i] = c.batch_prec == 64 ? 64 : o.factor_precision;
      }
      alloc_flags(b);
      build_plan(b);
      std::vector<uint8_t> fac(n);
      for (int32_t i = 0; i < n; ++i) {
        const Ctx &c = members[i]->c;
        fac[i] = (c.batch_factor_prec == b.prec[i] && c.band_prec == b.prec[i]) ? 0 : 1;
      }
      MAFEM_HIP(hipMemcpy(b.B.f.fac, fac.data(), n, hipMemcpyHostToDevice));
    } catch (const HipError &h) {
      b.err = std::string("HIP error ") + hipGetErrorString(h.code) + " in " + h.what;
      rc = MAFEM_EHIP;
    } catch (const BatchError &e) {
      b.err = e.msg;
      rc = e.code;
    } catch (const std::bad_alloc &) {
      b.err = "host allocation failed";
      rc = MAFEM_ENOMEM;
    }
  }
  if (rc != MAFEM_OK) {
    g_batch_err = bp->err;
    mafem_batch_destroy(bp);
    return rc;
  }
  *out = bp;
  return MAFEM_OK;
}

void mafem_batch_destroy(mafem_batch *bp) {
  if (!bp) return;
  (void)hipSetDevice(bp->dev);
  if (bp->B.stream) (void)hipStreamSynchronize(bp->B.stream);
  free_plan(*bp);
  for (void *p : bp->fowned) (void)hipFree(p);
  for (auto &e : bp->ev)
    if (e) (void)hipEventDestroy(e);
  if (bp->h_count) (void)hipHostFree(bp->h_count);
  if (bp->B.stream) (void)hipStreamDestroy(bp->B.stream);
  delete bp;
}

const char *mafem_batch_last_error(const mafem_batch *b) {
  return b ? b->err.c_str() : g_batch_err.c_str();
}

int mafem_batch_time_step(mafem_batch *bp, const double *time, const double *dt,
                          const uint8_t *active, int32_t predict, int32_t max_iter, int32_t *iters,
                          double *eps_hist, int32_t *status) {
  BATCH_GUARD(bp, {
    if (!time || !dt) {
      b.err = "null time or dt";
      return MAFEM_EINVAL;
    }
    if (max_iter < 1 || max_iter > 1000) {
      b.err = "max_iter must be in 1..1000";
      return MAFEM_EINVAL;
    }
    MAFEM_HIP(hipSetDevice(b.dev));
    BatchPlan &B = b.B;
    const int P = B.P;
    for (mafem_ctx *m : b.members) {
      MAFEM_HIP(hipStreamSynchronize(m->c.stream));
      // the batch factors into the member's band in its own (one-chain) layout:
      // a factorization the context kept for its own Krylov solves is gone
      m->c.have_factor = false;
    }
    if (plan_stale(b)) build_plan(b);
    ensure_hist(b, max_iter);
    BatchFlags &f = B.f;
    std::vector<uint8_t> ones8(P, 1);
    if (active)
      for (int p = 0; p < P; ++p) ones8[p] = active[p] ? 1 : 0;
    const std::vector<int32_t> zeros32(P, 0);
    MAFEM_HIP(hipMemcpyAsync(f.time, time, sizeof(double) * P, hipMemcpyHostToDevice, B.stream));
    MAFEM_HIP(hipMemcpyAsync(f.dt, dt, sizeof(double) * P, hipMemcpyHostToDevice, B.stream));
    MAFEM_HIP(hipMemcpyAsync(f.on, ones8.data(), P, hipMemcpyHostToDevice, B.stream));
    MAFEM_HIP(hipMemcpyAsync(f.iters, zeros32.data(), 4 * P, hipMemcpyHostToDevice, B.stream));
    std::vector<int32_t> nfac0(P);
    MAFEM_HIP(hipMemcpyAsync(nfac0.data(), f.nfac, 4 * P, hipMemcpyDeviceToHost, B.stream));
    b.h_count[3] = 0;
    MAFEM_HIP(hipMemcpyAsync(f.count + 3, b.h_count + 3, 4, hipMemcpyHostToDevice, B.stream));
    b_count(B);
    read_counts(b);
    int32_t pend = b.h_count[2];

    mafem_batch_stats st{};
    std::vector<int32_t> kits0(P);
    MAFEM_HIP(hipMemcpy(kits0.data(), f.kits_tot, 4 * P, hipMemcpyDeviceToHost));
    MAFEM_HIP(hipEventRecord(b.ev[0], B.stream));
    if (predict) b_predict(B);
    float t_asm = 0, t_fac = 0, t_sol = 0, t_upd = 0;
    int it = 0;
    for (; it < max_iter; ++it) {
      MAFEM_HIP(hipEventRecord(b.ev[1], B.stream));
      run(b, b.g_asm, issue_asm);
      MAFEM_HIP(hipEventRecord(b.ev[2], B.stream));
      // a failed solve moves a patch one rung up: stale factorization ->
      // fresh one -> FP64 storage -> static pivoting (further cycles on the
      // same factors, up to an Arnoldi-step budget) -> hard failure; the
      // bound here only guarantees the loop ends
      for (int attempt = 0; attempt < 32; ++attempt) {
        if (pend > 0) {
          MAFEM_HIP(hipEventRecord(b.ev[3], B.stream));
          run(b, b.g_fac, issue_fac);
          MAFEM_HIP(hipEventRecord(b.ev[4], B.stream));
          ++st.factor_passes;
          MAFEM_HIP(hipEventSynchronize(b.ev[4]));
          t_fac += ms_between(b.ev[3], b.ev[4]);
          pend = 0;
        }
        for (int cyc = 0; cyc < b.opts.max_cycles; ++cyc) {
          MAFEM_HIP(hipEventRecord(b.ev[5], B.stream));
          run(b, b.g_cyc, issue_cyc);
          MAFEM_HIP(hipEventRecord(b.ev[6], B.stream));
          ++st.gmres_cycles;
          read_counts(b);
          t_sol += ms_between(b.ev[5], b.ev[6]);
          if (b.h_count[1] == 0) break;
        }
        if (b.h_count[1] == 0) break;
        run(b, b.g_end, issue_end);
        read_counts(b);
        pend = b.h_count[2];
        if (b.h_count[4] > 0) escalate(b);
        if (b.h_count[1] == 0) break;
      }
      MAFEM_HIP(hipEventRecord(b.ev[3], B.stream));
      run(b, b.g_upd, issue_upd);
      MAFEM_HIP(hipEventRecord(b.ev[4], B.stream));
      read_counts(b);
      t_asm += ms_between(b.ev[1], b.ev[2]);
      t_upd += ms_between(b.ev[3], b.ev[4]);
      if (b.h_count[4] > 0) {
        escalate(b);
        b_count(B);
        read_counts(b);
      }
      pend = b.h_count[2];
      if (b.h_count[0] == 0) {
        ++it;
        break;
      }
    }
    MAFEM_HIP(hipEventRecord(b.ev[5], B.stream));
    MAFEM_HIP(hipEventSynchronize(b.ev[5]));
    st.total_us = 1e3 * ms_between(b.ev[0], b.ev[5]);
    st.assemble_us = 1e3 * t_asm;
    st.factor_us = 1e3 * t_fac;
    st.solve_us = 1e3 * t_sol;
    st.update_us = 1e3 * t_upd;

    std::vector<int32_t> h_it(P), h_nf(P), h_kt(P);
    MAFEM_HIP(hipMemcpy(h_kt.data(), f.kits_tot, 4 * P, hipMemcpyDeviceToHost));
    std::vector<double> h_eps(P), h_hist((size_t)P * max_iter);
    MAFEM_HIP(hipMemcpy(h_it.data(), f.iters, 4 * P, hipMemcpyDeviceToHost));
    MAFEM_HIP(hipMemcpy(h_nf.data(), f.nfac, 4 * P, hipMemcpyDeviceToHost));
    MAFEM_HIP(hipMemcpy(h_eps.data(), f.eps, sizeof(double) * P, hipMemcpyDeviceToHost));
    if (eps_hist)
      MAFEM_HIP(hipMemcpy(h_hist.data(), f.eps_hist, sizeof(double) * P * max_iter,
                          hipMemcpyDeviceToHost));
    read_counts(b);
    st.hard_failures = b.h_count[3];
    for (int p = 0; p < P; ++p) st.fp64_patches += b.prec[p] == 64 ? 1 : 0;
    int any_fail = 0;
    for (int p = 0; p < P; ++p) {
      const bool sel = ones8[p] != 0;
      const bool ok = h_eps[p] < b.h_pd[p].enr && h_it[p] > 0;
      if (iters) iters[p] = sel ? h_it[p] : 0;
      if (status) status[p] = !sel ? 2 : (ok ? 0 : 1);
      if (!sel) {
        if (eps_hist)
          for (int k = 0; k < max_iter; ++k) eps_hist[(size_t)p * max_iter + k] = 0.0;
        continue;
      }
      if (!ok) any_fail = 1;
      st.newton_iters = std::max(st.newton_iters, h_it[p]);
      st.patch_iters += h_it[p];
      st.patch_factors += h_nf[p] - nfac0[p];
      st.gmres_iters += h_kt[p] - kits0[p];
      st.elements += b.h_pd[p].numel;
      if (eps_hist)
        for (int k = 0; k < max_iter; ++k)
          eps_hist[(size_t)p * max_iter + k] = k < h_it[p] ? h_hist[(size_t)p * max_iter + k] : 0.0;
    }
    // multi-workgroup sweeps whose waits ran out invalidate the step
    for (int p = 0; p < P; ++p) {
      if (b.h_pd[p].sweep_w <= 1) continue;
      int e = 0;
      MAFEM_HIP(hipMemcpy(&e, b.h_pd[p].err, sizeof(int), hipMemcpyDeviceToHost));
      if (e != 0) {
        MAFEM_HIP(hipMemset(b.h_pd[p].err, 0, sizeof(int)));
        throw BatchError{"a triangular sweep wait ran out: this step's corrections are invalid "
                         "and so are the member states they updated (set them again "
                         "before continuing)", MAFEM_EHIP};
      }
    }
    // what each member's band holds now, for a later batch
    std::vector<uint8_t> h_fac(P);
    MAFEM_HIP(hipMemcpy(h_fac.data(), f.fac, P, hipMemcpyDeviceToHost));
    for (int p = 0; p < P; ++p) {
      Ctx &c = b.members[p]->c;
      c.batch_prec = b.prec[p];
      c.batch_factor_prec = (h_fac[p] == 0 && h_nf[p] > 0) || c.batch_factor_prec == b.prec[p]
                                ? (h_fac[p] == 0 ? b.prec[p] : 0)
                                : 0;
    }
    b.stats = st;
    if (any_fail) {
      b.err = "did not reach Newton--Raphson tolerance";
      return MAFEM_NOT_CONVERGED;
    }
    return MAFEM_OK;
  });
}

int mafem_batch_pull_force(mafem_batch *bp, double *out) {
  BATCH_GUARD(bp, {
    if (!out) {
      b.err = "null output";
      return MAFEM_EINVAL;
    }
    const int P = (int)b.members.size();
    for (int p = 0; p < P; ++p)
      if (!b.members[p]->c.pf_err.empty() || b.members[p]->c.motion == STATIC) {
        b.err = "member " + std::to_string(p) + ": " +
                (b.members[p]->c.pf_err.empty() ? std::string("f_pull needs 3-D velocity")
                                                : b.members[p]->c.pf_err);
        return MAFEM_EINVAL;
      }
    MAFEM_HIP(hipSetDevice(b.dev));
    // on the batch stream, after the members' last batched step
    for (int p = 0; p < P; ++p) {
      Ctx &c = b.members[p]->c;
      ensure_pull(c);
      launch_pull_force(c, b.B.stream, c.d_pf_el, c.d_pf_slot, c.d_pf_rel, c.d_pf_out);
      MAFEM_HIP(hipMemcpyAsync(c.h_red, c.d_pf_out, sizeof(double) * 6, hipMemcpyDeviceToHost,
                               b.B.stream));
    }
    MAFEM_HIP(hipStreamSynchronize(b.B.stream));
    for (int p = 0; p < P; ++p)
      for (int k = 0; k < 6; ++k) out[6 * p + k] = b.members[p]->c.h_red[k];
    return MAFEM_OK;
  });
}

int mafem_batch_get_stats(const mafem_batch *b, mafem_batch_stats *out) {
  if (!b || !out) return MAFEM_EINVAL;
  *out = b->stats;
  return MAFEM_OK;
}

}  // extern "C"
