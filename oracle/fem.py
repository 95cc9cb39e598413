"""Oracle restatement of the reference's assembly, solve and Newton loop.

Follows (reference paths):
  src/analysis/FiniteElement.jl:75-200  calc_r_K   (per-element residual and
      complex-step tangent come from csrc/oracle.c; the scatter through LM over
      active DOFs, then the serial Neumann edge loop, are restated here)
  src/analysis/FiniteElement.jl:11-63   time_step  (Newton, <= 14 iterations,
      eps = |du|_2 / nmdf, AssertionError on failure)
  src/analysis/FiniteElement.jl:408-423 update_xms
  src/Analysis.jl:17-102                run_analysis (predictor, then Newton)
  src/analysis/PullForce.jl:8-80        get_adj_maps, calc_pull_force
  src/Output.jl:30-118                  generate_output

The reference solves with UMFPACK (SuiteSparse through SparseArrays `\\`); this
oracle uses scipy's SuperLU, so the solution agrees to rounding only: the
linear-solve rounding is "parity unpinned" against UMFPACK itself.

Pinned by: the equilibrium invariant (flat F_PULL state at t = 0 with zero
pull speed gives r = 0), central finite differences of the residual against
the tangent, quadratic Newton convergence, the closed-form GeoDynStress
tests of the reference, flat-patch closed forms of the mesh-motion rows
(ALEV/ALEVB rm from a linear v^m field and a constant p^m, rp from a normal
v^m - v, EUL rm = am (v^m - (n.v) n) N J w, the tension row's surface
divergence) and the ALEV bending steady state converging to the analytic arc
under refinement (tests/test_oracle_*.py).  The pull force is pinned by
its closed form on the flat tension state (f = (kb/4) L x the integral of the
pulled element's edge basis functions for an edge element, 0 for an interior
one), by the sum of every element's residual over the pulled nodes, and by
the force curve of the reference's documented tether example
(docs/src/index.md:142-176, example-pull.png).
"""
import ctypes
import os

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from . import mesh as om

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcMesh(ctypes.Structure):
    _fields_ = [("numel", ctypes.c_int64), ("numnp", ctypes.c_int64),
                ("ndf", ctypes.c_int64),
                ("IX", ctypes.POINTER(ctypes.c_int64)),
                ("ID", ctypes.POINTER(ctypes.c_int64)),
                ("v_order", ctypes.c_int32 * 3), ("m_order", ctypes.c_int32 * 3),
                ("lam_order", ctypes.c_int32), ("p_order", ctypes.c_int32),
                ("mm_order", ctypes.c_int32 * 3),
                ("motion", ctypes.c_int32), ("scenario", ctypes.c_int32),
                ("kb", ctypes.c_double), ("kg", ctypes.c_double),
                ("zv", ctypes.c_double), ("pn", ctypes.c_double),
                ("adb", ctypes.c_double), ("am", ctypes.c_double),
                ("ek", ctypes.c_double)]


def lib():
    """Load oracle/liboracle.so (built by `make -C oracle`)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int64)
        L.orc_area_elements.argtypes = [ctypes.POINTER(OrcMesh), ip, dp, dp, dp,
                                        ctypes.c_double, ctypes.c_int64,
                                        ctypes.c_int64, dp, dp]
        L.orc_area_elements.restype = ctypes.c_int
        L.orc_bdry_elements.argtypes = [ctypes.POINTER(OrcMesh), ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int64, ip, ip, dp, dp, dp,
                                        ctypes.c_double, dp, dp]
        L.orc_bdry_elements.restype = ctypes.c_int
        L.orc_gds_point.argtypes = [ctypes.POINTER(OrcMesh), dp, dp, dp, dp]
        L.orc_gds_point.restype = None
        L.orc_area_residual.argtypes = [ctypes.POINTER(OrcMesh), ip, dp,
                                        ctypes.c_int64, dp, dp, dp]
        L.orc_area_residual.restype = None
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


class OracleMesh:
    """Keeps the ctypes view of a mesh (and the arrays it points to) alive."""

    def __init__(self, mesh, p):
        self.mesh, self.p = mesh, p
        self.IX = np.asfortranarray(mesh.IX, dtype=np.int64)
        self.ID = np.asfortranarray(mesh.ID, dtype=np.int64)
        self.uel = np.asarray(mesh.area_uel_ids, dtype=np.int64)
        self.ufns = np.ascontiguousarray(mesh.area_ufns, dtype=np.float64)
        s = OrcMesh()
        s.numel, s.numnp, s.ndf = mesh.numel, mesh.numnp, mesh.ndf
        s.IX = _ip(self.IX.ravel(order="F")) if False else self.IX.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int64))
        s.ID = self.ID.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        s.v_order[:] = om.v_order(mesh.dofs)
        s.m_order[:] = om.m_order(mesh.dofs)
        s.lam_order = om.lam_order(mesh.dofs)
        s.p_order = om.p_order(mesh.dofs)
        s.mm_order[:] = om.m_motion_order(p.motion, mesh.dofs)
        s.motion, s.scenario = p.motion, p.scenario
        s.kb, s.kg, s.zv, s.pn = p.kb, p.kg, p.zv, p.pn
        s.adb, s.am, s.ek = p.adb, p.am, p.ek
        self.s = s
        self.bdry = {}
        for b, (uids, ufns) in mesh.bdry_fns.items():
            els = np.asarray(mesh.bdry_elems[b], dtype=np.int64) - 1
            buel = np.asarray(uids, dtype=np.int64)
            self.bdry[b] = (els, buel, np.ascontiguousarray(ufns, dtype=np.float64))


def area_elements(om_, xms, cps, dt, e0=0, e1=None):
    """Per-element (r_el, K_el) of area elements [e0, e1); K_el[e, row, col]."""
    m = om_.mesh
    e1 = m.numel if e1 is None else e1
    ne = 9 * m.ndf
    X = np.asfortranarray(xms, dtype=np.float64)
    C = np.asfortranarray(cps, dtype=np.float64)
    r = np.zeros((e1 - e0, ne))
    K = np.zeros((e1 - e0, ne, ne))
    rc = lib().orc_area_elements(ctypes.byref(om_.s), _ip(om_.uel), _dp(om_.ufns),
                                 _dp(X), _dp(C), dt, e0, e1, _dp(r), _dp(K))
    assert rc == 0
    return r, K.transpose(0, 2, 1)


def bdry_elements(om_, bdry, ntype, nval, xms, cps, time, dt, bend_tm=1.0):
    """Per-element (r_el, K_el) of the edge elements of one boundary."""
    m = om_.mesh
    ne = 9 * m.ndf
    els, buel, ufns = om_.bdry[bdry]
    X = np.asfortranarray(xms, dtype=np.float64)
    C = np.asfortranarray(cps, dtype=np.float64)
    r = np.zeros((len(els), ne))
    K = np.zeros((len(els), ne, ne))
    rc = lib().orc_bdry_elements(ctypes.byref(om_.s), bdry, ntype, nval, time,
                                 bend_tm, len(els), _ip(els), _ip(buel), _dp(ufns),
                                 _dp(X), _dp(C), dt, _dp(r), _dp(K))
    if rc != 0:
        raise AssertionError("Neumann boundary condition not implemented")
    return r, K.transpose(0, 2, 1)


def _scatter(LMcols, r_el, K_el, nmdf):
    """Sum element blocks into (r, COO) over active DOFs (LM != 0)."""
    r = np.zeros(nmdf)
    rows, cols, vals = [], [], []
    for k in range(LMcols.shape[1]):
        lm = LMcols[:, k]
        act = np.nonzero(lm)[0]
        g = lm[act] - 1
        np.add.at(r, g, r_el[k, act])
        blk = K_el[k][np.ix_(act, act)]
        rows.append(np.repeat(g, len(g)))
        cols.append(np.tile(g, len(g)))
        vals.append(blk.ravel())
    return r, rows, cols, vals


def calc_r_K(om_, xms, cps, time, dt, bend_tm=1.0, chunk=512):
    """calc_r_K (FiniteElement.jl:75-200): returns (r_gl, K_gl as CSC)."""
    m = om_.mesh
    n = m.nmdf
    r_gl = np.zeros(n)
    K_gl = sps.csc_matrix((n, n))
    for e0 in range(0, m.numel, chunk):
        e1 = min(m.numel, e0 + chunk)
        r_el, K_el = area_elements(om_, xms, cps, dt, e0, e1)
        r, rows, cols, vals = _scatter(m.LM[:, e0:e1], r_el, K_el, n)
        r_gl += r
        K_gl = K_gl + sps.csc_matrix(
            (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
            shape=(n, n))
    for (bdry, ntype, nval) in m.inh_neu_bcs:
        r_el, K_el = bdry_elements(om_, bdry, ntype, nval, xms, cps, time, dt, bend_tm)
        els = np.asarray(m.bdry_elems[bdry]) - 1
        r, rows, cols, vals = _scatter(m.LM[:, els], r_el, K_el, n)
        r_gl += r
        K_gl = K_gl + sps.csc_matrix(
            (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
            shape=(n, n))
    K_gl.sum_duplicates()
    return r_gl, K_gl


def update_xms(motion, xms, u, dt, dofs):
    """update_xms! (FiniteElement.jl:408-423), in place."""
    mo = om.m_motion_order(motion, dofs)
    for j in range(3):
        if mo[j] != 0:
            xms[:, j] += dt * u[:, mo[j] - 1]


def apply_du(mesh, p, xms, cps, du, dt):
    """Δcps[ID_inv] .= Δu; cps .+= Δcps; update_xms! (FiniteElement.jl:41-46)."""
    dc = np.zeros_like(cps)
    nodes = np.fromiter((t[0] - 1 for t in mesh.ID_inv), dtype=np.int64)
    dofs_ = np.fromiter((t[1] - 1 for t in mesh.ID_inv), dtype=np.int64)
    dc[nodes, dofs_] = du
    cps += dc
    update_xms(p.motion, xms, dc, dt, mesh.dofs)


def time_step(om_, xms, cps, time, dt, max_iter=14, bend_tm=1.0):
    """time_step! (FiniteElement.jl:11-63); mutates xms, cps; returns eps list."""
    m, p = om_.mesh, om_.p
    eps = []
    it = 1
    while it < max_iter + 1:
        r, K = calc_r_K(om_, xms, cps, time, dt, bend_tm)
        du = -spla.spsolve(K.tocsc(), r)
        apply_du(m, p, xms, cps, du, dt)
        eps.append(np.linalg.norm(du) / m.nmdf)
        it += 1
        if eps[-1] < p.enr:
            break
    if not eps[-1] < p.enr:
        raise AssertionError("did not reach Newton--Raphson tolerance")
    return eps


def run_analysis(om_, xms, cps, dts, t0=0.0, bend_tm=1.0, max_iter=14):
    """run_analysis! (Analysis.jl:17-102) without file output.

    Returns the per-step eps histories.
    """
    m, p = om_.mesh, om_.p
    times = np.cumsum(dts) + t0
    hist = []
    for k, dt in enumerate(dts):
        update_xms(p.motion, xms, cps, dt, m.dofs)
        hist.append(time_step(om_, xms, cps, float(times[k]), float(dt),
                              max_iter=max_iter, bend_tm=bend_tm))
    return hist


def get_adj_maps(num1el, numel, IX, poly=2):
    """get_adj_maps (PullForce.jl:25-53): the (2 poly + 1)^2 element ids
    around the pulled one, by id arithmetic (1-based), and per element the
    pair (rows of rv_pull, rows of rv_el) of the pulled element's nodes it
    shares: local indices in increasing order on both sides, 3 rows each."""
    pull = om.get_pull_el_id(numel)
    pull_nodes = IX[:, pull - 1]
    ids = [pull + i * num1el + j for i in range(-poly, poly + 1)
           for j in range(-poly, poly + 1)]
    maps = []
    for e in ids:
        adj_nodes = IX[:, e - 1]
        rows_pull = [3 * a + c for a in range(9) if pull_nodes[a] in adj_nodes
                     for c in range(3)]
        rows_adj = [3 * b + c for b in range(9) if adj_nodes[b] in pull_nodes
                    for c in range(3)]
        maps.append((rows_pull, rows_adj))
    return ids, maps


def elem_rv(om_, el, xms, cps):
    """rv of calc_elem_dof_residuals (FiniteElement.jl:253-330) for one area
    element (0-based), layout comp + 3 (node - 1), from the oracle residual."""
    m = om_.mesh
    X = np.asfortranarray(xms, dtype=np.float64)
    C = np.asfortranarray(cps, dtype=np.float64)
    r = np.zeros(9 * m.ndf)
    lib().orc_area_residual(ctypes.byref(om_.s), _ip(om_.uel), _dp(om_.ufns), int(el),
                            _dp(X), _dp(C), _dp(r))
    vo = om.v_order(m.dofs)
    return np.array([r[(vo[c] - 1) + m.ndf * a] for a in range(9) for c in range(3)])


def calc_pull_force(om_, xms, cps, adj_el_ids, adj_node_map):
    """calc_pull_force (PullForce.jl:61-80): the 3-vector sum over the pulled
    nodes of the velocity residual rows of the adjacent elements."""
    rv_pull = np.zeros(27)
    for k, e in enumerate(adj_el_ids):
        rv = elem_rv(om_, e - 1, xms, cps)
        rp, ra = adj_node_map[k]
        rv_pull[rp] += rv[ra]
    return rv_pull.reshape(9, 3).sum(axis=0)


def pull_position(mesh, xms):
    """Mean position of the pulled element's 9 nodes (Analysis.jl:82-84)."""
    nodes = mesh.IX[:, om.get_pull_el_id(mesh.numel) - 1] - 1
    return xms[nodes, :].sum(axis=0) / 9.0


def generate_output(mesh, xms, cps):
    """generate_output (Output.jl:30-118): state sampled at the area Gauss
    points, the boundary Gauss points (BOTTOM, RIGHT, TOP, LEFT rows and
    columns, in boundary-element order) and the corners."""
    z1, z2 = mesh.num1el * 3 + 2, mesh.num2el * 3 + 2
    xout = np.zeros((z1, z2, 3))
    uout = np.zeros((z1, z2, mesh.ndf))
    for e2 in range(1, mesh.num2el + 1):
        for e1 in range(1, mesh.num1el + 1):
            el = e1 + mesh.num1el * (e2 - 1)
            nodes = mesh.IX[:, el - 1] - 1
            for g2 in range(1, 4):
                for g1 in range(1, 4):
                    gp = g1 + 3 * (g2 - 1)
                    N = mesh.area_ufns[mesh.area_uel_ids[el - 1] - 1, gp - 1, 1:10]
                    o1, o2 = (e1 - 1) * 3 + g1, (e2 - 1) * 3 + g2  # 0-based of +1
                    xout[o1, o2] = xms[nodes].T @ N
                    uout[o1, o2] = cps[nodes].T @ N
    for b in (om.BOTTOM, om.RIGHT, om.TOP, om.LEFT):
        uids, ufns = mesh.bdry_fns[b]
        out_id = 1
        for j, el in enumerate(mesh.bdry_elems[b]):
            nodes = mesh.IX[:, el - 1] - 1
            for g in range(3):
                N = ufns[uids[j] - 1, g, 1:10]
                o1, o2 = {om.BOTTOM: (out_id, 0), om.RIGHT: (z1 - 1, out_id),
                          om.TOP: (out_id, z2 - 1), om.LEFT: (0, out_id)}[b]
                xout[o1, o2] = xms[nodes].T @ N
                uout[o1, o2] = cps[nodes].T @ N
                out_id += 1
    for k, (o1, o2) in enumerate([(0, 0), (z1 - 1, 0), (0, z2 - 1), (z1 - 1, z2 - 1)]):
        nodes = mesh.IX[:, mesh.crnr_elems[k] - 1] - 1
        N = mesh.crnr_fns[k][1:10]
        xout[o1, o2] = xms[nodes].T @ N
        uout[o1, o2] = cps[nodes].T @ N
    return xout, uout


def build(p, **args):
    """Mesh + initial state for the parameters (Input.jl:22-121)."""
    mesh = om.make_mesh(p, **args)
    xms, cps = om.prepare_input(p, mesh, **args)
    return OracleMesh(mesh, p), xms, cps
