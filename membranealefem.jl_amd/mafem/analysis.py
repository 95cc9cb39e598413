"""Host mirror of the reference's analysis entry points on the HIP path.

`time_step`, `calc_r_K`, `update_xms` and `run_analysis` keep the reference's
names, arguments and error behaviour (src/analysis/FiniteElement.jl:11-200,
408-423; src/Analysis.jl:17-102); the work runs in libmafem on the GPU.
`DeviceMesh` is the device-resident context behind them (one per mesh).
"""
import ctypes
import os

import numpy as np
import scipy.sparse as sps

from . import _lib as L
from . import output as out
from .enums import BOTTOM, F_PULL, LEFT, RIGHT, TOP, Unknown
from .mesh import get_pull_el_id, m_motion_order, m_order, v_order
from .params import Params

_BDRY = (BOTTOM, RIGHT, TOP, LEFT)


class DeviceMesh:
    """A libmafem context: mesh tables and state resident on the GPU."""

    def __init__(self, mesh, p: Params, bend_tm=1.0):
        self.lib = L.load()
        self.mesh, self.p = mesh, p
        m = mesh
        keep = self._keep = []

        def a64(x):
            x = np.ascontiguousarray(np.asarray(x, dtype=np.int64).ravel(order="F"))
            keep.append(x)
            return L.ptr(x, L.i64p)

        def af(x):
            x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
            keep.append(x)
            return L.ptr(x, L.f64p)

        d = L.MeshDesc()
        d.num1el, d.num2el, d.numel = m.num1el, m.num2el, m.numel
        d.num1np, d.num2np, d.numnp = m.num1np, m.num2np, m.numnp
        d.ndf, d.nmdf = m.ndf, m.nmdf
        d.IX, d.ID = a64(m.IX), a64(m.ID)
        d.v_order[:] = v_order(m.dofs)
        d.m_order[:] = m_order(m.dofs)
        d.lam_order = m.dofs.get(Unknown.lam, 0)
        d.p_order = m.dofs.get(Unknown.pm, 0)
        d.area_n_uel = m.area_ufns.shape[1]
        d.area_uel_ids = a64(m.area_uel_ids)
        d.area_ufns = af(m.area_ufns)
        for k, b in enumerate(_BDRY):
            d.bdry_n_el[k] = len(m.bdry_elems[b])
            d.bdry_elems[k] = a64(m.bdry_elems[b])
            d.bdry_n_uel[k] = m.bdry_ufns[b].shape[1]
            d.bdry_uel_ids[k] = a64(m.bdry_uel_ids[b])
            d.bdry_ufns[k] = af(m.bdry_ufns[b])
        nb = np.array([int(t[0]) for t in m.inh_neu_bcs], dtype=np.int32)
        nt = np.array([int(t[1]) for t in m.inh_neu_bcs], dtype=np.int32)
        nv = np.array([float(t[2]) for t in m.inh_neu_bcs], dtype=np.float64)
        keep += [nb, nt, nv]
        d.n_neu = len(nb)
        d.neu_bdry, d.neu_type, d.neu_val = L.ptr(nb, L.i32p), L.ptr(nt, L.i32p), L.ptr(nv, L.f64p)
        if m.crnr_ufns is not None:
            d.crnr_ufns = af(np.ascontiguousarray(m.crnr_ufns))
        q = L.ParamsDesc(int(p.motion), int(p.scenario), p.kb, p.kg, p.zv, p.pn, p.adb, p.am,
                         p.ek, p.enr, bend_tm)
        h = ctypes.c_void_p()
        rc = self.lib.mafem_create(ctypes.byref(h), ctypes.byref(d), ctypes.byref(q))
        if rc != L.OK:
            raise L.LibError(rc, self.lib.mafem_last_error(None).decode())
        self.h = h
        self.nmdf = m.nmdf
        self._nnz = None

    # -- plumbing
    def _check(self, rc):
        if rc not in (L.OK, L.NOT_CONVERGED):
            raise L.LibError(rc, self.lib.mafem_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.mafem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state
    def set_state(self, xms, cps):
        x = np.asfortranarray(xms, dtype=np.float64)
        c = np.asfortranarray(cps, dtype=np.float64)
        self._check(self.lib.mafem_set_state(self.h, L.ptr(x, L.f64p), L.ptr(c, L.f64p)))

    def get_state(self):
        m = self.mesh
        x = np.zeros((m.numnp, 3), order="F")
        c = np.zeros((m.numnp, m.ndf), order="F")
        self._check(self.lib.mafem_get_state(self.h, L.ptr(x, L.f64p), L.ptr(c, L.f64p)))
        return x, c

    def set_prescribed(self, nodes, dofs, vals):
        n = np.ascontiguousarray(nodes, dtype=np.int64)
        d = np.ascontiguousarray(dofs, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        self._check(self.lib.mafem_set_prescribed(self.h, len(n), L.ptr(n, L.i64p),
                                                  L.ptr(d, L.i32p), L.ptr(v, L.f64p)))

    def predict(self, dt):
        self._check(self.lib.mafem_predict(self.h, float(dt)))

    def set_dof_ids(self, ID):
        """Replace the equation numbering (mesh.ID, ndf x numnp, 0 = prescribed)
        of this context; its mesh and state are kept."""
        a = np.ascontiguousarray(np.asarray(ID, dtype=np.int64).ravel(order="F"))
        nmdf = int(a.max())
        self._check(self.lib.mafem_set_dof_ids(self.h, L.ptr(a, L.i64p), nmdf))
        self.nmdf = nmdf
        self._nnz = None

    # -- hot path
    def nnz(self):
        if self._nnz is None:
            n = ctypes.c_int64()
            self._check(self.lib.mafem_nnz(self.h, ctypes.byref(n)))
            self._nnz = n.value
        return self._nnz

    def assemble(self, time, dt, want_K=True):
        n, nz = self.nmdf, self.nnz()
        r = np.zeros(n)
        if not want_K:
            self._check(self.lib.mafem_assemble(self.h, time, dt, L.ptr(r, L.f64p), None, None,
                                                None))
            return r, None
        cp = np.zeros(n + 1, dtype=np.int64)
        rv = np.zeros(nz, dtype=np.int64)
        v = np.zeros(nz)
        self._check(self.lib.mafem_assemble(self.h, time, dt, L.ptr(r, L.f64p), L.ptr(cp, L.i64p),
                                            L.ptr(rv, L.i64p), L.ptr(v, L.f64p)))
        K = sps.csc_matrix((v, rv - 1, cp - 1), shape=(n, n))
        return r, K

    def solve(self, rhs):
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros(self.nmdf)
        rel = ctypes.c_double()
        self._check(self.lib.mafem_solve(self.h, L.ptr(b, L.f64p), L.ptr(x, L.f64p),
                                         ctypes.byref(rel)))
        return x, rel.value

    def time_step(self, time, dt, max_iter=14):
        it = ctypes.c_int32()
        eps = np.zeros(max_iter)
        rc = self._check(self.lib.mafem_time_step(self.h, float(time), float(dt), int(max_iter),
                                                  ctypes.byref(it), L.ptr(eps, L.f64p)))
        return rc, list(eps[: it.value])

    def set_solver(self, mode="direct", precision=64, restart=30, max_iters=200,
                   refactor_iters=20, refine_sweeps=3, rtol=1e-13):
        """Linear solver of the Newton step: "direct" (band LU every iteration +
        FP64 refinement) or "krylov" (FP64 FGMRES preconditioned by a kept band
        LU, FP64 or FP32)."""
        o = L.SolverOpts(L.SOLVER_KRYLOV if mode == "krylov" else L.SOLVER_DIRECT,
                         int(precision), int(restart), int(max_iters), int(refactor_iters),
                         int(refine_sweeps), float(rtol))
        self._check(self.lib.mafem_set_solver(self.h, ctypes.byref(o)))

    def set_ordering(self, ordering="band"):
        """Fill-reducing ordering of the factorization: "band" (natural node
        order, band LU) or "nd" (nested dissection, multifrontal LU); takes
        effect at the next factorization."""
        o = {"band": L.ORDER_BAND, "nd": L.ORDER_ND}[ordering]
        self._check(self.lib.mafem_set_ordering(self.h, o))

    def get_solver(self):
        o = L.SolverOpts()
        self._check(self.lib.mafem_get_solver(self.h, ctypes.byref(o)))
        return {k: getattr(o, k) for k, _ in L.SolverOpts._fields_}

    def pull_force(self):
        """(x_pull, f_pull) of the device state: the mean position of the
        pulled element's control points and calc_pull_force
        (PullForce.jl:61-80, Analysis.jl:82-85)."""
        x = np.zeros(3)
        f = np.zeros(3)
        self._check(self.lib.mafem_pull_force(self.h, L.ptr(x, L.f64p), L.ptr(f, L.f64p)))
        return x, f

    def generate_output(self):
        """generate_output (Output.jl:30-118) of the device state: (xout,
        uout) of shapes (3 num1el + 2, 3 num2el + 2, 3) and (..., ndf)."""
        m = self.mesh
        shape = (3 * m.num1el + 2, 3 * m.num2el + 2)
        x = np.zeros(shape + (3,), order="F")
        u = np.zeros(shape + (m.ndf,), order="F")
        self._check(self.lib.mafem_generate_output(self.h, L.ptr(x, L.f64p), L.ptr(u, L.f64p)))
        return x, u

    def stats(self):
        s = L.Stats()
        self._check(self.lib.mafem_get_stats(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in L.Stats._fields_}


class Batch:
    """Many DeviceMesh contexts advanced through one time step together.

    No counterpart in the reference (one mesh per `solve`); each member's
    Newton step keeps the reference's definition (FiniteElement.jl:11-63).
    The linear solves run as FP64 GMRES(restart) right-preconditioned by a
    band LU kept in `precision` (32: mixed precision).  The members are
    borrowed: keep them alive while the batch exists.
    """

    def __init__(self, members, precision=32, restart=20, max_cycles=1, refactor_iters=6,
                 use_graphs=True, rtol=1e-10, escalate_iters=12):
        self.lib = L.load()
        self.members = list(members)
        n = len(self.members)
        arr = (ctypes.c_void_p * n)(*[m.h.value for m in self.members])
        o = L.BatchOpts(int(precision), int(restart), int(max_cycles), int(refactor_iters),
                        1 if use_graphs else 0, int(escalate_iters), float(rtol))
        h = ctypes.c_void_p()
        rc = self.lib.mafem_batch_create(ctypes.byref(h), n, arr, ctypes.byref(o))
        if rc != L.OK:
            raise L.LibError(rc, self.lib.mafem_batch_last_error(None).decode())
        self.h = h
        self.n = n

    def time_step(self, times, dts, predict=True, max_iter=14, active=None):
        """One backward-Euler step of the selected members (`active`: one flag
        per member, None = all); returns (iters, eps histories, status) with
        status 0 = converged, 1 = not converged, 2 = not selected."""
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(times, dtype=np.float64), (self.n,)))
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, dtype=np.float64), (self.n,)))
        it = np.zeros(self.n, dtype=np.int32)
        st = np.zeros(self.n, dtype=np.int32)
        eh = np.zeros((self.n, max_iter))
        act = None
        if active is not None:
            act = np.ascontiguousarray(np.asarray(active, dtype=np.uint8))
            assert act.shape == (self.n,)
        rc = self.lib.mafem_batch_time_step(self.h, L.ptr(t, L.f64p), L.ptr(d, L.f64p),
                                            None if act is None else L.ptr(act, L.u8p),
                                            1 if predict else 0, int(max_iter),
                                            L.ptr(it, L.i32p), L.ptr(eh, L.f64p),
                                            L.ptr(st, L.i32p))
        if rc not in (L.OK, L.NOT_CONVERGED):
            raise L.LibError(rc, self.lib.mafem_batch_last_error(self.h).decode())
        return it, [list(eh[k, : it[k]]) for k in range(self.n)], st

    def pull_force(self):
        """Per member (x_pull[3], f_pull[3]) after the last batch step, as an
        (n, 6) array (DeviceMesh.pull_force of every member)."""
        out = np.zeros((self.n, 6))
        rc = self.lib.mafem_batch_pull_force(self.h, L.ptr(out, L.f64p))
        if rc != L.OK:
            raise L.LibError(rc, self.lib.mafem_batch_last_error(self.h).decode())
        return out

    def stats(self):
        s = L.BatchStats()
        self.lib.mafem_batch_get_stats(self.h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in L.BatchStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.mafem_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_for(mesh, p, **args):
    """The DeviceMesh cached on a mesh (rebuilt if the parameters change)."""
    key = (p.motion, p.scenario, p.kb, p.kg, p.zv, p.pn, p.adb, p.am, p.enr,
           args.get("bend_tm", 1.0))
    cur = mesh._ctx
    if cur is None or cur[0] != key:
        if cur is not None:
            cur[1].close()
        mesh._ctx = (key, DeviceMesh(mesh, p, bend_tm=args.get("bend_tm", 1.0)))
    return mesh._ctx[1]


def update_xms(motion, xms, cps, dt, dofs):
    """update_xms! (FiniteElement.jl:408-423) on host arrays, in place."""
    for j, c in enumerate(m_motion_order(motion, dofs)):
        if c != 0:
            xms[:, j] += dt * cps[:, c - 1]


def calc_r_K(mesh, xms, cps, time, dt, p, **args):
    """calc_r_K (FiniteElement.jl:75-200): (r_gl, K_gl as scipy CSC)."""
    dev = device_for(mesh, p, **args)
    dev.set_state(xms, cps)
    return dev.assemble(float(time), float(dt))


def time_step(mesh, xms, cps, time, dt, p, **args):
    """time_step! (FiniteElement.jl:11-63): Newton on the GPU; mutates xms, cps.

    Raises AssertionError("did not reach Newton--Raphson tolerance") when the
    last ε = |Δu|_2 / nmdf is not below p.enr after at most 14 iterations.
    """
    dev = device_for(mesh, p, **args)
    dev.set_state(xms, cps)
    rc, eps = dev.time_step(time, dt)
    x, c = dev.get_state()
    xms[...] = x
    cps[...] = c
    if p.output:
        out.append(os.path.join(args["out_path"], args["out_file"]), out.iteration_lines(eps))
    if rc == L.NOT_CONVERGED:
        raise AssertionError("did not reach Newton--Raphson tolerance")
    return eps


def get_adj_maps(num1el, numel, IX, poly=2):
    """get_adj_maps (PullForce.jl:25-53): the 1-based ids of the (2 poly + 1)^2
    elements around the pulled one, taken by id arithmetic as the reference
    does, and per element the pair (rows of the pulled element's velocity
    residual, rows of the adjacent element's) of their shared nodes, 0-based,
    3 rows per node, local nodes in increasing order on both sides."""
    ix = np.asarray(IX)
    pull = get_pull_el_id(numel)
    pull_nodes = set(int(v) for v in ix[:, pull - 1])
    ids = [pull + i * num1el + j for i in range(-poly, poly + 1) for j in range(-poly, poly + 1)]
    maps = []
    for e in ids:
        if not 1 <= e <= numel:
            raise IndexError(f"adjacent element {e} outside 1..{numel}")
        adj = [int(v) for v in ix[:, e - 1]]
        rp = [3 * a + c for a in range(9) if int(ix[a, pull - 1]) in adj for c in range(3)]
        ra = [3 * b + c for b in range(9) if adj[b] in pull_nodes for c in range(3)]
        maps.append((rp, ra))
    return ids, maps


def lateral_pull_element(mesh, k, steps, fraction=0.4):
    """Tethered element at time step k (0-based) of a tether translated across
    `fraction` of the patch width over `steps` steps: the reference's element
    ceil(numel / 2) (which lies on the right edge for an even num1el) moved
    toward decreasing ζ1 by round(fraction num1el k / steps) elements.  An
    extension of this build: the reference has no moving tether."""
    shift = int(round(fraction * mesh.num1el * k / steps))
    return get_pull_el_id(mesh.numel) - shift


def move_tether(dev, mesh, p, old_el, new_el, speed):
    """Moves the F_PULL tether of a device context from element old_el to
    new_el (1-based): the context takes the numbering of the new Dirichlet
    set (tether_ids), the new tethered nodes get v_z (and v^m_z off LAG) =
    speed and their other velocity unknowns 0, and the released nodes'
    unknowns that stay prescribed return to 0; unknowns that become free keep
    their values.  Returns the new (ID, inh_dir_bcs)."""
    from .mesh import tether_ids
    ID, dirs = tether_ids(mesh, p, new_el, speed)
    dev.set_dof_ids(ID)
    inh = {(int(node), mesh.dofs[u]) for (u, node, _v) in dirs}
    nodes = set(int(v) for v in mesh.IX[:, old_el - 1]) | set(int(v) for v in mesh.IX[:, new_el - 1])
    pn, pd, pv = [], [], []
    for node in sorted(nodes):
        for d in range(1, mesh.ndf + 1):
            if ID[d - 1, node - 1] == 0:
                pn.append(node)
                pd.append(d)
                pv.append(float(speed) if (node, d) in inh else 0.0)
    dev.set_prescribed(pn, pd, pv)
    return ID, dirs


def calc_pull_force(mesh, xms, cps, adj_el_ids, adj_node_map, p, **args):
    """calc_pull_force (PullForce.jl:61-80) on the GPU: the pull force of the
    state (xms, cps).  The device evaluates get_adj_maps itself; maps other
    than get_adj_maps(mesh.num1el, mesh.numel, mesh.IX, 2) are rejected."""
    ids, maps = get_adj_maps(mesh.num1el, mesh.numel, mesh.IX, 2)
    if list(adj_el_ids) != ids or [tuple(map(list, m)) for m in adj_node_map] != \
            [tuple(m) for m in maps]:
        raise ValueError("only the adjacency of get_adj_maps with poly = 2 is supported")
    dev = device_for(mesh, p, **args)
    dev.set_state(xms, cps)
    return dev.pull_force()[1]


def run_analysis(mesh, xms, cps, p, **args):
    """run_analysis! (Analysis.jl:17-102): predictor + Newton per time step,
    with the state resident on the GPU; returns the ε history of each step.

    With p.output it writes the reference's files under args["out_path"]
    (times.txt, t<id>-xms.txt / t<id>-cps.txt after every step, f-pull.txt
    for F_PULL with the device-evaluated pull force, iteration lines in
    args["out_file"]); the state is downloaded only for those files.

    An optional `pull_schedule(t_id, time) -> speed` keyword re-applies the
    F_PULL prescribed speed before each step (what a restart with a new
    `pull_speed` does in the reference, src/Input.jl:103-108).
    """
    dev = device_for(mesh, p, **args)
    dts = list(args["dts"] if "dts" in args else args["Δts"])
    t0_id = int(args.get("t0_id", 0))
    times = np.cumsum(dts) + args.get("t0", 0.0)
    log = os.path.join(args["out_path"], args["out_file"]) if p.output else None
    if p.output:
        out.append(log, "\nStarting to run analysis...\n")
        out.append(os.path.join(args["out_path"], "times.txt"), out.times_text(times, t0_id))
    pulled = p.scenario == F_PULL
    if pulled:
        assert len(v_order(mesh.dofs)) == 3, "f_pull needs 3-D velocity"
        append_run = "in_path" in args and args.get("out_path") == args["in_path"]
        if p.output and not append_run:
            out.append(os.path.join(args["out_path"], "f-pull.txt"), out.PULL_HEADER)
    sched = args.get("pull_schedule")
    pres = None
    if sched is not None:
        pres = [(node, mesh.dofs[u]) for (u, node, _v) in mesh.inh_dir_bcs]
        if p.motion == 2:  # EUL also prescribes v^m_z at the pulled nodes
            pres += [(node, mesh.dofs[Unknown.vmz]) for (u, node, _v) in mesh.inh_dir_bcs
                     if u == Unknown.vz]
    dev.set_state(xms, cps)
    hist = []
    for k, dt in enumerate(dts):
        if p.output:
            out.append(log, f"\n-> Time {out.julia_float(times[k])}\n")
        if sched is not None:
            s = float(sched(k, float(times[k])))
            dev.set_prescribed([n for n, _ in pres], [d for _, d in pres], [s] * len(pres))
        dev.predict(dt)
        rc, eps = dev.time_step(float(times[k]), float(dt))
        hist.append(eps)
        if p.output:
            out.append(log, out.iteration_lines(eps))
        if rc == L.NOT_CONVERGED:
            x, c = dev.get_state()
            xms[...] = x
            cps[...] = c
            raise AssertionError("did not reach Newton--Raphson tolerance")
        if p.output:
            x, c = dev.get_state()
            out.write_state(args["out_path"], k + 1 + t0_id, x, c)
            if pulled:
                xp, fp = dev.pull_force()
                out.append(os.path.join(args["out_path"], "f-pull.txt"),
                           out.pull_line(times[k], xp, fp))
    if p.output:
        out.append(log, "\nCompleted running analysis...\n\n")
    x, c = dev.get_state()
    xms[...] = x
    cps[...] = c
    return hist


def generate_output(mesh, xms, cps, p=None, **args):
    """generate_output(mesh, xms, cps) (Output.jl:30-118) on the GPU: positions
    and unknowns at the area Gauss points, the edge Gauss points and the
    corners, as arrays (3 num1el + 2, 3 num2el + 2, 3) and (..., ndf).  Uses
    the mesh's cached context (or one built from `p`)."""
    if p is None:
        if mesh._ctx is None:
            raise ValueError("no context on this mesh: pass the Params")
        dev = mesh._ctx[1]
    else:
        dev = device_for(mesh, p, **args)
    dev.set_state(xms, cps)
    return dev.generate_output()


def solve(p, **args):
    """solve (src/MembraneAleFem.jl:44-60): parameters checked, mesh and
    initial state prepared, time steps run on the GPU; returns (mesh, xms, cps)."""
    from .mesh import make_mesh, prepare_input
    from .params import check_params
    if "Δts" in args and "dts" not in args:
        args["dts"] = args["Δts"]
    check_params(p, **args)
    mesh = make_mesh(p, **args)
    xms, cps = prepare_input(p, mesh, **args)
    run_analysis(mesh, xms, cps, p, **args)
    return mesh, xms, cps


def restart(p, args, **r_args):
    """restart (src/MembraneAleFem.jl:75-107) with the previous run's Params
    and keyword arguments given as objects (the reference deserializes them
    from files written by Julia's Serialization, which is not restated here):
    requires in_path, Δts (or dts), t0_id and t0; reads the state from
    t<t0_id>-xms.txt / t<t0_id>-cps.txt under in_path unless in_xms / in_cps
    are given; writes to in_path unless out_path is given."""
    for k in ("in_path", "t0_id", "t0"):
        assert k in r_args, f"\nneed to overwrite '{k}' during restart"
    assert "Δts" in r_args or "dts" in r_args, "\nneed to overwrite 'Δts' during restart"
    r = dict(r_args)
    r.setdefault("in_cps", f"t{r['t0_id']}-cps.txt")
    r.setdefault("in_xms", f"t{r['t0_id']}-xms.txt")
    r.setdefault("out_path", r["in_path"])
    merged = dict(args)
    merged.update(r)
    if "Δts" in r:
        merged["dts"] = r["Δts"]
    return solve(p, **merged)

