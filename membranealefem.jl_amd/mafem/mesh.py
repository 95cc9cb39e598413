"""Host-side mesh construction: the tables the reference's `Mesh` holds.

This mirrors the reference's input layer (which stays on the host in the
reference's own language): knot vectors (src/input/Spline.jl:20-183), tabulated
B-spline basis functions at Gauss points (src/input/GpBasisFn.jl:20-356), the
element/node numbering (src/input/Mesh.jl:94-302, 574-593) and the scenario
boundary conditions (src/input/Bc.jl:15-431).  The arrays use the reference's
memory layouts (1-based Int64, column-major), so they cross the C ABI as they
would from Julia.
"""
import math
import os
from dataclasses import dataclass, field

import numpy as np

from .enums import (ALEV, ALEVB, BOTTOM, EUL, F_BEND, F_CAVI, F_COUE, F_POIS, F_PULL, LAG,
                    LEFT, MOMENT, RIGHT, STATIC, STRETCH, TOP, Unknown)
from .params import GP1D, NEN, POLY

REC = 55
_XI = (-math.sqrt(3 / 5), 0.0, math.sqrt(3 / 5))
_WXI = (5 / 9, 8 / 9, 5 / 9)


# ------------------------------------------------------------------ knots

def uniform_knots(nel, p=POLY):
    """Clamped uniform knots on [0, 1] (Spline.jl:70-89)."""
    inner = [k / nel for k in range(1, nel)]
    return [0.0] * (p + 1) + inner + [1.0] * (p + 1)


def centre_refined_knots(nel, p=POLY):
    """Clamped knots refined towards the centre (get_fine_ζs, Spline.jl:122-183).

    Six wide elements per side over 1/3 of the length each; if at least a
    quarter of the remaining elements exceeds three, a second level puts
    `n_w2` elements over 1/9 on each side of the centre.
    """
    assert nel >= 18, "fine mesh requires at least 18 1-D elements"
    nw1, zw1 = 6, 1 / 3
    zf1 = 1.0 - 2 * zw1
    nf1 = nel - 2 * nw1 - 1
    d1 = zw1 / nw1
    left = [k * d1 for k in range(1, nw1 + 1)]
    right = [1.0 - zw1 + k * d1 for k in range(0, nw1)]
    nw2 = nf1 // 4
    if 2 * nw2 <= nw1:
        dc = zf1 / (nf1 + 1)
        centre = [zw1 + k * dc for k in range(1, nf1 + 1)]
    else:
        zw2 = 1 / 9
        nf2 = nf1 - 2 * nw2
        zf2 = zf1 - 2 * zw2
        d2 = zw2 / nw2
        l2 = [zw1 + k * d2 for k in range(1, nw2 + 1)]
        r2 = [1.0 - zw1 - zw2 + k * d2 for k in range(0, nw2)]
        dc = zf2 / (nf2 + 1)
        centre = l2 + [zw1 + zw2 + k * dc for k in range(1, nf2 + 1)] + r2
    return [0.0] * (p + 1) + left + centre + right + [1.0] * (p + 1)


def span_index(knots, z, p=POLY):
    """0-based span s with knots[s] <= z < knots[s+1]; the last span at z = max."""
    n = len(knots) - p - 2           # last valid span
    if z >= knots[n + 1]:
        return n
    return int(np.searchsorted(knots, z, side="right")) - 1


def basis_ders(knots, z, p=POLY, nd=2):
    """Nonzero B-splines N_{s-p..s}(z) and derivatives up to nd (rows: order).

    Builds the triangular table of lower-degree functions, then applies the
    derivative recurrence N'_{i,k} = k (N_{i,k-1}/(u_{i+k}-u_i) -
    N_{i+1,k-1}/(u_{i+k+1}-u_{i+1})) on it.
    """
    s = span_index(knots, z, p)
    u = knots
    # table[k][j] = N_{s-k+j, k}(z), j = 0..k
    table = [[1.0]]
    for k in range(1, p + 1):
        prev = table[-1]
        row = [0.0] * (k + 1)
        for j in range(k + 1):
            i = s - k + j
            a = 0.0
            if j >= 1:                      # N_{i,k-1} term
                den = u[i + k] - u[i]
                a += (z - u[i]) / den * prev[j - 1] if den != 0 else 0.0
            if j <= k - 1:                  # N_{i+1,k-1} term
                den = u[i + k + 1] - u[i + 1]
                a += (u[i + k + 1] - z) / den * prev[j] if den != 0 else 0.0
            row[j] = a
        table.append(row)
    out = np.zeros((nd + 1, p + 1))
    out[0] = table[p]
    for d in range(1, nd + 1):
        low = table[p - d]                   # N_{s-(p-d)+j, p-d}, j = 0..p-d
        for j in range(p + 1):
            i = s - p + j
            # d-th derivative of N_{i,p} as sum_c w[c] N_{i+c, p-d}
            w = [1.0]
            for k in range(p, p - d, -1):
                nw = [0.0] * (len(w) + 1)
                for c, wc in enumerate(w):
                    den1 = u[i + c + k] - u[i + c]
                    den2 = u[i + c + k + 1] - u[i + c + 1]
                    if den1 != 0:
                        nw[c] += wc * k / den1
                    if den2 != 0:
                        nw[c + 1] -= wc * k / den2
                w = nw
            acc = 0.0
            for c, wc in enumerate(w):
                jj = (i + c) - (s - (p - d))
                if 0 <= jj <= p - d:
                    acc += wc * low[jj]
            out[d, j] = acc
    return s, out


def unique_elements(knots, p=POLY):
    """Map each element to a unique element by the knot spacing around it
    (get_unique_1d_elements, Spline.jl:640-671); 1-based ids."""
    nk = len(knots)
    tol = math.sqrt(np.finfo(float).eps)
    prev = None
    ids, spans = [], []
    for k in range(p, nk - p - 1):           # 0-based span of element k - p
        ctx = [knots[i + 1] - knots[i] for i in range(k - p, k + p + 1)]
        same = prev is not None and all(
            a == b or abs(a - b) <= tol * max(abs(a), abs(b)) for a, b in zip(ctx, prev))
        if not same:
            spans.append((knots[k], knots[k + 1]))
        ids.append(len(spans))
        prev = ctx
    return ids, spans


class Line:
    """1-D tables on unique elements plus the end-point functions."""

    def __init__(self, knots):
        self.knots = knots
        self.uel_ids, spans = unique_elements(knots)
        self.nel = len(self.uel_ids)
        self.fns = []                          # [u][g] -> (w, ders[3x3])
        for lo, hi in spans:
            row = []
            for g in range(GP1D):
                z = _XI[g] * (hi - lo) / 2 + (hi + lo) / 2
                w = _WXI[g] * (hi - lo) / 2
                row.append((w, basis_ders(knots, z)[1]))
            self.fns.append(row)
        self.zmin = (1.0, basis_ders(knots, knots[0])[1])
        self.zmax = (1.0, basis_ders(knots, knots[-1])[1])


def tensor_record(f1, f2):
    """2-D basis record (w, N, dN_1, dN_2, ddN_11, ddN_22, ddN_12), node
    a = i1 + 3 i2 (GpBasisFn.jl:96-112)."""
    w1, d1 = f1
    w2, d2 = f2
    rec = np.empty(REC)
    rec[0] = w1 * w2
    rec[1:10] = np.outer(d2[0], d1[0]).ravel()
    rec[10:19] = np.outer(d2[0], d1[1]).ravel()
    rec[19:28] = np.outer(d2[1], d1[0]).ravel()
    rec[28:37] = np.outer(d2[0], d1[2]).ravel()
    rec[37:46] = np.outer(d2[2], d1[0]).ravel()
    rec[46:55] = np.outer(d2[1], d1[1]).ravel()
    return rec


# ------------------------------------------------------------------ mesh

def get_dofs(motion):
    """get_dofs (Bc.jl:414-431): Unknown -> column of cps (1-based)."""
    U = Unknown
    if motion in (LAG, STATIC):
        return {U.vx: 1, U.vy: 2, U.vz: 3, U.lam: 4}
    if motion == EUL:
        return {U.vx: 1, U.vy: 2, U.vz: 3, U.vmx: 4, U.vmy: 5, U.vmz: 6, U.lam: 7}
    if motion in (ALEV, ALEVB):
        return {U.vx: 1, U.vy: 2, U.vz: 3, U.vmx: 4, U.vmy: 5, U.vmz: 6, U.lam: 7, U.pm: 8}
    raise AssertionError(f"{motion} motion degrees of freedom not provided")


def v_order(dofs):
    return [dofs.get(Unknown.vx, 0), dofs.get(Unknown.vy, 0), dofs.get(Unknown.vz, 0)]


def m_order(dofs):
    return [dofs.get(Unknown.vmx, 0), dofs.get(Unknown.vmy, 0), dofs.get(Unknown.vmz, 0)]


def m_motion_order(motion, dofs):
    """get_m_motion_order (Mesh.jl:529-542)."""
    if motion == STATIC:
        return [0, 0, 0]
    return v_order(dofs) if motion == LAG else m_order(dofs)


def get_pull_el_id(numel):
    """Pulled element: ceil(numel / 2) (get_pull_el_id, PullForce.jl:8-13)."""
    return (numel + 1) // 2


@dataclass
class Mesh:
    num1el: int
    num2el: int
    numel: int
    num1np: int
    num2np: int
    numnp: int
    IX: np.ndarray                      # (9, numel) int64, 1-based, Fortran order
    bdry_elems: dict
    bdry_nodes: dict
    bdry_inner_nodes: dict
    crnr_nodes: dict
    kv1: list
    kv2: list
    area_uel_ids: np.ndarray            # (numel,) int64, 1-based
    area_ufns: np.ndarray               # (9, n_uel, 55): Julia Matrix(n_uel, 9) memory order
    bdry_uel_ids: dict                  # Boundary -> (n,) int64, 1-based
    bdry_ufns: dict                     # Boundary -> (3, n_uel, 55)
    dofs: dict
    ndf: int
    ID: np.ndarray                      # (ndf, numnp) int64, Fortran order, 0 = prescribed
    nmdf: int
    inh_dir_bcs: list
    inh_neu_bcs: list
    crnr_ufns: np.ndarray = None        # (4, 55): Corner order (Mesh.jl:228-233)
    _ctx: object = field(default=None, repr=False, compare=False)

    @property
    def LM(self):
        """reshape(ID[:, IX], ndf NEN, numel) (Mesh.jl:299)."""
        return self.ID[:, self.IX - 1].reshape(self.ndf * NEN, self.numel, order="F")

    @property
    def ID_inv(self):
        """(node, dof) of every equation, 1-based (Mesh.jl:291-296)."""
        d, n = np.nonzero(self.ID)
        order = np.argsort(self.ID[d, n])
        return np.stack([n[order] + 1, d[order] + 1], axis=1)


def make_mesh(p, **args):
    """Mesh(p; args...) (Mesh.jl:79, 94-302) for the flat topology."""
    assert p.scenario in (F_CAVI, F_COUE, F_POIS, F_PULL, F_BEND)
    n1e, n2e = p.num1el, p.num2el
    numel = n1e * n2e
    n1, n2 = n1e + POLY, n2e + POLY
    nn = n1 * n2
    bdry_elems = {BOTTOM: np.arange(1, n1e + 1), RIGHT: np.arange(n1e, numel + 1, n1e),
                  TOP: np.arange(numel - n1e + 1, numel + 1),
                  LEFT: np.arange(1, numel - n1e + 2, n1e)}
    bdry_nodes = {BOTTOM: np.arange(1, n1 + 1), RIGHT: np.arange(n1, nn + 1, n1),
                  TOP: np.arange(nn - n1 + 1, nn + 1), LEFT: np.arange(1, nn - n1 + 2, n1)}
    bdry_inner = {BOTTOM: bdry_nodes[BOTTOM] + n1, RIGHT: bdry_nodes[RIGHT] - 1,
                  TOP: bdry_nodes[TOP] - n1, LEFT: bdry_nodes[LEFT] + 1}
    crnr_nodes = {1: 1, 2: n1, 3: nn - n1 + 1, 4: nn}
    fine = p.scenario == F_PULL and n1e >= 18 and n2e >= 18
    kv1 = centre_refined_knots(n1e) if fine else uniform_knots(n1e)
    kv2 = centre_refined_knots(n2e) if fine else uniform_knots(n2e)
    l1, l2 = Line(kv1), Line(kv2)

    # area tables
    nu1, nu2 = len(l1.fns), len(l2.fns)
    u1 = np.asarray(l1.uel_ids, dtype=np.int64)
    u2 = np.asarray(l2.uel_ids, dtype=np.int64)
    area_uel = (u1[None, :] + (u2[:, None] - 1) * nu1).ravel()   # e = e1 + n1e e2
    area = np.empty((9, nu1 * nu2, REC))
    for a2 in range(nu2):
        for a1 in range(nu1):
            for g2 in range(GP1D):
                for g1 in range(GP1D):
                    area[g1 + 3 * g2, a1 + nu1 * a2] = tensor_record(l1.fns[a1][g1], l2.fns[a2][g2])
    # boundary tables (GpBasisFn.jl:258-277, Mesh.jl:220-225)
    def bdry(line, perp, along1):
        t = np.empty((GP1D, len(line.fns), REC))
        for u in range(len(line.fns)):
            for g in range(GP1D):
                t[g, u] = (tensor_record(line.fns[u][g], perp) if along1
                           else tensor_record(perp, line.fns[u][g]))
        return t
    bdry_ufns = {BOTTOM: bdry(l1, l2.zmin, True), RIGHT: bdry(l2, l1.zmax, False),
                 TOP: bdry(l1, l2.zmax, True), LEFT: bdry(l2, l1.zmin, False)}
    bdry_uel = {BOTTOM: u1, TOP: u1, RIGHT: u2, LEFT: u2}

    # IX (Mesh.jl:574-593): local node a = i1 + 3 i2 -> node (e1+i1) + n1 (e2+i2)
    e1, e2 = np.meshgrid(np.arange(n1e), np.arange(n2e), indexing="xy")
    e1, e2 = e1.ravel(), e2.ravel()
    ix = np.empty((9, numel), dtype=np.int64, order="F")
    for a2 in range(3):
        for a1 in range(3):
            ix[a1 + 3 * a2] = (e1 + a1) + n1 * (e2 + a2) + 1

    dofs, ndf, fixed, dir_bcs, neu_bcs = _scenario_bcs(p, nn, ix, bdry_nodes, bdry_inner,
                                                       crnr_nodes, **args)
    free = ~fixed                                     # (ndf, nn)
    ID = np.zeros((ndf, nn), dtype=np.int64, order="F")
    flat = free.ravel(order="F")                      # node-major, DOF fastest
    ID.ravel(order="F")[:] = np.where(flat, np.cumsum(flat), 0)
    crnr_ufns = np.array([tensor_record(l1.zmin, l2.zmin), tensor_record(l1.zmax, l2.zmin),
                          tensor_record(l1.zmin, l2.zmax), tensor_record(l1.zmax, l2.zmax)])
    return Mesh(n1e, n2e, numel, n1, n2, nn, ix, bdry_elems, bdry_nodes, bdry_inner, crnr_nodes,
                kv1, kv2, area_uel, area, bdry_uel, bdry_ufns, dofs, ndf, ID, int(ID.max()),
                dir_bcs, neu_bcs, crnr_ufns)


def tether_ids(mesh, p, pull_el, pull_speed):
    """(ID, inh_dir_bcs) of the F_PULL scenario on `mesh` with the tether on
    element `pull_el` (1-based) instead of ceil(numel / 2): the numbering and
    Dirichlet list generate_scenario would build (Mesh.jl:262-302,
    Bc.jl:188-269) for that element.  An extension of this build (the
    reference has no moving tether); pull_el = ceil(numel / 2) gives mesh.ID."""
    _, ndf, fixed, dirs, _ = _scenario_bcs(p, mesh.numnp, mesh.IX, mesh.bdry_nodes,
                                           mesh.bdry_inner_nodes, mesh.crnr_nodes,
                                           pull_speed=pull_speed, pull_el=pull_el)
    free = ~fixed
    ID = np.zeros((ndf, mesh.numnp), dtype=np.int64, order="F")
    flat = free.ravel(order="F")
    ID.ravel(order="F")[:] = np.where(flat, np.cumsum(flat), 0)
    return ID, dirs


def _scenario_bcs(p, nn, ix, bn, bin_, crnr, **args):
    """get_scenario_bc_info (Bc.jl:15-393): DOFs, fixed mask, Dirichlet and
    Neumann lists."""
    U = Unknown
    if p.scenario in (F_CAVI, F_COUE, F_POIS):
        dofs = {U.vx: 1, U.vy: 2, U.lam: 3}
    else:
        dofs = get_dofs(p.motion)
    ndf = len(dofs)
    fixed = np.zeros((ndf, nn), dtype=bool)
    dirs, neus = [], []

    def fix(u, nodes):
        fixed[dofs[u] - 1, np.asarray(nodes) - 1] = True

    ale = p.motion in (ALEV, ALEVB)
    if p.scenario == F_PULL:
        speed = args["pull_speed"]
        allb = np.concatenate([bn[b] for b in (BOTTOM, RIGHT, TOP, LEFT)])
        alli = np.concatenate([bin_[b] for b in (BOTTOM, RIGHT, TOP, LEFT)])
        fix(U.vz, allb)
        if ale:
            fix(U.vmx, allb); fix(U.vmy, allb); fix(U.vmz, allb)
        fix(U.vz, alli)
        if p.motion == ALEVB:
            fix(U.vmz, alli)
        # `pull_el` (1-based; an extension of this build, no reference
        # counterpart) moves the tether; the reference pulls element
        # ceil(numel / 2)
        pe = int(args.get("pull_el", get_pull_el_id(ix.shape[1])))
        for node in ix[:, pe - 1]:
            node = int(node)
            fix(U.vx, [node]); fix(U.vy, [node]); fix(U.vz, [node])
            dirs.append((U.vz, node, speed))
            if p.motion != LAG:
                fix(U.vmx, [node]); fix(U.vmy, [node]); fix(U.vmz, [node])
                dirs.append((U.vmz, node, speed))
        for b in (BOTTOM, RIGHT, TOP, LEFT):
            c = int(bn[b][len(bn[b]) // 2])
            fix(U.vx, [c]); fix(U.vy, [c])
            if p.motion != LAG:
                fix(U.vmx, [c]); fix(U.vmy, [c])
        lv = p.kb / 4
        neus = [(LEFT, STRETCH, lv), (RIGHT, STRETCH, lv), (TOP, STRETCH, lv),
                (BOTTOM, STRETCH, lv)]
        assert p.pn == 0.0, "F_PULL with a normal pressure is not implemented"
    elif p.scenario == F_BEND:
        for b in (TOP, BOTTOM):
            fix(U.vy, bn[b])
            if ale:
                fix(U.vmy, bn[b])
        fix(U.vx, bn[LEFT]); fix(U.vy, bn[LEFT]); fix(U.vz, bn[LEFT])
        if ale:
            fix(U.vmx, bn[LEFT]); fix(U.vmy, bn[LEFT]); fix(U.vmz, bn[LEFT])
        fix(U.vz, bn[RIGHT])
        if ale:
            fix(U.vmz, bn[RIGHT])
        neus = [(LEFT, MOMENT, args["bend_mf"]), (RIGHT, MOMENT, args["bend_mf"])]
    elif p.scenario == F_CAVI:
        for b in (BOTTOM, TOP, LEFT, RIGHT):
            fix(U.vx, bn[b]); fix(U.vy, bn[b])
        for i in bn[TOP]:
            if i not in (crnr[3], crnr[4]):
                dirs.append((U.vx, int(i), 1.0))
        c = nn // 2 + 1
        fix(U.lam, [c])
        dirs.append((U.lam, c, 0.0))
    elif p.scenario == F_COUE:
        fix(U.vx, bn[BOTTOM]); fix(U.vy, bn[BOTTOM])
        fix(U.vx, bn[TOP]); fix(U.vy, bn[TOP])
        dirs += [(U.vx, int(i), 3.0) for i in bn[TOP]]
        fix(U.vy, bn[LEFT]); fix(U.vy, bn[RIGHT])
        neus = [(LEFT, STRETCH, 4.0), (RIGHT, STRETCH, 4.0)]
    else:  # F_POIS
        for b in (TOP, BOTTOM):
            fix(U.vx, bn[b]); fix(U.vy, bn[b])
        fix(U.vy, bn[LEFT]); fix(U.vy, bn[RIGHT])
        neus = [(LEFT, STRETCH, 4.0), (RIGHT, STRETCH, 8.0)]
    return dofs, ndf, fixed, dirs, neus


def collocation_cps_1d(knots, f):
    """Control values reproducing f at the collocation points: both ends and
    the midpoint of every knot span (get_1d_bspline_cps / collocate_ζ,
    Spline.jl:522-538, 581-626 for degree 2)."""
    uk = sorted(set(knots))
    z = [uk[0]] + [(uk[i] + uk[i + 1]) / 2 for i in range(len(uk) - 1)] + [uk[-1]]
    nb = len(knots) - POLY - 1
    assert len(z) == nb
    A = np.zeros((nb, nb))
    for j, zz in enumerate(z):
        s, d = basis_ders(knots, zz)
        A[j, s - POLY:s + 1] = d[0]
    return np.linalg.solve(A, np.array([f(zz) for zz in z]))


def _bend_initial_state(p, mesh, **args):
    """F_BEND initial state (Input.jl:44-75): x = L ζ1, y = L ζ2, z = 0;
    velocities (and mesh velocities unless LAG) drawn uniformly, vx, vy from
    (U - 1/2) / numel and vz from U / numel, then zero on the boundary nodes;
    λ = kb / 4 / L^2 (Δt_1 / bend_tm)^2.  The reference draws with Julia's
    unseeded rand(); here the draws come from numpy's default_rng(bend_seed)
    (bend_seed = 0 unless given), in the same order."""
    rng = np.random.default_rng(int(args.get("bend_seed", 0)))
    L = p.length
    nn, ne = mesh.numnp, mesh.numel
    c1 = collocation_cps_1d(mesh.kv1, lambda z: L * z)
    c2 = collocation_cps_1d(mesh.kv2, lambda z: L * z)
    xms = np.zeros((nn, 3), order="F")
    xms[:, 0] = np.tile(c1, mesh.num2np)
    xms[:, 1] = np.repeat(c2, mesh.num1np)
    cps = np.zeros((nn, mesh.ndf), order="F")
    d = mesh.dofs
    cps[:, d[Unknown.vx] - 1] = (rng.random(nn) - 0.5) / ne
    cps[:, d[Unknown.vy] - 1] = (rng.random(nn) - 0.5) / ne
    cps[:, d[Unknown.vz] - 1] = rng.random(nn) / ne
    if p.motion != LAG:
        cps[:, d[Unknown.vmx] - 1] = (rng.random(nn) - 0.5) / ne
        cps[:, d[Unknown.vmy] - 1] = (rng.random(nn) - 0.5) / ne
        cps[:, d[Unknown.vmz] - 1] = rng.random(nn) / ne
    for b in (BOTTOM, RIGHT, TOP, LEFT):
        cps[np.asarray(mesh.bdry_nodes[b]) - 1, :] = 0.0
    dts = args["dts"] if "dts" in args else args["Δts"]
    cps[:, d[Unknown.lam] - 1] = p.kb / 4 / L ** 2 * (dts[0] / args["bend_tm"]) ** 2
    return xms, cps


def prepare_input(p, mesh, **args):
    """prepare_input (Input.jl:22-121).

    Positions fit x = L(ζ1 - 1/2), y = L(ζ2 - 1/2), z = 0 (the reference's
    2-D collocation solve factorises into these 1-D solves); λ = kb/4; then
    the inhomogeneous Dirichlet values.  With `in_xms`/`in_cps` arrays given,
    they are used instead (the restart path, Input.jl:35-37).
    """
    if "in_xms" in args and "in_cps" in args:
        ix, ic = args["in_xms"], args["in_cps"]
        if isinstance(ix, str):  # file names under in_path, read as readdlm does
            from .output import readdlm
            ix = readdlm(os.path.join(args["in_path"], ix))
            ic = readdlm(os.path.join(args["in_path"], ic))
        xms = np.array(ix, dtype=np.float64, order="F")
        cps = np.array(ic, dtype=np.float64, order="F")
        if xms.shape != (mesh.numnp, 3) or cps.shape != (mesh.numnp, mesh.ndf):
            raise ValueError("restart state does not match the mesh")
    elif p.scenario == F_BEND:
        xms, cps = _bend_initial_state(p, mesh, **args)
    else:
        L = p.length
        c1 = collocation_cps_1d(mesh.kv1, lambda z: L * (z - 0.5))
        c2 = collocation_cps_1d(mesh.kv2, lambda z: L * (z - 0.5))
        xms = np.zeros((mesh.numnp, 3), order="F")
        xms[:, 0] = np.tile(c1, mesh.num2np)
        xms[:, 1] = np.repeat(c2, mesh.num1np)
        cps = np.zeros((mesh.numnp, mesh.ndf), order="F")
        cps[:, mesh.dofs[Unknown.lam] - 1] = p.kb / 4
    for unk, node, val in mesh.inh_dir_bcs:
        cps[node - 1, mesh.dofs[unk] - 1] = val
        if p.scenario == F_PULL and p.motion == EUL:
            cps[node - 1, mesh.dofs[Unknown.vmz] - 1] = val
    if p.output:  # initial condition files (Input.jl:111-114)
        from .output import write_state
        write_state(args["out_path"], int(args.get("t0_id", 0)), xms, cps)
    return xms, cps
