"""Oracle restatement of the reference's mesh, boundary conditions and inputs.

Follows (reference paths):
  src/input/Params.jl  Params defaults :36-55, constants :161-179
  src/input/Enums.jl   enum values :24-135
  src/input/Mesh.jl    generate_mesh :94-248, generate_scenario :262-302,
                       DOF order maps :477-542, construct_IX :574-593
  src/input/Bc.jl      scenario BCs :15-393, get_dofs :414-431
  src/Input.jl         prepare_input :22-121, get_pull_el_id :131-136

Pinned by: the unknown counts in tests/golden/mesh_counts.json (4 x 4 rows
counted by hand per unknown on the 6 x 6 node grid; nmdf at the benchmark
sizes), the F_PULL Dirichlet sets rebuilt from grid coordinates (boundary
ring, inner ring, the pulled element's 3 x 3 node block, edge-centre nodes)
and compared entry by entry for every motion (tests/test_oracle_ale.py), the
structure of IX (tensor-product node numbering), and agreement with the
independently written product-side mesh builder.
"""
from dataclasses import dataclass, field

import numpy as np

from . import spline as sp

# enums (Enums.jl)
F_CAVI, F_COUE, F_POIS, F_PULL, F_BEND = 1, 2, 3, 4, 5
STATIC, EUL, LAG, ALEV, ALEVB = 1, 2, 3, 4, 5
BOTTOM, RIGHT, TOP, LEFT = 1, 2, 3, 4
SHEAR, STRETCH, MOMENT = 1, 2, 3
# Dof.Unknown (Dof.jl:29-34)
VX, VY, VZ, VMX, VMY, VMZ, LAM, PM = 1, 2, 3, 4, 5, 6, 7, 8


@dataclass
class Params:
    """Params (Params.jl:36-55)."""
    motion: int = ALEVB
    scenario: int = F_PULL
    num1el: int = 17
    num2el: int = 17
    output: bool = False
    length: float = 64.0
    kb: float = 1.0
    kg: float = -0.5
    zv: float = 1.0
    pn: float = 0.0
    poly: int = 2
    gp1d: int = 3
    nders: int = 2
    nen: int = 9
    adb: float = None
    am: float = 1.0
    ek: float = 1e-15
    enr: float = 1e-12

    def __post_init__(self):
        if self.adb is None:
            self.adb = self.length ** 2


def get_dofs(motion):
    """get_dofs (Bc.jl:414-431)."""
    if motion in (LAG, STATIC):
        return {VX: 1, VY: 2, VZ: 3, LAM: 4}
    if motion == EUL:
        return {VX: 1, VY: 2, VZ: 3, VMX: 4, VMY: 5, VMZ: 6, LAM: 7}
    if motion in (ALEV, ALEVB):
        return {VX: 1, VY: 2, VZ: 3, VMX: 4, VMY: 5, VMZ: 6, LAM: 7, PM: 8}
    raise AssertionError("motion degrees of freedom not provided")


def v_order(dofs):
    return [dofs.get(VX, 0), dofs.get(VY, 0), dofs.get(VZ, 0)]


def m_order(dofs):
    return [dofs.get(VMX, 0), dofs.get(VMY, 0), dofs.get(VMZ, 0)]


def lam_order(dofs):
    return dofs.get(LAM, 0)


def p_order(dofs):
    return dofs.get(PM, 0)


def m_motion_order(motion, dofs):
    """get_m_motion_order (Mesh.jl:529-542)."""
    if motion == STATIC:
        return [0, 0, 0]
    if motion == LAG:
        return v_order(dofs)
    return m_order(dofs)


def get_pull_el_id(numel):
    """get_pull_el_id (PullForce.jl:8-13): ceil(numel/2)."""
    return -(-numel // 2)


@dataclass
class Mesh:
    num1el: int = 0
    num2el: int = 0
    numel: int = 0
    num1np: int = 0
    num2np: int = 0
    numnp: int = 0
    IX: np.ndarray = None
    bdry_elems: dict = field(default_factory=dict)
    bdry_nodes: dict = field(default_factory=dict)
    bdry_inner_nodes: dict = field(default_factory=dict)
    crnr_nodes: dict = field(default_factory=dict)
    kv1: object = None
    kv2: object = None
    area_uel_ids: list = None
    area_ufns: np.ndarray = None
    bdry_fns: dict = field(default_factory=dict)
    dofs: dict = None
    ndf: int = 0
    ID: np.ndarray = None
    ID_inv: list = None
    nmdf: int = 0
    LM: np.ndarray = None
    inh_dir_bcs: list = None
    inh_neu_bcs: list = None


def construct_IX(kv1, kv2, num1np):
    """construct_IX (Mesh.jl:574-593), 1-based node ids, NEN x numel."""
    ix = np.zeros((9, kv1.nel * kv2.nel), dtype=np.int64)
    for e2 in range(1, kv2.nel + 1):
        for e1 in range(1, kv1.nel + 1):
            s2 = sp.get_bspline_indices(kv2, ks_id=e2 + 2)
            s1 = sp.get_bspline_indices(kv1, ks_id=e1 + 2)
            for n2 in range(1, 4):
                for n1 in range(1, 4):
                    ix[n1 + (n2 - 1) * 3 - 1, e1 + (e2 - 1) * kv1.nel - 1] = \
                        s1[n1 - 1] + num1np * (s2[n2 - 1] - 1)
    return ix


def make_mesh(p, **args):
    """generate_mesh (Mesh.jl:94-248) for the FLAT topology."""
    m = Mesh()
    m.num1el, m.num2el = p.num1el, p.num2el
    m.numel = p.num1el * p.num2el
    m.num1np = p.num1el + p.poly
    m.num2np = p.num2el + p.poly
    m.numnp = m.num1np * m.num2np
    n1e, ne = p.num1el, m.numel
    n1, nn = m.num1np, m.numnp
    m.bdry_elems = {
        BOTTOM: list(range(1, n1e + 1)),
        RIGHT: list(range(n1e, ne + 1, n1e)),
        TOP: list(range(ne - n1e + 1, ne + 1)),
        LEFT: list(range(1, ne - n1e + 2, n1e)),
    }
    m.bdry_nodes = {
        BOTTOM: list(range(1, n1 + 1)),
        RIGHT: list(range(n1, nn + 1, n1)),
        TOP: list(range(nn - n1 + 1, nn + 1)),
        LEFT: list(range(1, nn - n1 + 2, n1)),
    }
    m.bdry_inner_nodes = {
        BOTTOM: [i + n1 for i in range(1, n1 + 1)],
        RIGHT: [i - 1 for i in range(n1, nn + 1, n1)],
        TOP: [i - n1 for i in range(nn - n1 + 1, nn + 1)],
        LEFT: [i + 1 for i in range(1, nn - n1 + 2, n1)],
    }
    m.crnr_nodes = {1: 1, 2: n1, 3: nn - n1 + 1, 4: nn}
    if p.scenario == F_PULL and p.num1el >= 18 and p.num2el >= 18:
        m.kv1 = sp.KnotVector(sp.get_fine_zs(p.num1el, p.poly), p.poly, sp.CLAMPED)
        m.kv2 = sp.KnotVector(sp.get_fine_zs(p.num2el, p.poly), p.poly, sp.CLAMPED)
    else:
        m.kv1 = sp.KnotVector(p.num1el, p.poly, sp.CLAMPED)
        m.kv2 = sp.KnotVector(p.num2el, p.poly, sp.CLAMPED)
    l1 = sp.LineBasis(m.kv1, p.gp1d)
    l2 = sp.LineBasis(m.kv2, p.gp1d)
    m.area_uel_ids, m.area_ufns = sp.area_basis(l1, l2)
    m.bdry_fns = {
        BOTTOM: sp.bdry_basis(l1, l2.zmin, True),
        RIGHT: sp.bdry_basis(l2, l1.zmax, False),
        TOP: sp.bdry_basis(l1, l2.zmax, True),
        LEFT: sp.bdry_basis(l2, l1.zmin, False),
    }
    m.crnr_fns = [sp.gp_basis_2d(l1.zmin, l2.zmin), sp.gp_basis_2d(l1.zmax, l2.zmin),
                  sp.gp_basis_2d(l1.zmin, l2.zmax), sp.gp_basis_2d(l1.zmax, l2.zmax)]
    m.crnr_elems = [1, n1e, ne - n1e + 1, ne]
    m.IX = construct_IX(m.kv1, m.kv2, n1)
    _generate_scenario(m, p, **args)
    return m


def _generate_scenario(m, p, **args):
    """generate_scenario (Mesh.jl:262-302)."""
    dofs, ndf, ID, dir_bcs, neu_bcs = _scenario_bc_info(m, p, **args)
    idx = 1
    for node in range(m.numnp):
        for d in range(ndf):
            if ID[d, node] != -1:
                ID[d, node] = idx
                idx += 1
            else:
                ID[d, node] = 0
    m.dofs, m.ndf, m.ID = dofs, ndf, ID
    m.nmdf = int(ID.max())
    m.ID_inv = [(node + 1, d + 1) for node in range(m.numnp) for d in range(ndf)
                if ID[d, node] != 0]
    # LM = reshape(ID[:, IX], ndf*NEN, numel)
    m.LM = ID[:, m.IX - 1].reshape(ndf * 9, m.numel, order="F")
    m.inh_dir_bcs, m.inh_neu_bcs = dir_bcs, neu_bcs


def retarget_pull(m, p, pull_el, pull_speed):
    """The F_PULL numbering and Dirichlet lists of mesh m rebuilt, in place,
    with the tether on element pull_el (1-based): generate_scenario
    (Mesh.jl:262-302) with the pulled element of Bc.jl:227-243 replaced
    (extension; the reference always pulls ceil(numel / 2))."""
    _generate_scenario(m, p, pull_speed=pull_speed, pull_el=pull_el)


def _scenario_bc_info(m, p, **args):
    """get_scenario_bc_info and the per-scenario functions (Bc.jl:15-393)."""
    bn = m.bdry_nodes
    if p.scenario in (F_CAVI, F_COUE, F_POIS):
        dofs = {VX: 1, VY: 2, LAM: 3}
        ndf = 3
        ID = np.zeros((ndf, m.numnp), dtype=np.int64)
        dirs, neus = [], []

        def mark(d, i):
            ID[dofs[d] - 1, i - 1] = -1
        if p.scenario == F_CAVI:                       # Bc.jl:58-106
            for b in (BOTTOM, TOP, LEFT, RIGHT):
                for i in bn[b]:
                    mark(VX, i)
                    mark(VY, i)
                    if b == TOP and i not in (m.crnr_nodes[3], m.crnr_nodes[4]):
                        dirs.append((VX, i, 1.0))
            c = m.numnp // 2 + 1
            mark(LAM, c)
            dirs.append((LAM, c, 0.0))
        elif p.scenario == F_COUE:                     # Bc.jl:122-157
            for i in bn[BOTTOM]:
                mark(VX, i)
                mark(VY, i)
            for i in bn[TOP]:
                mark(VX, i)
                mark(VY, i)
                dirs.append((VX, i, 3.0))
            for i in bn[LEFT] + bn[RIGHT]:
                mark(VY, i)
            neus = [(LEFT, STRETCH, 4.0), (RIGHT, STRETCH, 4.0)]
        else:                                          # Bc.jl:286-316
            for b in (TOP, BOTTOM):
                for i in bn[b]:
                    mark(VX, i)
                    mark(VY, i)
            for b in (LEFT, RIGHT):
                for i in bn[b]:
                    mark(VY, i)
            neus = [(LEFT, STRETCH, 4.0), (RIGHT, STRETCH, 8.0)]
        return dofs, ndf, ID, dirs, neus

    dofs = get_dofs(p.motion)
    ndf = len(dofs)
    ID = np.zeros((ndf, m.numnp), dtype=np.int64)
    dirs = []

    def mark(d, i):
        ID[dofs[d] - 1, i - 1] = -1

    if p.scenario == F_PULL:                           # Bc.jl:188-269
        speed = args["pull_speed"]
        for b in (BOTTOM, RIGHT, TOP, LEFT):
            for i in bn[b]:
                mark(VZ, i)
                if p.motion in (ALEV, ALEVB):
                    mark(VMX, i)
                    mark(VMY, i)
                    mark(VMZ, i)
        for b in (BOTTOM, RIGHT, TOP, LEFT):
            for i in m.bdry_inner_nodes[b]:
                mark(VZ, i)
                if p.motion == ALEVB:
                    mark(VMZ, i)
        # `pull_el` moves the tether (extension, no reference counterpart)
        pe = int(args.get("pull_el", get_pull_el_id(m.numel)))
        for node in m.IX[:, pe - 1]:
            node = int(node)
            mark(VX, node)
            mark(VY, node)
            mark(VZ, node)
            dirs.append((VZ, node, speed))
            if p.motion != LAG:
                mark(VMX, node)
                mark(VMY, node)
                mark(VMZ, node)
                dirs.append((VMZ, node, speed))
        for b in (BOTTOM, RIGHT, TOP, LEFT):
            lst = bn[b]
            c = lst[len(lst) // 2]            # bdry_nodes[b][floor(end/2)+1]
            mark(VX, c)
            mark(VY, c)
            if p.motion != LAG:
                mark(VMX, c)
                mark(VMY, c)
        lv = p.kb / 4
        neus = [(LEFT, STRETCH, lv), (RIGHT, STRETCH, lv), (TOP, STRETCH, lv),
                (BOTTOM, STRETCH, lv)]
        assert p.pn == 0.0
        return dofs, ndf, ID, dirs, neus

    if p.scenario == F_BEND:                           # Bc.jl:347-393
        for b in (TOP, BOTTOM):
            for i in bn[b]:
                mark(VY, i)
                if p.motion in (ALEV, ALEVB):
                    mark(VMY, i)
        for i in bn[LEFT]:
            mark(VX, i)
            mark(VY, i)
            mark(VZ, i)
            if p.motion in (ALEV, ALEVB):
                mark(VMX, i)
                mark(VMY, i)
                mark(VMZ, i)
        for i in bn[RIGHT]:
            mark(VZ, i)
            if p.motion in (ALEV, ALEVB):
                mark(VMZ, i)
        neus = [(LEFT, MOMENT, args["bend_mf"]), (RIGHT, MOMENT, args["bend_mf"])]
        return dofs, ndf, ID, dirs, neus
    raise AssertionError("scenario not implemented")


def prepare_input(p, mesh, **args):
    """prepare_input (Input.jl:22-121), non-random scenarios only.

    Positions are the collocation fit of x = L(ζ1 - 1/2), y = L(ζ2 - 1/2),
    z = 0; λ = kb/4; then the inhomogeneous Dirichlet values.  F_BEND
    (Input.jl:44-75): x = L ζ1, y = L ζ2, seeded uniform velocities.
    """
    xms = np.zeros((mesh.numnp, 3))
    cps = np.zeros((mesh.numnp, mesh.ndf))
    L = p.length
    if p.scenario == F_BEND:
        # Input.jl:44-75, with Julia's unseeded rand() replaced by
        # numpy's default_rng(bend_seed) drawn in the same order
        rng = np.random.default_rng(int(args.get("bend_seed", 0)))
        nn, ne = mesh.numnp, mesh.numel
        xms[:, 0] = sp.get_2d_bspline_cps(mesh.kv1, mesh.kv2, lambda a, b: L * a)
        xms[:, 1] = sp.get_2d_bspline_cps(mesh.kv1, mesh.kv2, lambda a, b: L * b)
        cols = [VX, VY, VZ] + ([] if p.motion == LAG else [VMX, VMY, VMZ])
        for u in cols:
            shift = 0.0 if u in (VZ, VMZ) else 0.5
            cps[:, mesh.dofs[u] - 1] = (rng.random(nn) - shift) / ne
        for nodes in mesh.bdry_nodes.values():
            cps[np.asarray(nodes) - 1, :] = 0.0
        dts = args["dts"] if "dts" in args else args["Δts"]
        cps[:, mesh.dofs[LAM] - 1] = p.kb / 4 / L ** 2 * (dts[0] / args["bend_tm"]) ** 2
    else:
        xms[:, 0] = sp.get_2d_bspline_cps(mesh.kv1, mesh.kv2, lambda a, b: L * (a - 0.5))
        xms[:, 1] = sp.get_2d_bspline_cps(mesh.kv1, mesh.kv2, lambda a, b: L * (b - 0.5))
        cps[:, mesh.dofs[LAM] - 1] = p.kb / 4
    for (unk, node, val) in mesh.inh_dir_bcs:
        cps[node - 1, mesh.dofs[unk] - 1] = val
        if p.scenario == F_PULL and p.motion == EUL:
            cps[node - 1, mesh.dofs[VMZ] - 1] = val
    return xms, cps
