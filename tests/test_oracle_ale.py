"""Oracle pins for the mesh-motion (EUL / ALE) physics and the F_PULL boundary sets.

* Dirichlet sets of F_PULL (src/input/Bc.jl:188-269) rebuilt here from grid
  coordinates -- boundary ring, inner ring, the 3 x 3 node block of element
  ceil(numel/2), the edge-centre nodes -- and compared entry by entry with the
  oracle's ID for every motion; the unknown counts are checked against
  tests/golden/mesh_counts.json (hand-counted 4 x 4 rows and the benchmark
  sizes).
* Flat-patch closed forms of the mesh-motion rows of the element residual
  (src/analysis/FiniteElement.jl:298-313, GeoDynStress.jl:156-159): on the flat
  patch x = L(zeta - 1/2) the geometry is a_alpha = L e_alpha, J = L^2, n = e_z,
  b = 0, so
    ALEV/ALEVB, v^m = G x + c (G with G_12 != G_21):
      sigma_m^{ab} = zv (G_ab + G_ba) / L^2,
      rm_x(a) = zv L (2 G_11 I1_a + (G_12 + G_21) I2_a),
      rm_y(a) = zv L ((G_12 + G_21) I1_a + 2 G_22 I2_a),
      rm_z(a) = -L^2 P I0_a                       (constant p^m = P),
      rp(a)   = -L^2 (c_z - v_z) I0_a             (the Dohrmann--Bochev term
                                                   cancels -adb p^m / zv for
                                                   constant p^m);
    EUL, constant v^m = c and v = d:
      rm(a) = am L^2 (c_x, c_y, c_z - d_z) I0_a;
    every motion, v = G_v x: rlambda(a) = L^2 (G_v11 + G_v22) I0_a,
  with I0_a, Ialpha_a the Gauss sums of w N_a and w N_a,alpha over the
  element (basis records pinned by the reference's own unit vectors).  A
  dropped symmetrisation, a transposed G, a wrong sign or the full v^m - v
  in the EUL row fails these.
* The ALEV bending steady state approaches the analytic arc of
  src/Output.jl:141-176 (H = 1/2, lambda = 1/4, K = 0) with an error that
  falls under mesh refinement.
"""
import json
import os

import numpy as np
import pytest

from oracle import fem, mesh as om

GOLD = os.path.join(os.path.dirname(__file__), "golden", "mesh_counts.json")
MOTION = {"LAG": om.LAG, "EUL": om.EUL, "ALEV": om.ALEV, "ALEVB": om.ALEVB}
UNK = {"vx": om.VX, "vy": om.VY, "vz": om.VZ, "vmx": om.VMX, "vmy": om.VMY, "vmz": om.VMZ,
       "lambda": om.LAM, "pm": om.PM}


def _fixed_from_geometry(n, motion):
    """Prescribed (unknown -> boolean grid [i2, i1]) from the F_PULL rules."""
    m = n + 2
    i1, i2 = np.meshgrid(np.arange(1, m + 1), np.arange(1, m + 1))
    ring = (i1 == 1) | (i1 == m) | (i2 == 1) | (i2 == m)
    inner = (i1 == 2) | (i1 == m - 1) | (i2 == 2) | (i2 == m - 1)
    pe = -(-n * n // 2)
    e1, e2 = (pe - 1) % n + 1, (pe - 1) // n + 1
    pull = (i1 >= e1) & (i1 <= e1 + 2) & (i2 >= e2) & (i2 <= e2 + 2)
    c = m // 2 + 1
    centre = (((i2 == 1) | (i2 == m)) & (i1 == c)) | (((i1 == 1) | (i1 == m)) & (i2 == c))
    ale = motion in (om.ALEV, om.ALEVB)
    fixed = {
        om.VX: pull | centre,
        om.VY: pull | centre,
        om.VZ: ring | inner | pull,
        om.LAM: np.zeros_like(ring),
    }
    if motion != om.LAG:
        fixed[om.VMX] = pull | centre | (ring if ale else False)
        fixed[om.VMY] = pull | centre | (ring if ale else False)
        fixed[om.VMZ] = pull | (ring if ale else False) | (inner if motion == om.ALEVB else False)
    if ale:
        fixed[om.PM] = np.zeros_like(ring)
    return fixed


@pytest.mark.parametrize("motion", [om.LAG, om.EUL, om.ALEV, om.ALEVB])
@pytest.mark.parametrize("n", [4, 5, 8, 17, 20])
def test_pull_dirichlet_sets_match_grid_rules(n, motion):
    p = om.Params(motion=motion, num1el=n, num2el=n, length=1.0)
    mesh = om.make_mesh(p, pull_speed=0.5)
    fixed = _fixed_from_geometry(n, motion)
    assert set(fixed) == set(mesh.dofs)
    for unk, col in mesh.dofs.items():
        got = (mesh.ID[col - 1, :] == 0).reshape(n + 2, n + 2)  # node = i1 + m (i2 - 1)
        assert np.array_equal(got, fixed[unk]), (n, motion, unk)
    # the prescribed values: v_z (and v^m_z off LAG) = pull speed on the 9 pulled nodes
    pulled = {(u, node) for (u, node, val) in mesh.inh_dir_bcs if val == 0.5}
    assert len(pulled) == (9 if motion == om.LAG else 18)


def test_hand_counted_unknowns_4x4():
    gold = json.load(open(GOLD))["hand_counted_4x4"]
    for name, row in gold.items():
        p = om.Params(motion=MOTION[name], num1el=4, num2el=4, length=1.0)
        mesh = om.make_mesh(p, pull_speed=0.5)
        assert mesh.nmdf == row["nmdf"], name
        for unk, cnt in row["by_unknown"].items():
            assert int(np.count_nonzero(mesh.ID[mesh.dofs[UNK[unk]] - 1, :])) == cnt, (name, unk)


@pytest.mark.parametrize("row", json.load(open(GOLD))["config_sizes"],
                         ids=lambda r: f"{r['n']}-{r['motion']}")
def test_unknown_counts_at_config_sizes(row):
    n = row["n"]
    p = om.Params(motion=MOTION[row["motion"]], num1el=n, num2el=n, length=1.0)
    mesh = om.make_mesh(p, pull_speed=0.5)
    assert mesh.nmdf == row["nmdf"]
    fixed = _fixed_from_geometry(n, p.motion)
    assert sum(int((~f).sum()) for f in fixed.values()) == row["nmdf"]


# ------------------------------------------------- flat-patch closed forms

def _flat(motion, n=5, L=1.3, **kw):
    p = om.Params(motion=motion, num1el=n, num2el=n, length=L, **kw)
    o, xms, cps = fem.build(p, pull_speed=0.0)
    return p, o, xms, np.zeros_like(cps)


def _gauss_sums(mesh, el):
    """(I0[9], I1[9], I2[9]): sums over the element's Gauss points of w N_a, w N_a,1, w N_a,2."""
    rec = mesh.area_ufns[mesh.area_uel_ids[el] - 1]          # [gp, 55]
    w = rec[:, 0:1]
    return (w * rec[:, 1:10]).sum(0), (w * rec[:, 10:19]).sum(0), (w * rec[:, 19:28]).sum(0)


def _rows(r_el, mesh, unk):
    return r_el[mesh.dofs[unk] - 1::mesh.ndf]


@pytest.mark.parametrize("motion", [om.ALEV, om.ALEVB])
def test_ale_mesh_motion_rows_closed_form(motion):
    p, o, xms, cps = _flat(motion, zv=0.7, kb=1.9, kg=-0.3)
    m = o.mesh
    G = np.array([[0.31, -0.47, 0.0], [0.83, -0.12, 0.0], [0.0, 0.0, 0.0]])
    c = np.array([0.05, -0.02, 0.23])
    v = np.array([0.0, 0.0, 0.08])                          # v only along the normal
    P = 0.37
    vm = xms @ G.T + c                                       # control points of G x + c
    for k, u in enumerate((om.VMX, om.VMY, om.VMZ)):
        cps[:, m.dofs[u] - 1] = vm[:, k]
    cps[:, m.dofs[om.VZ] - 1] = v[2]
    cps[:, m.dofs[om.PM] - 1] = P
    r_el, _ = fem.area_elements(o, xms, cps, 1e-3)
    L, zv = p.length, p.zv
    s12 = G[0, 1] + G[1, 0]
    for el in range(m.numel):
        I0, I1, I2 = _gauss_sums(m, el)
        scale = zv * L
        np.testing.assert_allclose(_rows(r_el[el], m, om.VMX), scale * (2 * G[0, 0] * I1 + s12 * I2),
                                   rtol=0, atol=1e-13)
        np.testing.assert_allclose(_rows(r_el[el], m, om.VMY), scale * (s12 * I1 + 2 * G[1, 1] * I2),
                                   rtol=0, atol=1e-13)
        np.testing.assert_allclose(_rows(r_el[el], m, om.VMZ), -L * L * P * I0, rtol=0, atol=1e-13)
        np.testing.assert_allclose(_rows(r_el[el], m, om.PM), -L * L * (c[2] - v[2]) * I0,
                                   rtol=0, atol=1e-13)


def test_eul_mesh_motion_row_closed_form():
    p, o, xms, cps = _flat(om.EUL, am=1.7)
    m = o.mesh
    c = np.array([0.21, -0.34, 0.55])
    d = np.array([0.4, 0.9, -0.15])
    for k, u in enumerate((om.VMX, om.VMY, om.VMZ)):
        cps[:, m.dofs[u] - 1] = c[k]
    for k, u in enumerate((om.VX, om.VY, om.VZ)):
        cps[:, m.dofs[u] - 1] = d[k]
    r_el, _ = fem.area_elements(o, xms, cps, 1e-3)
    L = p.length
    want = p.am * L * L * np.array([c[0], c[1], c[2] - d[2]])
    for el in range(m.numel):
        I0, _, _ = _gauss_sums(m, el)
        for k, u in enumerate((om.VMX, om.VMY, om.VMZ)):
            np.testing.assert_allclose(_rows(r_el[el], m, u), want[k] * I0, rtol=0, atol=1e-13)


@pytest.mark.parametrize("motion", [om.LAG, om.EUL, om.ALEV, om.ALEVB])
def test_tension_row_divergence_closed_form(motion):
    p, o, xms, cps = _flat(motion)
    m = o.mesh
    Gv = np.array([[0.6, 0.25, 0.0], [-0.4, 0.15, 0.0], [0.3, -0.2, 0.0]])
    v = xms @ Gv.T
    for k, u in enumerate((om.VX, om.VY, om.VZ)):
        cps[:, m.dofs[u] - 1] = v[:, k]
    cps[:, m.dofs[om.LAM] - 1] = 0.9                        # DB term cancels for constant lambda
    r_el, _ = fem.area_elements(o, xms, cps, 1e-3)
    L = p.length
    for el in range(m.numel):
        I0, _, _ = _gauss_sums(m, el)
        np.testing.assert_allclose(_rows(r_el[el], m, om.LAM), L * L * (Gv[0, 0] + Gv[1, 1]) * I0,
                                   rtol=0, atol=1e-13)


# ------------------------------------------------- ALEV bending steady state

def _alev_bend_errors(n):
    from test_oracle_bend import gauss_fields
    p = om.Params(motion=om.ALEV, scenario=om.F_BEND, num1el=n, num2el=n, length=1.0,
                  output=False)
    dts = [0.5] * 16 + [1.0] * 8 + [4.0] * 8 + [16.0] * 10 + [64.0] * 10 + [256.0] * 10
    o, xms, cps = fem.build(p, bend_tm=2.0, bend_mf=0.5, dts=dts, bend_seed=0)
    t = 0.0
    for dt in dts:
        t += dt
        fem.update_xms(p.motion, xms, cps, dt, o.mesh.dofs)
        fem.time_step(o, xms, cps, t, dt, bend_tm=2.0)
    f = gauss_fields(o, xms, cps)
    H = np.array([x[1] for x in f])
    K = np.array([x[2] for x in f])
    lam = np.array([x[3] for x in f])
    return np.abs(H - 0.5).max(), np.abs(lam - 0.25).max(), np.abs(K).max()


@pytest.mark.slow
def test_alev_bending_reaches_the_arc_under_refinement():
    eH4, el4, eK4 = _alev_bend_errors(4)
    eH6, el6, eK6 = _alev_bend_errors(6)
    assert eK4 < 1e-12 and eK6 < 1e-12
    assert eH6 < 0.02 and el6 < 0.02
    assert eH6 < 0.5 * eH4 and el6 < 0.5 * el4
