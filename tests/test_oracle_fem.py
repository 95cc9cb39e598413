"""Oracle element physics and Newton loop, pinned by closed forms and invariants.

* GeoDynStress closed forms of test/analysis/geo_dyn_stress_tests.jl
  (paraboloid geometry; Couette and Poiseuille stress; total area), plus the
  textbook mean/Gaussian curvature of a graph z = f(x, y), which fixes the sign
  and index conventions of b, H and K independently of the code's formula.
* Flat F_PULL state with zero pull speed is an equilibrium: r = 0.
* The complex-step tangent equals central finite differences of the residual
  (area elements; the motion column includes the dt-scaled position term).
* Newton converges quadratically.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import fem, mesh as om, spline as sp


def _gds_all(o, xms, cps):
    """Yield (el, gp, out[26]) over all area Gauss points."""
    m = o.mesh
    out = np.zeros(26)
    for el in range(m.numel):
        nodes = m.IX[:, el] - 1
        X9 = np.asfortranarray(xms[nodes, :])
        C9 = np.asfortranarray(cps[nodes, :])
        for gp in range(9):
            rec = np.ascontiguousarray(m.area_ufns[m.area_uel_ids[el] - 1, gp])
            fem.lib().orc_gds_point(ctypes.byref(o.s), fem._dp(rec), fem._dp(X9),
                                    fem._dp(C9), fem._dp(out))
            yield el, gp, rec, out.copy()


def _flat_xy(m):
    x_c = sp.get_1d_bspline_cps(m.kv1, lambda z: 4 * z + 2)
    y_c = sp.get_1d_bspline_cps(m.kv2, lambda z: 2 * z - 3)
    xms = np.zeros((m.numnp, 3))
    for node in range(m.numnp):
        xms[node, 0] = x_c[node % m.num1np]
        xms[node, 1] = y_c[node // m.num1np]
    return xms


def test_gds_curved_paraboloid():
    p = om.Params(motion=om.STATIC, scenario=om.F_CAVI, num1el=2, num2el=3)
    m = om.make_mesh(p)
    o = fem.OracleMesh(m, p)
    A, B = 2.2, 1.7
    xms = _flat_xy(m)
    xms[:, 2] = sp.get_2d_bspline_cps(m.kv1, m.kv2,
                                      lambda a, b: A * (4 * a - 2) ** 2 + B * (2 * b - 1) ** 2)
    cps = np.zeros((m.numnp, m.ndf))
    eps = np.finfo(float).eps
    for el, gp, rec, o_ in _gds_all(o, xms, cps):
        x, y, z = o_[0:3]
        a1, a2 = o_[3:6], o_[6:9]
        aup = o_[9:13].reshape(2, 2)
        J, H, K = o_[13], o_[14], o_[15]
        b = o_[19:23].reshape(2, 2)
        assert z == pytest.approx(A * (x - 4) ** 2 + B * (y + 2) ** 2, rel=1e-12)
        aa = np.array([[4., 0.], [0., 2.], [8 * A * (x - 4), 4 * B * (y + 2)]])
        assert np.linalg.norm(np.stack([a1, a2], 1) - aa) < 4e2 * eps
        aab = np.stack([a1, a2], 1).T @ np.stack([a1, a2], 1)
        assert np.linalg.norm(aup @ aab - np.eye(2)) < 1e2 * eps
        assert abs(np.trace(aup @ b) / 2 - H) < 1e1 * eps
        assert abs(np.linalg.det(aup @ b) - K) < 1e1 * eps
        # textbook curvatures of the graph z = f(x, y), upward normal
        fx, fy, fxx, fyy = 2 * A * (x - 4), 2 * B * (y + 2), 2 * A, 2 * B
        g = 1 + fx * fx + fy * fy
        H_ex = ((1 + fy * fy) * fxx + (1 + fx * fx) * fyy) / (2 * g ** 1.5)
        K_ex = fxx * fyy / g ** 2
        assert H == pytest.approx(H_ex, rel=1e-12)
        assert K == pytest.approx(K_ex, rel=1e-12)
        assert J == pytest.approx(8 * math.sqrt(g), rel=1e-12)


def test_gds_flat_couette_poiseuille_area():
    p = om.Params(motion=om.STATIC, scenario=om.F_CAVI, num1el=3, num2el=2)
    m = om.make_mesh(p)
    o = fem.OracleMesh(m, p)
    xms = _flat_xy(m)
    w, U = 12.0, 16.0
    c_vs = sp.get_1d_bspline_cps(m.kv2, lambda z: 2 * w * z)
    p_vs = sp.get_1d_bspline_cps(m.kv1, lambda z: -4 * U * (z * z - z))
    cps_c = np.zeros((m.numnp, m.ndf))
    cps_p = np.zeros((m.numnp, m.ndf))
    for node in range(m.numnp):
        cps_c[node, 0] = c_vs[node // m.num1np]
        cps_p[node, 1] = p_vs[node % m.num1np]
    eps = np.finfo(float).eps
    area = 0.0
    for (el, gp, rec, oc), (_, _, _, op) in zip(_gds_all(o, xms, cps_c), _gds_all(o, xms, cps_p)):
        area += op[13] * rec[0]
        x = oc[0]
        assert np.linalg.norm(oc[16:19] - p.zv * np.array([0., 0., 2 * w / 8])) < 2e2 * eps
        assert np.linalg.norm(op[16:19] - p.zv * np.array([0., 0., U * (4 - x) / 8])) < 3e2 * eps
        assert op[14] == 0.0 and op[15] == 0.0           # flat: H = K = 0
    assert area == pytest.approx(8.0, rel=1e-12)


@pytest.mark.parametrize("motion", [om.LAG, om.EUL, om.ALEV, om.ALEVB])
@pytest.mark.parametrize("n", [8, 19])
def test_flat_equilibrium(motion, n):
    p = om.Params(motion=motion, num1el=n, num2el=n, length=1.0)
    o, xms, cps = fem.build(p, pull_speed=0.0)
    r, K = fem.calc_r_K(o, xms, cps, 0.0, 1e-3)
    assert np.abs(r).max() < 1e-14
    assert K.shape == (o.mesh.nmdf, o.mesh.nmdf)


def _perturbed_state(motion, n=6, seed=0):
    p = om.Params(motion=motion, num1el=n, num2el=n, length=1.0)
    o, xms, cps = fem.build(p, pull_speed=0.3)
    rng = np.random.default_rng(seed)
    xms = xms + rng.normal(0, 0.02, xms.shape)
    cps = cps + rng.normal(0, 0.1, cps.shape)
    return o, xms, cps


@pytest.mark.parametrize("motion", [om.LAG, om.EUL, om.ALEVB])
def test_tangent_matches_finite_differences(motion):
    o, xms, cps = _perturbed_state(motion)
    m, dt = o.mesh, 0.05
    el = 7
    r_el, K_el = fem.area_elements(o, xms, cps, dt, el, el + 1)
    K_el = K_el[0]
    mm = om.m_motion_order(o.p.motion, m.dofs)
    nodes = m.IX[:, el] - 1
    ne = 9 * m.ndf
    K_fd = np.zeros((ne, ne))
    h = 1e-6

    def res(x, c):
        out = np.zeros(ne)
        fem.lib().orc_area_residual(ctypes.byref(o.s), fem._ip(o.uel), fem._dp(o.ufns), el,
                                    fem._dp(np.asfortranarray(x)), fem._dp(np.asfortranarray(c)),
                                    fem._dp(out))
        return out
    for a in range(9):
        for d in range(m.ndf):
            if m.ID[d, nodes[a]] == 0:
                continue
            col = d + m.ndf * a
            cp, cm = cps.copy(), cps.copy()
            cp[nodes[a], d] += h
            cm[nodes[a], d] -= h
            K_fd[:, col] = (res(xms, cp) - res(xms, cm)) / (2 * h)
            if d + 1 in mm:
                i = mm.index(d + 1)
                xp, xm = xms.copy(), xms.copy()
                xp[nodes[a], i] += h
                xm[nodes[a], i] -= h
                K_fd[:, col] += dt * (res(xp, cps) - res(xm, cps)) / (2 * h)
    assert np.allclose(r_el[0], res(xms, cps), rtol=0, atol=1e-13)
    rel = np.linalg.norm(K_el - K_fd) / np.linalg.norm(K_el)
    assert rel < 1e-7


def test_newton_quadratic_convergence():
    p = om.Params(motion=om.ALEVB, num1el=8, num2el=8, length=1.0, enr=1e-10)
    o, xms, cps = fem.build(p, pull_speed=0.5)
    fem.update_xms(p.motion, xms, cps, 1e-3, o.mesh.dofs)
    eps = fem.time_step(o, xms, cps, 1e-3, 1e-3)
    assert eps[-1] < 1e-10
    assert len(eps) >= 2
    # quadratic: e_{k+1} / e_k^2 stays bounded while e_k falls by orders
    ratios = [eps[k + 1] / eps[k] ** 2 for k in range(len(eps) - 1)]
    assert max(ratios[:-1]) < 1e2      # the last step sits at the rounding floor
    assert eps[2] < 1e-4 * eps[1]


def test_newton_failure_asserts():
    p = om.Params(motion=om.LAG, num1el=4, num2el=4, length=1.0, enr=1e-30)
    o, xms, cps = fem.build(p, pull_speed=0.5)
    with pytest.raises(AssertionError, match="did not reach Newton--Raphson tolerance"):
        fem.time_step(o, xms, cps, 1e-3, 1e-3, max_iter=3)
