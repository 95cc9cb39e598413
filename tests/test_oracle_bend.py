"""Pure bending (F_BEND with MOMENT edges), pinned by its analytic solution.

The reference's documented example (docs/src/index.md:109-137: LAG, F_BEND,
2 x 2 elements, length 1, bend_tm 2, bend_mf = kb / (2 L) = 0.5, 16 steps of
Δt = 0.5) relaxes to a cylinder of mean curvature H = 1/2 (radius 1) under
tension λ = kb/4 = 0.25 (the exact values its analyze_output compares with,
src/Output.jl:141-176; check_params fixes the half-angle to 1/2,
src/input/Params.jl:99-100).  On 2 x 2 quadratic elements the oracle's end
state is within the discretization error of that solution (|H - 1/2| < 0.02,
|λ - 1/4| < 0.015, K = 0 to rounding) and does not depend on the random
initial velocities (Input.jl:52-60; seeded here).
"""
import ctypes

import numpy as np
import pytest

from oracle import fem, mesh as om


def run_bend(seed, steps=16):
    p = om.Params(motion=om.LAG, scenario=om.F_BEND, num1el=2, num2el=2, length=1.0,
                  output=False)
    args = dict(bend_tm=2.0, bend_mf=0.5, dts=[0.5] * steps, bend_seed=seed)
    o, xms, cps = fem.build(p, **args)
    its = []
    times = np.cumsum(args["dts"])
    for k, dt in enumerate(args["dts"]):
        fem.update_xms(p.motion, xms, cps, dt, o.mesh.dofs)
        its.append(len(fem.time_step(o, xms, cps, float(times[k]), dt, bend_tm=2.0)))
    return o, xms, cps, its


def gauss_fields(o, xms, cps):
    """(x, H, K, λ) at every area Gauss point."""
    m = o.mesh
    out = np.zeros(26)
    lam = m.dofs[om.LAM] - 1
    res = []
    for el in range(m.numel):
        nodes = m.IX[:, el] - 1
        X9 = np.asfortranarray(xms[nodes, :])
        C9 = np.asfortranarray(cps[nodes, :])
        for gp in range(9):
            rec = np.ascontiguousarray(m.area_ufns[m.area_uel_ids[el] - 1, gp])
            fem.lib().orc_gds_point(ctypes.byref(o.s), fem._dp(rec), fem._dp(X9), fem._dp(C9),
                                    fem._dp(out))
            res.append((out[0:3].copy(), out[14], out[15], rec[1:10] @ C9[:, lam]))
    return res


def test_documented_bending_reaches_the_arc():
    o, xms, cps, its = run_bend(0)
    assert its[-1] <= 4
    fields = gauss_fields(o, xms, cps)
    H = np.array([f[1] for f in fields])
    K = np.array([f[2] for f in fields])
    lam = np.array([f[3] for f in fields])
    assert np.abs(H - 0.5).max() < 0.02
    assert np.abs(K).max() < 1e-12
    assert np.abs(lam - 0.25).max() < 0.015
    # the left edge stays at the origin line; the arc bends out of the plane
    x = np.array([f[0] for f in fields])
    assert x[:, 2].max() < 0.0


def test_bending_end_state_does_not_depend_on_initial_velocities():
    _, x0, c0, _ = run_bend(0)
    _, x1, c1, _ = run_bend(1)
    assert np.allclose(x0, x1, rtol=0, atol=1e-10)
    assert np.allclose(c0, c1, rtol=0, atol=1e-10)


def test_bending_initial_state_layout():
    p = om.Params(motion=om.ALEV, scenario=om.F_BEND, num1el=3, num2el=3, length=2.0,
                  output=False)
    o, xms, cps = fem.build(p, bend_tm=2.0, bend_mf=0.25, dts=[0.5], bend_seed=3)
    m = o.mesh
    assert np.allclose(xms[:, 2], 0.0)
    assert xms[:, 0].min() == pytest.approx(0.0) and xms[:, 0].max() == pytest.approx(2.0)
    lam = cps[:, m.dofs[om.LAM] - 1]
    assert np.allclose(lam, 1.0 / 4 / 4.0 * (0.5 / 2.0) ** 2)
    for nodes in m.bdry_nodes.values():
        rows = np.asarray(nodes) - 1
        assert np.all(cps[rows][:, [m.dofs[u] - 1 for u in (om.VX, om.VY, om.VZ, om.VMX)]] == 0)
    inner = np.setdiff1d(np.arange(m.numnp), np.concatenate(
        [np.asarray(v) - 1 for v in m.bdry_nodes.values()]))
    vz = cps[inner, m.dofs[om.VZ] - 1]
    assert np.all((vz > 0) & (vz < 1.0 / m.numel))
