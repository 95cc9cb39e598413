"""Scenarios beside F_PULL on the device, through the C ABI: F_BEND and the
flat 2-D STATIC scenarios F_CAVI, F_COUE, F_POIS.

F_BEND (MOMENT edges with the time ramp min(t / bend_tm, 1),
FiniteElement.jl:377-380; boundary conditions Bc.jl:347-393): the
reference's documented pure-bending example run with the host mirror's
`solve` agrees with the oracle step by step (state <= 1e-8 relative, same
Newton counts) and ends on the analytic arc (H = 1/2, λ = 1/4 within the
2 x 2 discretization error, K = 0).
"""
import numpy as np
import pytest

import mafem
from oracle import fem, mesh as om

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("motion", [mafem.LAG, mafem.ALEV])
def test_documented_bending_matches_oracle(motion):
    steps = 16
    args = dict(bend_tm=2.0, bend_mf=0.5, t0=0.0, t0_id=0, dts=[0.5] * steps, bend_seed=0)
    p = mafem.Params(motion=motion, scenario=mafem.F_BEND, num1el=2, num2el=2, length=1.0,
                     output=False)
    mesh = mafem.make_mesh(p, **args)
    xms, cps = mafem.prepare_input(p, mesh, **args)
    po = om.Params(motion=int(motion), scenario=om.F_BEND, num1el=2, num2el=2, length=1.0,
                   output=False)
    o, xo, co = fem.build(po, **args)
    assert np.array_equal(xms, xo) and np.array_equal(cps, co)
    print(f"bending motion={motion.name}", flush=True)
    hist = mafem.run_analysis(mesh, xms, cps, p, **args)
    times = np.cumsum(args["dts"])
    for k, dt in enumerate(args["dts"]):
        fem.update_xms(po.motion, xo, co, dt, o.mesh.dofs)
        eps = fem.time_step(o, xo, co, float(times[k]), dt, bend_tm=2.0)
        assert len(hist[k]) == len(eps), (k, hist[k], eps)
    assert np.linalg.norm(xms - xo) <= 1e-8 * np.linalg.norm(xo)
    assert np.linalg.norm(cps - co) <= 1e-8 * np.linalg.norm(co)
    if motion == mafem.LAG:
        from test_oracle_bend import gauss_fields
        fields = gauss_fields(o, xms, cps)
        assert max(abs(f[1] - 0.5) for f in fields) < 0.02
        assert max(abs(f[3] - 0.25) for f in fields) < 0.015
        assert max(abs(f[2]) for f in fields) < 1e-12


@pytest.mark.parametrize("scen,n1,n2,L,zv", [(mafem.F_COUE, 6, 7, 2.0, 0.5),
                                             (mafem.F_POIS, 7, 6, 1.0, 1.0),
                                             (mafem.F_CAVI, 6, 6, 1.0, 1.0),
                                             (mafem.F_CAVI, 17, 16, 1.0, 0.3)])
def test_flat_2d_scenarios_match_oracle(scen, n1, n2, L, zv):
    """STATIC 2-D scenarios (vx, vy, λ): assembled r and K <= 1e-12 and the
    steady state after two steps <= 1e-8 vs the oracle with equal Newton
    counts; Couette and Poiseuille also against their closed forms."""
    print(f"2-D scenario {scen.name} {n1}x{n2}", flush=True)
    p = mafem.Params(motion=mafem.STATIC, scenario=scen, num1el=n1, num2el=n2, length=L, zv=zv,
                     output=False)
    mesh = mafem.make_mesh(p)
    xms, cps = mafem.prepare_input(p, mesh)
    po = om.Params(motion=om.STATIC, scenario=int(scen), num1el=n1, num2el=n2, length=L, zv=zv,
                   output=False)
    o, xo, co = fem.build(po)
    assert np.allclose(xms, xo, rtol=0, atol=1e-14 * L) and np.array_equal(cps, co)
    xo, co = xms.copy(), cps.copy()   # the same initial state for both
    r, K = mafem.calc_r_K(mesh, xms, cps, 0.1, 0.1, p)
    r_o, K_o = fem.calc_r_K(o, xo, co, 0.1, 0.1)
    assert np.linalg.norm(r - r_o) <= 1e-12 * np.linalg.norm(r_o)
    assert np.linalg.norm((K - K_o).toarray()) <= 1e-12 * np.linalg.norm(K_o.toarray())
    hist = mafem.run_analysis(mesh, xms, cps, p, dts=[0.1, 0.1])
    for k in range(2):
        eps = fem.time_step(o, xo, co, 0.1 * (k + 1), 0.1)
        assert len(hist[k]) == len(eps)
    assert np.linalg.norm(cps - co) <= 1e-8 * np.linalg.norm(co)
    xout, uout = mafem.generate_output(mesh, xms, cps, p)
    x, y = xout[..., 0], xout[..., 1]
    if scen == mafem.F_COUE:
        assert np.abs(uout[..., 0] - 3 * (y / L + 0.5)).max() < 1e-10
        assert np.abs(uout[..., 2] - 4.0).max() < 1e-10
    elif scen == mafem.F_POIS:
        g = 4.0 / L
        assert np.abs(uout[..., 0] - g / (2 * zv) * (y + L / 2) * (L / 2 - y)).max() < 1e-10
        assert np.abs(uout[..., 2] - (6.0 + g * x)).max() < 1e-10
    else:
        assert np.abs(uout[..., 0] - uout[::-1, :, 0]).max() < 1e-10
