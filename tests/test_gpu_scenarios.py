"""Scenarios beside F_PULL on the device, through the C ABI.

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
