"""Static pivoting of the direct solve, through the C ABI.

The band LU pivots inside each 64-row diagonal tile.  On some EUL tangents a
tile has (numerically) zero pivots: the first factorization then gives a
failed solve, the context switches to static pivoting (pivots below
eps max|K| replaced by +-eps max|K|; the refined solve or FGMRES finishes on
the nearby factorization) and the Newton step converges.  The case is the
serving workload's 96x96 EUL request (seed 0xC5, request 12) whose first step
failed before; the reference algorithm (oracle with SuperLU) takes the same
3 Newton iterations on it (eps 5.8e-5, 3.8e-5, 6.2e-11; about 4 minutes on
the host, so not rerun here).
"""
import numpy as np
import pytest

import mafem
from mafem.patches import Patch, serving_specs

pytestmark = pytest.mark.gpu


def test_zero_pivot_tangent_converges_with_static_pivoting():
    print("static pivoting: 96x96 EUL request", flush=True)
    sp = serving_specs(24)[12]
    assert int(sp.motion) == int(mafem.EUL) and sp.n == 96
    pt = Patch(sp)
    try:
        pt.dev.set_solver("direct", precision=64)
        pt.set_pull(sp.speed_at(0))
        pt.dev.predict(1e-3)
        rc, eps = pt.dev.time_step(1e-3, 1e-3)
        st = pt.dev.stats()
    finally:
        pt.close()
    print(f"  eps {eps} perturbations {st['pivot_perturbations']}", flush=True)
    assert rc == 0 and len(eps) == 3
    assert eps[-1] < 1e-10 and abs(eps[1] - 3.76e-5) < 1e-7
    assert st["pivot_perturbations"] > 0


def test_well_posed_tangent_uses_no_perturbation():
    p = mafem.Params(motion=mafem.ALEVB, num1el=24, num2el=24, length=1.0, enr=1e-10,
                     output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.5)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.5)
    dev = mafem.DeviceMesh(mesh, p)
    try:
        dev.set_state(xms, cps)
        dev.predict(1e-3)
        rc, eps = dev.time_step(1e-3, 1e-3)
        assert rc == 0
        assert dev.stats()["pivot_perturbations"] == 0
    finally:
        dev.close()
