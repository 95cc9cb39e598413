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


@pytest.mark.parametrize("ordering", ["band", "nd"])
def test_zero_pivot_tangent_converges_with_static_pivoting(ordering):
    print(f"static pivoting: 96x96 EUL request, {ordering} ordering", flush=True)
    sp = serving_specs(24)[12]
    assert int(sp.motion) == int(mafem.EUL) and sp.n == 96
    pt = Patch(sp)
    try:
        pt.dev.set_solver("direct", precision=64)
        pt.dev.set_ordering(ordering)
        pt.set_pull(sp.speed_at(0))
        pt.dev.predict(1e-3)
        rc, eps = pt.dev.time_step(1e-3, 1e-3)
        st = pt.dev.stats()
    finally:
        pt.close()
    print(f"  eps {eps} perturbations {st['pivot_perturbations']}", flush=True)
    assert rc == 0 and len(eps) == 3
    assert eps[-1] < 1e-10 and abs(eps[1] - 3.76e-5) < 1e-7
    if ordering == "band":  # its 64-row diagonal tiles meet numerically zero pivots
        assert st["pivot_perturbations"] > 0


@pytest.mark.parametrize("precision", [32, 64])
def test_batched_zero_pivot_member_converges_with_static_pivoting(precision):
    """The same request as a member of a batch: its fresh FP64 factorization
    (after the FP32 one moved it to FP64) fails, the member refactors with
    static pivoting, GMRES runs the cycles the replaced pivots need (about 100
    Arnoldi steps) and the member takes the reference's 3 Newton iterations
    with no inexact correction; the other member keeps the result it has
    alone."""
    print(f"static pivoting: batched 96x96 EUL request, precision {precision}", flush=True)
    specs = serving_specs(24)
    pts = [Patch(specs[12]), Patch(specs[0])]
    solo = Patch(specs[0])
    try:
        for pt in pts + [solo]:
            pt.set_pull(pt.spec.speed_at(0))
        b = mafem.Batch([pt.dev for pt in pts], precision=precision)
        try:
            it, eps, st = b.time_step(np.full(2, 1e-3), 1e-3, predict=True)
            hard = b.stats()["hard_failures"]
        finally:
            b.close()
        b1 = mafem.Batch([solo.dev], precision=precision)
        try:
            it1, eps1, st1 = b1.time_step(np.full(1, 1e-3), 1e-3, predict=True)
        finally:
            b1.close()
    finally:
        for pt in pts + [solo]:
            pt.close()
    print(f"  iters {list(it)} eps {eps[0]}", flush=True)
    assert list(st) == [0, 0] and it[0] == 3 and hard == 0
    assert eps[0][-1] < 1e-10 and abs(eps[0][1] - 3.76e-5) < 1e-7
    assert it[1] == it1[0] and np.allclose(eps[1], eps1[0], rtol=1e-6, atol=1e-13)


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
