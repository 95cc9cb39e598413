"""Two-ended band LU (top part factored downward, bottom part upward, the
separator last) against the one-chain band LU on the same tangents.

The two layouts order the elimination differently, so their solutions agree
to the conditioning of K rather than bit for bit; both must solve the system
to the direct path's target (|K du + r| / |r| <= 1e-12), agree with each other
and with SuperLU, and the Newton trajectory through the two-ended solver must
keep the oracle's iteration counts.  Contexts read MAFEM_TWO_ENDED when they
are created.
"""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import mafem
from mafem.analysis import DeviceMesh

pytestmark = pytest.mark.gpu


def _state(motion, n, seed=0):
    p = mafem.Params(motion=motion, num1el=n, num2el=n, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.5)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.5)
    rng = np.random.default_rng(seed)
    xms[:, 2] += rng.normal(0, 1e-3, mesh.numnp)
    free = mesh.ID.T != 0
    cps += np.where(free, rng.normal(0, 1e-2, cps.shape), 0.0)
    return p, mesh, xms, cps


def _device(mesh, p, two_ended):
    old = os.environ.get("MAFEM_TWO_ENDED")
    os.environ["MAFEM_TWO_ENDED"] = "1" if two_ended else "0"
    try:
        dev = DeviceMesh(mesh, p)
        dev.set_ordering("band")  # these tests are about the band's two layouts
        return dev
    finally:
        if old is None:
            del os.environ["MAFEM_TWO_ENDED"]
        else:
            os.environ["MAFEM_TWO_ENDED"] = old


@pytest.mark.parametrize("motion,n,precision", [(mafem.ALEVB, 16, 64), (mafem.EUL, 20, 64),
                                                (mafem.LAG, 24, 64), (mafem.ALEVB, 40, 64),
                                                (mafem.ALEVB, 24, 32)])
def test_two_ended_solve_matches_one_chain(motion, n, precision):
    print(f"two-ended solve motion={motion.name} n={n} precision={precision}", flush=True)
    p, mesh, xms, cps = _state(motion, n)
    sols = {}
    for te in (False, True):
        dev = _device(mesh, p, te)
        if precision == 32:
            dev.set_solver("krylov", precision=32, rtol=1e-12)
        dev.set_state(xms, cps)
        r, K = dev.assemble(1e-3, 1e-3)
        du, rel = dev.solve(-r)
        st = dev.stats()
        assert st["factor_chains"] == (2 if te else 1), st
        assert rel <= 1e-12, (te, rel)
        assert np.linalg.norm(K @ du + r) / np.linalg.norm(r) <= 1e-11
        sols[te] = (du, K, r)
        dev.close()
    du1, K, r = sols[False]
    du2 = sols[True][0]
    assert np.linalg.norm(du2 - du1) / np.linalg.norm(du1) <= 1e-8
    ref = spla.spsolve(K.tocsc(), -r)
    assert np.linalg.norm(du2 - ref) / np.linalg.norm(ref) <= 1e-7


def test_two_ended_sweep_splits_agree():
    """Every workgroup split of the two-ended sweeps gives the same solution."""
    print("two-ended sweep splits ALEVB 40", flush=True)
    p, mesh, xms, cps = _state(mafem.ALEVB, 40, seed=3)
    outs = []
    for g in ("1", "2", "4"):
        # the setting fixes the split at creation and keeps the autotuner off
        # at the first solve
        os.environ["MAFEM_SWEEP_GROUPS"] = g
        try:
            dev = _device(mesh, p, True)
            dev.set_state(xms, cps)
            r, _ = dev.assemble(1e-3, 1e-3, want_K=False)
            du, rel = dev.solve(-r)
        finally:
            del os.environ["MAFEM_SWEEP_GROUPS"]
        assert dev.stats()["sweep_groups"] == int(g)
        assert rel <= 1e-12
        outs.append(du)
        dev.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
