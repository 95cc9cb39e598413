"""Nested-dissection multifrontal factorization (MAFEM_ORDER_ND) on the GPU.

The ordering replaces the band's natural node order; its solves must meet the
same bars as the band's (relative residual <= 1e-12 after refinement,
agreement with SuperLU to the conditioning of K, trajectories equal to the
oracle's with the same Newton counts), and its factorization alone (no
refinement) must already solve to the accuracy the tile-local pivoting gives.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import mafem
from oracle import fem, mesh as om

pytestmark = pytest.mark.gpu


def _setup(motion, n1, n2=None, seed=0):
    n2 = n1 if n2 is None else n2
    p = mafem.Params(motion=motion, num1el=n1, num2el=n2, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.5)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.5)
    rng = np.random.default_rng(seed)
    xms[:, 2] += rng.normal(0, 1e-3, mesh.numnp)
    free = mesh.ID.T != 0
    cps += np.where(free, rng.normal(0, 1e-2, cps.shape), 0.0)
    return p, mesh, xms, cps


def _solve(dev, r, ordering, refine=3):
    dev.set_ordering(ordering)
    dev.set_solver("direct", refine_sweeps=refine, rtol=1e-13)
    try:
        du, rel = dev.solve(-r)
        st = dev.stats()
    finally:
        dev.set_ordering("band")
        dev.set_solver("direct")
    return du, rel, st


@pytest.mark.parametrize("motion,n1,n2", [(mafem.LAG, 8, 8), (mafem.ALEVB, 16, 16),
                                          (mafem.EUL, 12, 12), (mafem.ALEV, 12, 9),
                                          (mafem.ALEVB, 5, 14), (mafem.LAG, 1, 1)])
def test_nd_solve_matches_band_and_superlu(motion, n1, n2):
    print(f"nd solve motion={motion.name} {n1}x{n2}", flush=True)
    p, mesh, xms, cps = _setup(motion, n1, n2)
    dev = mafem.analysis.device_for(mesh, p)
    dev.set_state(xms, cps)
    r, K = dev.assemble(1e-3, 1e-3)
    du, rel, st = _solve(dev, r, "nd")
    assert st["fronts"] >= 1 and st["front_steps"] >= 1 and st["factor_chains"] == 0
    assert rel <= 1e-12
    assert np.linalg.norm(K @ du + r) / np.linalg.norm(r) <= 1e-11
    db, relb, _ = _solve(dev, r, "band")
    assert relb <= 1e-12
    # the solution is unique only when K is (the single 1x1 LAG element has
    # rigid motions left): then both orderings and SuperLU agree
    if K.shape[0] > 200 or np.linalg.cond(K.toarray()) < 1e12:
        ref = spla.spsolve(K.tocsc(), -r)
        assert np.linalg.norm(du - ref) / np.linalg.norm(ref) <= 1e-7
        assert np.linalg.norm(du - db) / np.linalg.norm(db) <= 1e-7


@pytest.mark.parametrize("motion,n", [(mafem.ALEVB, 16), (mafem.LAG, 12)])
def test_nd_factorization_alone_solves(motion, n):
    """No refinement: the multifrontal factors themselves solve K du = -r."""
    p, mesh, xms, cps = _setup(motion, n)
    dev = mafem.analysis.device_for(mesh, p)
    dev.set_state(xms, cps)
    r, K = dev.assemble(1e-3, 1e-3)
    du, rel, st = _solve(dev, r, "nd", refine=0)
    print(f"nd unrefined relres {rel:.3e} ({st['fronts']} fronts)", flush=True)
    assert rel <= 1e-9


def test_nd_time_steps_match_oracle():
    """ALEVB 16x16, pull ramped towards 2.0 over 200 steps, 4 steps: same
    Newton counts and states within 1e-8 of the oracle with nested dissection."""
    n, steps = 16, 4
    p = mafem.Params(motion=mafem.ALEVB, num1el=n, num2el=n, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.0)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.0)
    rng = np.random.default_rng(0xC2)
    xms[:, 2] += rng.normal(0, 1e-4, mesh.numnp)
    po = om.Params(motion=int(mafem.ALEVB), num1el=n, num2el=n, length=1.0, enr=1e-10)
    o, _, _ = fem.build(po, pull_speed=0.0)
    sched = lambda k, t: 2.0 * (k + 1) / 200
    x_o, c_o = xms.copy(), cps.copy()
    m = o.mesh
    times = np.cumsum([1e-3] * steps)
    its_o = []
    for k in range(steps):
        for (u, node, _v) in m.inh_dir_bcs:
            c_o[node - 1, m.dofs[u] - 1] = sched(k, times[k])
        fem.update_xms(po.motion, x_o, c_o, 1e-3, m.dofs)
        its_o.append(len(fem.time_step(o, x_o, c_o, float(times[k]), 1e-3)))
    dev = mafem.analysis.device_for(mesh, p)
    dev.set_ordering("nd")
    try:
        hist = mafem.run_analysis(mesh, xms, cps, p, dts=[1e-3] * steps, t0=0.0,
                                  pull_schedule=sched)
        st = dev.stats()
    finally:
        dev.set_ordering("band")
    assert st["fronts"] > 0
    assert [len(h) for h in hist] == its_o
    assert np.linalg.norm(cps - c_o) / np.linalg.norm(c_o) <= 1e-8
    assert np.linalg.norm(xms - x_o) / np.linalg.norm(x_o) <= 1e-8


@pytest.mark.parametrize("precision", [32, 64])
def test_nd_krylov_preconditioner(precision):
    """FGMRES preconditioned by kept multifrontal factors (FP32 or FP64)."""
    p, mesh, xms, cps = _setup(mafem.ALEVB, 20)
    dev = mafem.analysis.device_for(mesh, p)
    dev.set_state(xms, cps)
    r, K = dev.assemble(1e-3, 1e-3)
    dev.set_ordering("nd")
    dev.set_solver("krylov", precision=precision, rtol=1e-13)
    try:
        du, rel = dev.solve(-r)
        st = dev.stats()
        dev.set_state(xms * (1 + 1e-6), cps)
        r2, K2 = dev.assemble(1e-3, 1e-3)
        du2, rel2 = dev.solve(-r2)
        st2 = dev.stats()
    finally:
        dev.set_ordering("band")
        dev.set_solver("direct")
    print(f"nd krylov fp{precision}: {st['gmres_iters']} + {st2['gmres_iters']} iterations",
          flush=True)
    assert rel <= 1e-13 and st["n_factor"] == 1
    assert np.linalg.norm(K @ du + r) / np.linalg.norm(r) <= 1e-12
    assert rel2 <= 1e-13 and st2["n_factor"] == 0
    assert np.linalg.norm(K2 @ du2 + r2) / np.linalg.norm(r2) <= 1e-12


def test_nd_benchmark_patch():
    """64x64 ALEVB: nested dissection and the two-ended band solve the same
    tangent to 1e-12 and agree; the fronts' dependent chain is the plan's."""
    n = 64
    p = mafem.Params(motion=mafem.ALEVB, num1el=n, num2el=n, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.0)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.0)
    rng = np.random.default_rng(7)
    xms[:, 2] += rng.normal(0, 1e-4, mesh.numnp)
    dev = mafem.analysis.device_for(mesh, p)
    dev.set_state(xms, cps)
    r, K = dev.assemble(1e-3, 1e-3)
    du, rel, st = _solve(dev, r, "nd")
    db, relb, stb = _solve(dev, r, "band")
    print(f"64x64 nd: relres {rel:.2e}, {st['fronts']} fronts, {st['front_steps']} steps, "
          f"factor {st['factor_us']:.0f} us, solve {st['solve_us']:.0f} us; band factor "
          f"{stb['factor_us']:.0f} us, solve {stb['solve_us']:.0f} us", flush=True)
    assert st["front_steps"] == 48
    assert rel <= 1e-12 and relb <= 1e-12
    assert np.linalg.norm(K @ du + r) / np.linalg.norm(r) <= 1e-11
    assert np.linalg.norm(du - db) / np.linalg.norm(db) <= 1e-8


@pytest.mark.parametrize("la,inv", [("1", "1"), ("1", "100000"), ("0", "1"), ("3", "16")])
def test_nd_panel_forms_on_every_level(monkeypatch, la, inv):
    """The panels as products with inverses formed by the LUs (levels of at
    least MAFEM_ND_INV_FRONTS fronts) and the staged panels (the others), the
    look-ahead steps on the top MAFEM_ND_LA_FRONTS-front levels: every mix
    solves the 20x20 ALEVB tangent like SuperLU, its factors alone to 1e-9."""
    monkeypatch.setenv("MAFEM_ND_LA_FRONTS", la)
    monkeypatch.setenv("MAFEM_ND_INV_FRONTS", inv)
    p, mesh, xms, cps = _setup(mafem.ALEVB, 20)
    dev = mafem.analysis.DeviceMesh(mesh, p)
    try:
        dev.set_state(xms, cps)
        r, K = dev.assemble(1e-3, 1e-3)
        dev.set_ordering("nd")
        dev.set_solver("direct", refine_sweeps=0, rtol=1e-13)
        du, rel = dev.solve(-r)
    finally:
        dev.close()
    print(f"nd la={la} inv={inv}: unrefined relres {rel:.2e}", flush=True)
    assert rel <= 1e-9
    ref = spla.spsolve(K.tocsc(), -r)
    assert np.linalg.norm(du - ref) / np.linalg.norm(ref) <= 1e-7
