"""Parity of the HIP path (through the C ABI) with the oracle.

Bars (from the hot-path contract):
  * assembled r and K against the oracle's complex-step restatement:
    relative Frobenius error <= 1e-12;
  * converged state after N backward-Euler steps: relative error <= 1e-8,
    with the same Newton iteration count in every step;
  * linear solve: |K du - rhs| / |rhs| <= 1e-12 and agreement with a sparse
    direct solve (SuperLU) to the conditioning of K.
At the benchmark size (64^2 ALEVB) size-independent properties are checked:
the flat equilibrium r = 0, exact linearity of r in the tension unknowns,
and the solve residual.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import mafem
from oracle import fem, mesh as om

pytestmark = pytest.mark.gpu

RTOL_ASM = 1e-12
RTOL_STATE = 1e-8


def setup(motion, n, speed=0.5, enr=1e-10, seed=0, perturb=True):
    p = mafem.Params(motion=motion, num1el=n, num2el=n, length=1.0, enr=enr, output=False)
    mesh = mafem.make_mesh(p, pull_speed=speed)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=speed)
    if perturb:
        rng = np.random.default_rng(seed)
        xms[:, 2] += rng.normal(0, 1e-3, mesh.numnp)
        free = mesh.ID.T != 0                        # (numnp, ndf)
        cps += np.where(free, rng.normal(0, 1e-2, cps.shape), 0.0)
    po = om.Params(motion=int(motion), num1el=n, num2el=n, length=1.0, enr=enr)
    o, _, _ = fem.build(po, pull_speed=speed)
    return p, mesh, xms, cps, o


def rel_fro(A, B):
    return np.linalg.norm((A - B).toarray()) / np.linalg.norm(B.toarray())


ASM_CASES = [(mafem.LAG, 8), (mafem.EUL, 10), (mafem.ALEV, 12), (mafem.ALEVB, 16),
             (mafem.ALEVB, 20), (mafem.LAG, 19), (mafem.EUL, 3), (mafem.ALEVB, 1)]


@pytest.mark.parametrize("motion,n", ASM_CASES)
def test_assemble_matches_oracle(motion, n):
    print(f"assemble parity motion={motion.name} n={n}", flush=True)
    p, mesh, xms, cps, o = setup(motion, n)
    dt, t = 1e-3, 2e-3
    r, K = mafem.calc_r_K(mesh, xms, cps, t, dt, p)
    r_o, K_o = fem.calc_r_K(o, xms, cps, t, dt)
    assert np.linalg.norm(r - r_o) / np.linalg.norm(r_o) <= RTOL_ASM
    assert rel_fro(K, K_o) <= RTOL_ASM
    # the exported pattern is a structural superset of the oracle's
    Ks = K.tocsr()
    Ko = K_o.tocsr()
    for i in range(0, mesh.nmdf, max(1, mesh.nmdf // 50)):
        assert set(Ko.indices[Ko.indptr[i]:Ko.indptr[i + 1]]) <= \
            set(Ks.indices[Ks.indptr[i]:Ks.indptr[i + 1]])


def test_assembly_is_deterministic():
    p, mesh, xms, cps, _ = setup(mafem.ALEVB, 16)
    dev = mafem.analysis.device_for(mesh, p)
    dev.set_state(xms, cps)
    r1, K1 = dev.assemble(0.0, 1e-3)
    r2, K2 = dev.assemble(0.0, 1e-3)
    assert np.array_equal(r1, r2) and np.array_equal(K1.data, K2.data)


@pytest.mark.parametrize("motion,n", [(mafem.LAG, 8), (mafem.ALEVB, 16), (mafem.EUL, 12)])
def test_solve_matches_direct(motion, n):
    print(f"solve parity motion={motion.name} n={n}", flush=True)
    p, mesh, xms, cps, _ = setup(motion, n)
    dev = mafem.analysis.device_for(mesh, p)
    dev.set_state(xms, cps)
    r, K = dev.assemble(1e-3, 1e-3)
    du, rel = dev.solve(-r)
    assert rel <= 1e-12
    assert np.linalg.norm(K @ du + r) / np.linalg.norm(r) <= 1e-11
    ref = spla.spsolve(K.tocsc(), -r)
    assert np.linalg.norm(du - ref) / np.linalg.norm(ref) <= 1e-7


def _oracle_run(o, xms, cps, dts, sched=None):
    m, po = o.mesh, o.p
    times = np.cumsum(dts)
    its = []
    for k, dt in enumerate(dts):
        if sched is not None:
            s = sched(k, times[k])
            for (u, node, _v) in m.inh_dir_bcs:
                cps[node - 1, m.dofs[u] - 1] = s
        fem.update_xms(po.motion, xms, cps, dt, m.dofs)
        its.append(len(fem.time_step(o, xms, cps, float(times[k]), float(dt))))
    return its


def test_time_steps_lag_8x8_matches_oracle():
    """Config 1: LAG 8x8, pull speed 0.5, 20 steps of dt = 1e-3, εnr = 1e-10."""
    print("time-step parity LAG 8x8 x20", flush=True)
    p, mesh, xms, cps, o = setup(mafem.LAG, 8, perturb=False)
    dts = [1e-3] * 20
    x_o, c_o = xms.copy(), cps.copy()
    its_o = _oracle_run(o, x_o, c_o, dts)
    hist = mafem.run_analysis(mesh, xms, cps, p, dts=dts, t0=0.0)
    assert [len(h) for h in hist] == its_o
    assert np.linalg.norm(cps - c_o) / np.linalg.norm(c_o) <= RTOL_STATE
    assert np.linalg.norm(xms - x_o) / np.linalg.norm(x_o) <= RTOL_STATE


@pytest.mark.parametrize("motion,n,steps", [(mafem.ALEVB, 16, 4), (mafem.EUL, 10, 3),
                                            (mafem.ALEV, 12, 3)])
def test_time_steps_ramped_pull_matches_oracle(motion, n, steps):
    print(f"time-step parity motion={motion.name} n={n} steps={steps}", flush=True)
    p, mesh, xms, cps, o = setup(motion, n, speed=0.0, perturb=False)
    rng = np.random.default_rng(0xC2)
    xms[:, 2] += rng.normal(0, 1e-4, mesh.numnp)
    dts = [1e-3] * steps
    sched = lambda k, t: 2.0 * (k + 1) / 200          # ramp towards 2.0 over 200 steps
    x_o, c_o = xms.copy(), cps.copy()
    its_o = _oracle_run(o, x_o, c_o, dts, sched)
    hist = mafem.run_analysis(mesh, xms, cps, p, dts=dts, t0=0.0, pull_schedule=sched)
    assert [len(h) for h in hist] == its_o
    assert np.linalg.norm(cps - c_o) / np.linalg.norm(c_o) <= RTOL_STATE
    assert np.linalg.norm(xms - x_o) / np.linalg.norm(x_o) <= RTOL_STATE


def test_time_step_not_converged_raises():
    p, mesh, xms, cps, _ = setup(mafem.LAG, 4, enr=1e-30, perturb=False)
    mafem.update_xms(p.motion, xms, cps, 1e-3, mesh.dofs)
    with pytest.raises(AssertionError, match="did not reach Newton--Raphson tolerance"):
        mafem.time_step(mesh, xms, cps, 1e-3, 1e-3, p)
    assert mesh._ctx[1].stats()["newton_iters"] == 14


def test_benchmark_size_properties():
    """64x64 ALEVB (the plain-FP64 benchmark patch): flat equilibrium r = 0,
    r exactly linear in λ at fixed geometry (K's λ columns), solve residual."""
    print("64x64 ALEVB properties", flush=True)
    n = 64
    p = mafem.Params(motion=mafem.ALEVB, num1el=n, num2el=n, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.0)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.0)
    dev = mafem.analysis.device_for(mesh, p)
    dev.set_state(xms, cps)
    r0, _ = dev.assemble(0.0, 1e-3, want_K=False)
    assert np.abs(r0).max() < 1e-13
    rng = np.random.default_rng(7)
    x1 = xms.copy()
    x1[:, 2] += rng.normal(0, 1e-4, mesh.numnp)
    c1 = cps.copy()
    dev.set_state(x1, c1)
    r1, K1 = dev.assemble(1e-3, 1e-3)
    lam = mesh.dofs[mafem.Unknown.lam] - 1
    dl = np.zeros_like(c1)
    dl[:, lam] = np.where(mesh.ID[lam] != 0, rng.normal(0, 1e-2, mesh.numnp), 0.0)
    dev.set_state(x1, c1 + dl)
    r2, _ = dev.assemble(1e-3, 1e-3, want_K=False)
    du = np.zeros(mesh.nmdf)
    ids = mesh.ID[lam]
    du[ids[ids != 0] - 1] = dl[ids != 0, lam]
    lin = K1 @ du
    # exact up to the rounding of r itself (contributions cancel across Gauss points)
    assert np.linalg.norm((r2 - r1) - lin) / np.linalg.norm(lin) < 1e-8
    dev.set_state(x1, c1)
    r1b, K1b = dev.assemble(1e-3, 1e-3)
    sol, rel = dev.solve(-r1b)
    assert rel <= 1e-12
    assert np.linalg.norm(K1b @ sol + r1b) / np.linalg.norm(r1b) <= 1e-11


def test_benchmark_size_assembly_matches_oracle():
    """One full assembly at 64x64 ALEVB against the oracle."""
    print("64x64 ALEVB assembly parity", flush=True)
    p, mesh, xms, cps, o = setup(mafem.ALEVB, 64, speed=0.02)
    r, K = mafem.calc_r_K(mesh, xms, cps, 1e-3, 1e-3, p)
    r_o, K_o = fem.calc_r_K(o, xms, cps, 1e-3, 1e-3)
    assert np.linalg.norm(r - r_o) / np.linalg.norm(r_o) <= RTOL_ASM
    D = (K - K_o)
    assert spla.norm(D) / spla.norm(K_o) <= RTOL_ASM


@pytest.mark.parametrize("precision", [64, 32])
def test_krylov_solver_time_steps_match_oracle(precision):
    """FGMRES with a kept band-LU preconditioner (FP64 or FP32 storage): the
    FP64 residual target keeps Newton counts and states at the oracle's."""
    print(f"krylov time-step parity precision={precision}", flush=True)
    p, mesh, xms, cps, o = setup(mafem.ALEVB, 16, speed=0.0, perturb=False)
    rng = np.random.default_rng(0xC2)
    xms[:, 2] += rng.normal(0, 1e-4, mesh.numnp)
    dts = [1e-3] * 4
    sched = lambda k, t: 2.0 * (k + 1) / 200
    x_o, c_o = xms.copy(), cps.copy()
    its_o = _oracle_run(o, x_o, c_o, dts, sched)
    dev = mafem.analysis.device_for(mesh, p)
    dev.set_solver("krylov", precision=precision, rtol=1e-13)
    hist = mafem.run_analysis(mesh, xms, cps, p, dts=dts, t0=0.0, pull_schedule=sched)
    st = dev.stats()
    assert [len(h) for h in hist] == its_o
    assert np.linalg.norm(cps - c_o) / np.linalg.norm(c_o) <= RTOL_STATE
    assert np.linalg.norm(xms - x_o) / np.linalg.norm(x_o) <= RTOL_STATE
    assert st["last_relres"] <= 1e-13
    dev.set_solver("direct")


@pytest.mark.parametrize("precision", [64, 32])
def test_krylov_solve_matches_direct(precision):
    p, mesh, xms, cps, _ = setup(mafem.ALEVB, 20)
    dev = mafem.analysis.device_for(mesh, p)
    dev.set_state(xms, cps)
    r, K = dev.assemble(1e-3, 1e-3)
    dev.set_solver("krylov", precision=precision, rtol=1e-13)
    du, rel = dev.solve(-r)
    st = dev.stats()
    assert rel <= 1e-13
    assert np.linalg.norm(K @ du + r) / np.linalg.norm(r) <= 1e-12
    ref = spla.spsolve(K.tocsc(), -r)
    assert np.linalg.norm(du - ref) / np.linalg.norm(ref) <= 1e-7
    assert st["n_factor"] == 1 and st["gmres_iters"] >= 0
    # a second solve on a perturbed state reuses the kept factorization
    dev.set_state(xms * (1 + 1e-6), cps)
    r2, K2 = dev.assemble(1e-3, 1e-3)
    du2, rel2 = dev.solve(-r2)
    assert rel2 <= 1e-13 and dev.stats()["n_factor"] == 0
    assert np.linalg.norm(K2 @ du2 + r2) / np.linalg.norm(r2) <= 1e-12
    dev.set_solver("direct")


def test_solver_options_validation():
    p, mesh, xms, cps, _ = setup(mafem.LAG, 4, perturb=False)
    dev = mafem.analysis.device_for(mesh, p)
    assert dev.get_solver()["mode"] == 0
    with pytest.raises(mafem._lib.LibError):
        dev.set_solver("krylov", precision=16)
    with pytest.raises(mafem._lib.LibError):
        dev.set_solver("krylov", restart=0)


@pytest.mark.parametrize("groups", [2, 5, 16])
def test_multi_workgroup_sweeps_match_single(groups, monkeypatch):
    """The triangular sweeps split over several workgroups (rows published
    through sentinel-filled outputs) sum every row in the single-workgroup
    sweep's order: the solutions are bit-identical."""
    print(f"sweep groups {groups}", flush=True)
    monkeypatch.setenv("MAFEM_ORDERING", "band")  # the band's sweeps
    p, mesh, xms, cps, _ = setup(mafem.ALEVB, 16)
    rhs = np.random.default_rng(3).normal(size=mesh.nmdf)
    sols = []
    for g in (1, groups):
        monkeypatch.setenv("MAFEM_SWEEP_GROUPS", str(g))
        dev = mafem.DeviceMesh(mesh, p)
        dev.set_state(xms, cps)
        dev.assemble(2e-3, 1e-3, want_K=False)
        x, rel = dev.solve(rhs)
        assert rel <= 1e-12
        sols.append(x)
        dev.close()
    assert np.array_equal(sols[1], sols[0])


@pytest.mark.parametrize("motion,n,groups", [(mafem.ALEVB, 48, 8), (mafem.LAG, 64, 16),
                                              (mafem.EUL, 40, 6), (mafem.LAG, 48, 2)])
def test_multi_workgroup_sweeps_larger_patches(motion, n, groups, monkeypatch):
    """Multi-workgroup sweeps on wider bands: the direct solve (band LU + FP64
    refinement) of the Newton right-hand side equals the single-workgroup
    sweep's bit for bit."""
    print(f"sweep groups {groups} motion={motion.name} n={n}", flush=True)
    monkeypatch.setenv("MAFEM_ORDERING", "band")  # the band's sweeps
    p = mafem.Params(motion=motion, num1el=n, num2el=n, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.3)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.3)
    sols = []
    for g in (1, groups):
        monkeypatch.setenv("MAFEM_SWEEP_GROUPS", str(g))
        dev = mafem.DeviceMesh(mesh, p)
        dev.set_state(xms, cps)
        r, _ = dev.assemble(1e-3, 1e-3, want_K=False)
        x, rel = dev.solve(-r)
        print(f"  groups {g}: relres {rel:.2e}", flush=True)
        assert rel <= 1e-10
        sols.append(x)
        dev.close()
    assert np.array_equal(sols[1], sols[0])


def test_direct_solve_on_a_band_beyond_one_dispatch(monkeypatch):
    """A 256x256 ALEVB patch: its FP64 band holds 8321 x 131 x 64^2 = 4.5e9
    entries (36 GB), more than a dispatch's 32-bit work-item count.  The band
    is zeroed by a grid-stride kernel over a capped grid, and the direct solve
    of the Newton right-hand side is finite and meets the refinement target."""
    print("256x256 ALEVB direct solve", flush=True)
    monkeypatch.setenv("MAFEM_ORDERING", "band")
    p = mafem.Params(motion=mafem.ALEVB, num1el=256, num2el=256, length=1.0, enr=1e-10,
                     output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.3)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.3)
    dev = mafem.DeviceMesh(mesh, p)
    try:
        dev.set_state(xms, cps)
        r, _ = dev.assemble(1e-3, 1e-3, want_K=False)
        x, rel = dev.solve(-r)
    finally:
        dev.close()
    print(f"  relres {rel:.2e}", flush=True)
    assert np.isfinite(x).all() and np.abs(x).max() > 0.0
    assert rel <= 1e-10


def test_sweep_autotuner_changes_time_only(monkeypatch):
    """The sweep split is autotuned on the first factorization; every
    candidate gives bit-identical results, so a time step with the tuned
    split equals one with a fixed split."""
    print("sweep autotuner", flush=True)
    monkeypatch.setenv("MAFEM_ORDERING", "band")  # the band's sweep split
    p = mafem.Params(motion=mafem.LAG, num1el=64, num2el=64, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.3)
    xms0, cps0 = mafem.prepare_input(p, mesh, pull_speed=0.3)
    out = []
    for forced in (None, "2"):
        if forced is None:
            monkeypatch.delenv("MAFEM_SWEEP_GROUPS", raising=False)
        else:
            monkeypatch.setenv("MAFEM_SWEEP_GROUPS", forced)
        dev = mafem.DeviceMesh(mesh, p)
        dev.set_state(xms0, cps0)
        dev.predict(1e-3)
        rc, eps = dev.time_step(1e-3, 1e-3)
        assert rc == 0
        g = dev.stats()["sweep_groups"]
        print(f"  forced={forced}: sweep groups {g}", flush=True)
        if forced is None:
            assert g in (2, 4, 8, 16)
        out.append(dev.get_state())
        dev.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
