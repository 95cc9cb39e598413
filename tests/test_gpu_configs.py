"""The benchmark configurations on the GPU, each anchored to the oracle.

* 64 x 64 ALEVB tether-pull ramp (the plain FP64 workload): the first three
  time steps from the benchmark's seeded initial state match the oracle's
  trajectory (tests/golden/c2_alevb64_ramp.npz, written by
  tests/golden/make_c2_trajectory.py): equal Newton iteration counts per step
  and the state after the last step within 1e-8 relative.
* 256 x 256 ALEVB (the precision form's size): one Newton solve with the FP32
  band LU inside FP64 FGMRES reaches its FP64 residual target and agrees with
  the FP64 direct solve of the same tangent; the FP32 band of this patch has
  more entries than one dispatch counts (4.5e9), so the solve also covers the
  grid-stride band zeroing.
* Batched form: every member of a mixed batch drawn from the batched
  workload's generator (sizes, formulations, kb, zv, tether magnitudes)
  matches the oracle run of the same patch (same Newton counts, state within
  1e-8); one step of the full 512-patch workload converges for every member,
  and its smallest members match the oracle.
* Serving form: requests drawn from the serving stream's generator at its
  sizes (32 ... 128) and formulations (LAG, EUL, ALEVB) give the same result
  bit for bit alone and in the mix.
"""
import os

import numpy as np
import pytest

import mafem
from mafem.patches import Patch, PatchSpec, batched_specs, serving_specs
from oracle import fem, mesh as om

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "c2_alevb64_ramp.npz")
DT = 1e-3


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_c2_ramp_matches_oracle_trajectory():
    print("config 2: 64x64 ALEVB ramp, 3 steps vs the oracle trajectory", flush=True)
    g = np.load(GOLD)
    n = 64
    p = mafem.Params(motion=mafem.ALEVB, num1el=n, num2el=n, length=1.0, kb=1.0, zv=1.0, kg=-0.5,
                     enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.0)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.0)
    rng = np.random.default_rng(0xC2)
    xms[:, 2] += rng.normal(0.0, 1e-6, mesh.numnp)
    # the same initial state as the oracle's (positions agree to rounding:
    # the two collocation fits sum in different orders)
    assert abs(float(xms.sum()) - float(g["x0_sum"])) <= 1e-12
    assert float(cps.sum()) == pytest.approx(float(g["c0_sum"]), rel=1e-15)
    dev = mafem.DeviceMesh(mesh, p)
    dev.set_state(xms, cps)
    pres = [(node, mesh.dofs[u]) for (u, node, _v) in mesh.inh_dir_bcs]
    nodes, dofs = [a for a, _ in pres], [b for _, b in pres]
    steps = len(g["iters"])
    for k in range(steps):
        dev.set_prescribed(nodes, dofs, [2.0 * (k + 1) / 200] * len(pres))
        dev.predict(DT)
        rc, eps = dev.time_step((k + 1) * DT, DT)
        assert rc == 0
        assert len(eps) == int(g["iters"][k]), (k, eps, g["eps"][k])
        # the iterations before convergence follow the oracle's
        np.testing.assert_allclose(eps[:-1], g["eps"][k][:len(eps) - 1], rtol=1e-6)
    x, c = dev.get_state()
    dev.close()
    assert rel(c, g["cps"]) <= 1e-8
    assert rel(x, g["xms"]) <= 1e-8


def test_c3_mixed_precision_solve_256():
    print("config 3 size: 256x256 ALEVB, FP32 band LU in FP64 FGMRES vs FP64 direct", flush=True)
    n = 256
    p = mafem.Params(motion=mafem.ALEVB, num1el=n, num2el=n, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.02)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.02)
    mafem.update_xms(p.motion, xms, cps, DT, mesh.dofs)
    sols = {}
    for prec in (32, 64):
        dev = mafem.DeviceMesh(mesh, p)
        # both solves to relative residual 1e-12: the solutions' difference is
        # bounded by cond(K) times the residuals, and cond(K) is ~1e6 or more here
        if prec == 32:
            dev.set_solver("krylov", precision=32, restart=30, max_iters=2000, rtol=1e-12)
        dev.set_state(xms, cps)
        r, _ = dev.assemble(DT, DT, want_K=False)
        du, relres = dev.solve(-r)
        st = dev.stats()
        print(f"  FP{prec}: relres {relres:.2e}, GMRES {st['gmres_iters']}, "
              f"factor {st['factor_us'] / 1e3:.0f} ms, solve {st['solve_us'] / 1e3:.0f} ms",
              flush=True)
        assert np.isfinite(du).all()
        assert relres <= 1e-12, (prec, relres)
        sols[prec] = du
        dev.close()
    assert rel(sols[32], sols[64]) <= 1e-6


# ------------------------------------------------------------- batched form

def _oracle_patch(spec, steps):
    po = om.Params(motion=int(spec.motion), num1el=spec.n, num2el=spec.n, length=1.0,
                   kb=spec.kb, zv=spec.zv, kg=-0.5, enr=1e-10)
    o, xms, cps = fem.build(po, pull_speed=0.0)
    m = o.mesh
    its = []
    for k in range(steps):
        s = spec.speed_at(k)
        for (u, node, _v) in m.inh_dir_bcs:
            cps[node - 1, m.dofs[u] - 1] = s
            if int(spec.motion) == int(mafem.EUL) and u == om.VZ:
                cps[node - 1, m.dofs[om.VMZ] - 1] = s
        fem.update_xms(po.motion, xms, cps, DT, m.dofs)
        its.append(len(fem.time_step(o, xms, cps, (k + 1) * DT, DT)))
    return its, xms, cps


def _batch_steps(patches, steps, **opts):
    b = mafem.Batch([pt.dev for pt in patches], **opts)
    its = []
    for k in range(steps):
        for pt in patches:
            pt.set_pull(pt.spec.speed_at(k))
        it, eps, st = b.time_step((k + 1) * DT, DT)
        assert (st == 0).all(), (k, st, eps)
        its.append(it.copy())
    stats = b.stats()
    b.close()
    return np.array(its).T, stats


def test_c4_batch_members_match_oracle():
    print("config 4 mix: batch members from the batched generator vs the oracle", flush=True)
    specs = [s for s in batched_specs(64, seed=0xC4, sizes=(16, 24))][:6]
    steps = 3
    patches = [Patch(s) for s in specs]
    its, stats = _batch_steps(patches, steps)
    assert stats["hard_failures"] == 0
    for k, (s, pt) in enumerate(zip(specs, patches)):
        its_o, x_o, c_o = _oracle_patch(s, steps)
        x, c = pt.dev.get_state()
        assert list(its[k]) == its_o, (k, s, its[k], its_o)
        assert rel(c, c_o) <= 1e-8, (k, s)
        assert rel(x, x_o) <= 1e-8, (k, s)
        pt.close()


@pytest.mark.timeout(600)
def test_c4_full_scale_step():
    print("config 4 full scale: one step of 512 patches 16..96", flush=True)
    specs = batched_specs(512, seed=0xC4, sizes=(16, 96))
    patches = [Patch(s) for s in specs]
    print("  512 patches on the device", flush=True)
    b = mafem.Batch([pt.dev for pt in patches])
    for pt in patches:
        pt.set_pull(pt.spec.speed_at(0))
    it, eps, st = b.time_step(DT, DT)
    stats = b.stats()
    b.close()
    print(f"  iterations {int(it.sum())}, hard failures {stats['hard_failures']}, "
          f"FP64 members {stats['fp64_patches']}", flush=True)
    assert (st == 0).all()
    assert all(e[-1] < 1e-10 for e in eps)
    small = [k for k, s in enumerate(specs) if s.n <= 18][:3]
    assert small
    for k in small:
        its_o, x_o, c_o = _oracle_patch(specs[k], 1)
        x, c = patches[k].dev.get_state()
        assert it[k] == its_o[0], (k, specs[k])
        assert rel(c, c_o) <= 1e-8 and rel(x, x_o) <= 1e-8, (k, specs[k])
    for pt in patches:
        pt.close()


# ------------------------------------------------------------- serving form

@pytest.mark.timeout(600)
def test_c5_serving_mix_alone_and_in_mix():
    from mafem.serving import Request, Server
    print("config 5 mix: serving requests n 32..128, LAG/EUL/ALEVB, alone vs mix", flush=True)
    pool = serving_specs(200, seed=0xC5)
    picked = []
    for n in (32, 48, 64, 96, 128):
        for mo in (mafem.LAG, mafem.EUL, mafem.ALEVB):
            s = next((s for s in pool if s.n == n and int(s.motion) == int(mo)), None)
            if s is not None and len(picked) < 6 and all(q.n != n for q in picked):
                picked.append(s)
    assert {s.n for s in picked} >= {32, 64, 128}
    steps = [3, 2, 3, 2, 2, 2][:len(picked)]
    reqs = [Request(k, s, steps=steps[k], arrival=0.02 * k) for k, s in enumerate(picked)]
    srv = Server(max_members=8)
    mixed = srv.run(reqs)
    for k, r in enumerate(mixed):
        assert r.failed < 0 and r.state is not None, (k, picked[k], r.error)
        alone = Server(max_members=1).run([Request(k, picked[k], steps=steps[k])])[0]
        assert alone.iters == r.iters, k
        assert np.array_equal(alone.state[0], r.state[0]), k
        assert np.array_equal(alone.state[1], r.state[1]), k
