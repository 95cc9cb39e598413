"""Batched patches (many contexts advanced through one time step together).

Bars:
  * each member's converged state after N steps agrees with the same patch
    run alone on the single-patch FP64 direct path (itself pinned to the
    oracle) within 1e-8 relative, with the same Newton iteration counts;
  * a patch's trajectory is bit-identical alone (a batch of one) and in a
    mixed batch, and with or without graph replay: every batched kernel
    computes each patch from that patch's data only, in a fixed order;
  * the FP64-factor batch agrees with the FP64 direct path to 1e-8.
"""
import numpy as np
import pytest

import mafem

pytestmark = pytest.mark.gpu

RTOL_STATE = 1e-8

# (motion, n, kb, zv, pull speed): mixed formulations, sizes (uniform and
# centre-refined knots) and parameters
PATCHES = [(mafem.LAG, 8, 1.0, 1.0, 0.5), (mafem.ALEVB, 10, 0.3, 2.0, 0.4),
           (mafem.EUL, 12, 2.5, 0.5, 0.3), (mafem.ALEV, 9, 1.0, 1.0, 0.2),
           (mafem.ALEVB, 19, 0.7, 1.5, 0.5), (mafem.LAG, 17, 4.0, 0.2, 0.3)]
STEPS = 4
DT = 1e-3


def make(motion, n, kb, zv, speed, seed):
    p = mafem.Params(motion=motion, num1el=n, num2el=n, length=1.0, kb=kb, zv=zv, enr=1e-10,
                     output=False)
    mesh = mafem.make_mesh(p, pull_speed=speed)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=speed)
    rng = np.random.default_rng(seed)
    xms[:, 2] += rng.normal(0, 1e-6, mesh.numnp)
    return p, mesh, xms, cps


def run_alone(spec, seed):
    p, mesh, xms, cps = make(*spec, seed)
    dev = mafem.DeviceMesh(mesh, p)
    dev.set_state(xms, cps)
    its = []
    for k in range(STEPS):
        dev.predict(DT)
        rc, eps = dev.time_step((k + 1) * DT, DT)
        assert rc == 0
        its.append(len(eps))
    x, c = dev.get_state()
    dev.close()
    return its, x, c


def run_batch(specs, seeds, **opts):
    devs = []
    for spec, seed in zip(specs, seeds):
        p, mesh, xms, cps = make(*spec, seed)
        d = mafem.DeviceMesh(mesh, p)
        d.set_state(xms, cps)
        devs.append(d)
    b = mafem.Batch(devs, **opts)
    its = []
    for k in range(STEPS):
        it, eps, st = b.time_step((k + 1) * DT, DT)
        assert (st == 0).all(), (k, st, eps)
        its.append(it.copy())
    states = [d.get_state() for d in devs]
    stats = b.stats()
    b.close()
    for d in devs:
        d.close()
    return np.array(its).T, states, stats


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def alone():
    return [run_alone(s, 100 + k) for k, s in enumerate(PATCHES)]


def test_batch_matches_single_patch_fp64(alone):
    print("batch of 6 mixed patches, FP32 factor + FP64 refinement", flush=True)
    its, states, stats = run_batch(PATCHES, [100 + k for k in range(len(PATCHES))])
    for k, (its_a, x_a, c_a) in enumerate(alone):
        x, c = states[k]
        assert list(its[k]) == its_a, (k, its[k], its_a)
        assert rel(x, x_a) <= RTOL_STATE, k
        assert rel(c, c_a) <= RTOL_STATE, k
    assert stats["patch_iters"] == sum(its[:, -1])
    assert stats["hard_failures"] == 0


def test_batch_fp64_factor_matches_single_patch(alone):
    print("batch of 6 mixed patches, FP64 factor", flush=True)
    its, states, _ = run_batch(PATCHES, [100 + k for k in range(len(PATCHES))],
                               precision=64)
    for k, (its_a, x_a, c_a) in enumerate(alone):
        x, c = states[k]
        assert list(its[k]) == its_a
        assert rel(c, c_a) <= RTOL_STATE


def test_alone_and_in_mix_are_identical():
    print("alone vs in the mix vs without graphs", flush=True)
    seeds = [100 + k for k in range(len(PATCHES))]
    _, mix, _ = run_batch(PATCHES, seeds)
    _, mix_ng, _ = run_batch(PATCHES, seeds, use_graphs=False)
    for k in (1, 4):
        _, one, _ = run_batch([PATCHES[k]], [seeds[k]])
        assert np.array_equal(one[0][0], mix[k][0]) and np.array_equal(one[0][1], mix[k][1]), k
    for k in range(len(PATCHES)):
        assert np.array_equal(mix[k][1], mix_ng[k][1]), k


def test_fp32_failure_moves_patch_to_fp64():
    """A large centre-refined LAG patch whose tangent is too ill-conditioned
    for an FP32 factorization: the batch moves it to FP64 storage and its
    state still matches the FP64 direct path."""
    from mafem.patches import Patch, PatchSpec
    print("escalation: LAG 83x83 kb=0.53 zv=8.5 with two small patches", flush=True)
    specs = [PatchSpec(mafem.LAG, 83, 0.5286395248414045, 8.539281469289673, 0.8995302812009891,
                       478865490),
             PatchSpec(mafem.ALEVB, 12, 1.0, 1.0, 0.5, 7), PatchSpec(mafem.LAG, 10, 2.0, 0.5, 0.4, 8)]
    batch_pts = [Patch(s) for s in specs]
    b = mafem.Batch([pt.dev for pt in batch_pts])
    for k in range(2):
        for pt in batch_pts:
            pt.set_pull(pt.spec.speed_at(k))
        it, eps, st = b.time_step((k + 1) * DT, DT)
        assert (st == 0).all(), (k, st, eps)
    stats = b.stats()
    assert stats["fp64_patches"] >= 1
    got = [pt.dev.get_state() for pt in batch_pts]
    b.close()
    for pt, (xb, cb) in zip(batch_pts, got):
        pt.close()
    for s, (xb, cb) in zip(specs, got):
        ref = Patch(s)
        ref.dev.set_solver("direct", precision=64, rtol=1e-13)
        for k in range(2):
            ref.set_pull(s.speed_at(k))
            ref.dev.predict(DT)
            rc, _ = ref.dev.time_step((k + 1) * DT, DT)
            assert rc == 0
        x, c = ref.dev.get_state()
        ref.close()
        assert rel(cb, c) <= RTOL_STATE and rel(xb, x) <= RTOL_STATE


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_multi_workgroup_sweeps_match_single_patch(graphs):
    """Long chains inside a batch sweep over several workgroups; a 48x48 ALEVB
    patch (one of them) next to small ones still matches the FP64 direct path."""
    from mafem.patches import Patch, PatchSpec
    print("batch with a multi-workgroup sweep member", flush=True)
    specs = [PatchSpec(mafem.ALEVB, 48, 1.0, 1.0, 0.5, 11), PatchSpec(mafem.LAG, 8, 1.0, 1.0, 0.5, 12),
             PatchSpec(mafem.EUL, 48, 2.0, 0.5, 0.4, 13), PatchSpec(mafem.LAG, 48, 0.5, 2.0, 0.6, 14),
             PatchSpec(mafem.ALEVB, 64, 1.0, 1.0, 0.3, 15)]
    pts = [Patch(s) for s in specs]
    b = mafem.Batch([pt.dev for pt in pts], use_graphs=graphs)
    for k in range(2):
        for pt in pts:
            pt.set_pull(pt.spec.speed_at(k))
        it, eps, st = b.time_step((k + 1) * DT, DT)
        print(f"  step {k}: iters {list(it)} status {list(st)} stats {b.stats()}", flush=True)
        assert (st == 0).all(), (k, st, eps)
    got = [pt.dev.get_state() for pt in pts]
    b.close()
    for pt in pts:
        pt.close()
    for s, (xb, cb) in zip(specs, got):
        ref = Patch(s)
        ref.dev.set_solver("direct", precision=64, rtol=1e-13)
        for k in range(2):
            ref.set_pull(s.speed_at(k))
            ref.dev.predict(DT)
            assert ref.dev.time_step((k + 1) * DT, DT)[0] == 0
        x, c = ref.dev.get_state()
        ref.close()
        assert rel(cb, c) <= RTOL_STATE and rel(xb, x) <= RTOL_STATE
