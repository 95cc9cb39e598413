"""Pull force on the device (mafem_pull_force, mafem_batch_pull_force) and the
reference's output files, through the C ABI.

Bars:
  * device pull force and pulled-element position against the oracle's
    calc_pull_force restatement on the same perturbed states: relative error
    <= 1e-12 (FP64 residual evaluation; only the summation order differs);
  * the reference's documented tether example run end to end on the device
    (host mirror `solve` with output): f-pull.txt read back as the docs do
    reproduces the figure's curve within 0.03 (about one pixel) and the
    oracle's first steps within 1e-8;
  * a restart from the written text state continues bit for bit like the
    uninterrupted run (written doubles read back exactly);
  * the batched form equals the per-context form member by member.
"""
import json
import os

import numpy as np
import pytest

import mafem
from mafem import output as out
from oracle import fem, mesh as om

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def perturbed(motion, n, seed, speed=0.4):
    p = mafem.Params(motion=motion, num1el=n, num2el=n, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=speed)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=speed)
    rng = np.random.default_rng(seed)
    xms[:, 2] += rng.normal(0, 1e-2, mesh.numnp)
    cps += rng.normal(0, 1e-1, cps.shape)
    po = om.Params(motion=int(motion), num1el=n, num2el=n, length=1.0, enr=1e-10)
    o, _, _ = fem.build(po, pull_speed=speed)
    return p, mesh, xms, cps, o


@pytest.mark.parametrize("motion,n", [(mafem.LAG, 8), (mafem.EUL, 9), (mafem.ALEV, 12),
                                      (mafem.ALEVB, 20), (mafem.ALEVB, 64)])
def test_pull_force_matches_oracle(motion, n):
    print(f"pull force parity motion={motion.name} n={n}", flush=True)
    p, mesh, xms, cps, o = perturbed(motion, n, seed=n)
    dev = mafem.DeviceMesh(mesh, p)
    try:
        dev.set_state(xms, cps)
        x, f = dev.pull_force()
    finally:
        dev.close()
    ids, maps = fem.get_adj_maps(n, o.mesh.numel, o.mesh.IX)
    f_o = fem.calc_pull_force(o, xms, cps, ids, maps)
    x_o = fem.pull_position(o.mesh, xms)
    assert np.linalg.norm(f - f_o) <= 1e-12 * np.linalg.norm(f_o)
    assert np.allclose(x, x_o, rtol=1e-14, atol=1e-15)
    # the host mirror of calc_pull_force takes the same maps
    ids_h, maps_h = mafem.get_adj_maps(mesh.num1el, mesh.numel, mesh.IX)
    f_h = mafem.calc_pull_force(mesh, xms, cps, ids_h, maps_h, p)
    assert np.array_equal(f_h, f)


def test_pull_force_flat_closed_form():
    """Flat tension state, right-edge pulled element (even n): f = (kb/4) L x
    the integral of its three ζ2 basis functions, 0, 0."""
    p = mafem.Params(motion=mafem.ALEVB, num1el=20, num2el=20, length=4.0, kb=0.3,
                     output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.5)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.5)
    dev = mafem.DeviceMesh(mesh, p)
    try:
        dev.set_state(xms, cps)
        _, f = dev.pull_force()
    finally:
        dev.close()
    po = om.Params(motion=om.ALEVB, num1el=20, num2el=20, length=4.0, kb=0.3)
    t = om.make_mesh(po, pull_speed=0.5).kv2.zs
    el2 = 10
    integral = sum((t[j + 2] - t[j - 1]) / 3.0 for j in (el2, el2 + 1, el2 + 2))
    assert np.allclose(f, [0.3 / 4 * 4.0 * integral, 0.0, 0.0], rtol=1e-13, atol=1e-14)


def test_documented_tether_example(tmp_path):
    fig = json.load(open(os.path.join(HERE, "golden", "docs_pull_force.json")))
    ex = fig["example"]
    p = mafem.Params(motion=mafem.LAG, scenario=mafem.F_PULL, num1el=ex["num1el"],
                     num2el=ex["num2el"], length=ex["length"])
    print("documented tether example on the device", flush=True)
    mesh, xms, cps = mafem.solve(p, pull_speed=ex["pull_speed"], t0=0.0, t0_id=0,
                                 Δts=[ex["dt"]] * ex["steps"], out_path=str(tmp_path),
                                 out_file="out.txt")
    data = np.loadtxt(tmp_path / "f-pull.txt", skiprows=1)
    assert (tmp_path / "f-pull.txt").read_text().startswith(out.PULL_HEADER)
    assert data.shape == (32, 7)
    tether_z, tether_f = data[:, 3], data[:, 6]
    zf = np.array([z for z, _ in fig["zp_fz"]])
    ff = np.array([f for _, f in fig["zp_fz"]])
    assert np.allclose(data[:, 0], 0.5 * np.arange(1, 33))
    assert np.allclose(tether_z, zf, atol=1e-9)
    assert np.abs(tether_f - ff).max() < 0.03, np.abs(tether_f - ff).max()
    # state files: the last one is the returned state, exactly
    assert np.array_equal(out.readdlm(tmp_path / "t32-xms.txt"), xms)
    assert np.array_equal(out.readdlm(tmp_path / "t32-cps.txt"), cps)
    assert (tmp_path / "times.txt").read_text().splitlines()[:2] == ["32 times:", "1\t0.5"]
    # first steps against the oracle
    po = om.Params(motion=om.LAG, scenario=om.F_PULL, num1el=ex["num1el"], num2el=ex["num2el"],
                   length=ex["length"], output=False)
    o, xo, co = fem.build(po, pull_speed=ex["pull_speed"])
    ids, maps = fem.get_adj_maps(ex["num1el"], o.mesh.numel, o.mesh.IX)
    for k in range(3):
        fem.update_xms(po.motion, xo, co, ex["dt"], o.mesh.dofs)
        fem.time_step(o, xo, co, ex["dt"] * (k + 1), ex["dt"])
        f_o = fem.calc_pull_force(o, xo, co, ids, maps)
        assert abs(data[k, 6] - f_o[2]) <= 1e-8 * abs(f_o[2])
        assert abs(data[k, 3] - fem.pull_position(o.mesh, xo)[2]) <= 1e-12


def test_restart_from_text_state_is_exact(tmp_path):
    p = mafem.Params(motion=mafem.ALEVB, scenario=mafem.F_PULL, num1el=12, num2el=12,
                     length=1.0, enr=1e-10)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    args = dict(pull_speed=0.5, t0=0.0, t0_id=0, out_file="out.txt")
    print("restart: straight run", flush=True)
    _, x4, c4 = mafem.solve(p, Δts=[1e-3] * 4, out_path=str(a), **args)
    print("restart: from t2", flush=True)
    _, x4r, c4r = mafem.restart(p, dict(args, out_path=str(a)), in_path=str(a), out_path=str(b),
                                t0_id=2, t0=2e-3, Δts=[1e-3] * 2)
    assert np.array_equal(x4r, x4) and np.array_equal(c4r, c4)
    assert (b / "t4-cps.txt").read_text() == (a / "t4-cps.txt").read_text()
    fa = (a / "f-pull.txt").read_text().splitlines()
    fb = (b / "f-pull.txt").read_text().splitlines()
    assert fb[0] == fa[0] and len(fb) == 3
    for la, lb in zip(fa[3:5], fb[1:]):
        ta, tb_ = la.split("\t"), lb.split("\t")
        # times are t0 + cumsum(Δts) in both runs, equal up to the last bit
        assert abs(float(ta[0]) - float(tb_[0])) <= 1e-15 and ta[1:] == tb_[1:]


def test_batch_pull_force_matches_single():
    specs = [(mafem.LAG, 8, 1), (mafem.ALEVB, 12, 2), (mafem.EUL, 9, 3)]
    devs, states = [], []
    try:
        for motion, n, seed in specs:
            p = mafem.Params(motion=motion, num1el=n, num2el=n, length=1.0, enr=1e-10,
                             output=False)
            mesh = mafem.make_mesh(p, pull_speed=0.3)
            xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.3)
            d = mafem.DeviceMesh(mesh, p)
            d.set_state(xms, cps)
            devs.append(d)
        batch = mafem.Batch(devs, precision=64)
        try:
            it, eps, st = batch.time_step(np.full(3, 1e-3), 1e-3, predict=True)
            assert list(st) == [0, 0, 0]
            got = batch.pull_force()
        finally:
            batch.close()
        for k, d in enumerate(devs):
            x, f = d.pull_force()
            assert np.array_equal(got[k, :3], x) and np.array_equal(got[k, 3:], f)
            assert np.abs(f).max() > 0
    finally:
        for d in devs:
            d.close()


@pytest.mark.parametrize("motion,n1,n2", [(mafem.LAG, 4, 3), (mafem.ALEVB, 20, 19),
                                          (mafem.EUL, 9, 9), (mafem.ALEVB, 64, 64)])
def test_generate_output_matches_oracle(motion, n1, n2):
    print(f"generate_output motion={motion.name} {n1}x{n2}", flush=True)
    p = mafem.Params(motion=motion, num1el=n1, num2el=n2, length=2.0, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.3)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.3)
    rng = np.random.default_rng(n1 + 100 * n2)
    xms += rng.normal(0, 1e-2, xms.shape)
    cps += rng.normal(0, 1e-1, cps.shape)
    xout, uout = mafem.generate_output(mesh, xms, cps, p)
    po = om.Params(motion=int(motion), num1el=n1, num2el=n2, length=2.0)
    mo = om.make_mesh(po, pull_speed=0.3)
    xo, uo = fem.generate_output(mo, xms, cps)
    assert xout.shape == xo.shape and uout.shape == uo.shape
    assert np.abs(xout - xo).max() <= 1e-14 * np.abs(xo).max()
    assert np.abs(uout - uo).max() <= 1e-14 * np.abs(uo).max()
