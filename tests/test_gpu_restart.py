"""Restart parity: the oracle continues from the state files the device wrote.

The device run writes t<id>-xms.txt / t<id>-cps.txt after every step
(src/Analysis.jl:75-78); restart reads them back with readdlm
(src/MembraneAleFem.jl:75-107, src/Input.jl:35-37).  Here the oracle takes
the device-written text of step 5, parsed with numpy alone, and continues for
5 steps; its iteration counts and final state must match the oracle's own
straight 10-step run.  The written numbers are the shortest decimal strings
that read back to the same doubles, so the text carries the state exactly
(checked against the device's downloaded state).
"""
import numpy as np
import pytest

import mafem
from mafem import output as out
from oracle import fem, mesh as om

pytestmark = pytest.mark.gpu

DT = 1e-3


def _oracle_steps(o, xms, cps, k0, k1):
    its = []
    for k in range(k0, k1):
        fem.update_xms(o.p.motion, xms, cps, DT, o.mesh.dofs)
        its.append(len(fem.time_step(o, xms, cps, (k + 1) * DT, DT)))
    return its


@pytest.mark.parametrize("motion,n", [(mafem.LAG, 8), (mafem.ALEVB, 10)])
def test_oracle_continues_from_device_written_state(tmp_path, motion, n):
    print(f"restart parity motion={motion.name} n={n}", flush=True)
    p = mafem.Params(motion=motion, scenario=mafem.F_PULL, num1el=n, num2el=n, length=1.0,
                     enr=1e-10)
    args = dict(pull_speed=0.5, t0=0.0, t0_id=0, out_file="out.txt")
    _, x5_dev, c5_dev = mafem.solve(p, Δts=[DT] * 5, out_path=str(tmp_path), **args)
    x5 = np.loadtxt(tmp_path / "t5-xms.txt", delimiter="\t")
    c5 = np.loadtxt(tmp_path / "t5-cps.txt", delimiter="\t")
    assert np.array_equal(x5, x5_dev) and np.array_equal(c5, c5_dev)
    # the text itself: Julia's Float64 printing, e.g. "0.0", "0.25", "1.0e-5"
    first = (tmp_path / "t5-cps.txt").read_text().splitlines()[0].split("\t")
    assert all(out.julia_float(float(v)) == v for v in first)

    po = om.Params(motion=int(motion), num1el=n, num2el=n, length=1.0, enr=1e-10)
    o, xs, cs = fem.build(po, pull_speed=0.5)
    its_straight = _oracle_steps(o, xs, cs, 0, 10)
    o2, _, _ = fem.build(po, pull_speed=0.5)
    xr, cr = x5.copy(), c5.copy()
    its_cont = _oracle_steps(o2, xr, cr, 5, 10)
    assert its_cont == its_straight[5:]
    assert np.linalg.norm(cr - cs) / np.linalg.norm(cs) <= 1e-8
    assert np.linalg.norm(xr - xs) / np.linalg.norm(xs) <= 1e-8
    # and the device restart from the same files agrees with the oracle
    _, x10, c10 = mafem.restart(p, dict(args, out_path=str(tmp_path)), in_path=str(tmp_path),
                                t0_id=5, t0=5 * DT, Δts=[DT] * 5)
    assert np.linalg.norm(c10 - cs) / np.linalg.norm(cs) <= 1e-8
    assert np.linalg.norm(x10 - xs) / np.linalg.norm(xs) <= 1e-8
