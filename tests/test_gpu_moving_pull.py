"""Tether translated across the patch (the precision form's workload).

The reference has no moving tether: the pulled element is ceil(numel / 2)
(src/input/Bc.jl:227-243).  This build moves it by giving a context the
numbering of the new Dirichlet set (mafem_set_dof_ids) and setting the new
prescribed values.  Pins:
  * at zero translation the moving path is the fixed path, bit for bit;
  * with the tether stepping one element every two time steps, the device
    run matches the oracle restatement moved the same way (its F_PULL
    scenario rebuilt for the new element): equal Newton iteration counts and
    states within 1e-8 relative.
"""
import numpy as np
import pytest

import mafem
from mafem.analysis import DeviceMesh, lateral_pull_element, move_tether
from mafem.mesh import tether_ids
from oracle import fem, mesh as om

pytestmark = pytest.mark.gpu

DT = 1e-3


def _setup(motion, n, speed):
    p = mafem.Params(motion=motion, num1el=n, num2el=n, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=speed)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=speed)
    return p, mesh, xms, cps


def _device_run(motion, n, speed, steps, fraction):
    p, mesh, xms, cps = _setup(motion, n, speed)
    dev = DeviceMesh(mesh, p)
    dev.set_state(xms, cps)
    el = lateral_pull_element(mesh, 0, steps, fraction)
    its, els = [], []
    for k in range(steps):
        new = lateral_pull_element(mesh, k, steps, fraction)
        if new != el:
            move_tether(dev, mesh, p, el, new, speed)
            el = new
        els.append(el)
        dev.predict(DT)
        rc, eps = dev.time_step((k + 1) * DT, DT)
        assert rc == 0, (k, eps)
        its.append(len(eps))
    x, c = dev.get_state()
    dev.close()
    return its, els, x, c


def test_zero_translation_is_the_fixed_path():
    print("moving tether: zero translation vs fixed, ALEVB 16", flush=True)
    p, mesh, xms, cps = _setup(mafem.ALEVB, 16, 0.5)
    ID, dirs = tether_ids(mesh, p, lateral_pull_element(mesh, 0, 6, 0.0), 0.5)
    assert np.array_equal(ID, mesh.ID)
    hist = mafem.run_analysis(mesh, xms, cps, p, dts=[DT] * 4, t0=0.0)
    its, els, x, c = _device_run(mafem.ALEVB, 16, 0.5, 4, 0.0)
    assert len(set(els)) == 1
    assert its == [len(h) for h in hist]
    assert np.array_equal(x, xms) and np.array_equal(c, cps)


@pytest.mark.parametrize("motion,n", [(mafem.ALEVB, 16), (mafem.LAG, 12), (mafem.EUL, 10)])
def test_moving_tether_matches_oracle(motion, n):
    print(f"moving tether vs oracle motion={motion.name} n={n}", flush=True)
    steps, speed = 6, 0.5
    fraction = 3.0 / n          # one element every two steps
    its, els, x, c = _device_run(motion, n, speed, steps, fraction)
    assert len(set(els)) == 3
    po = om.Params(motion=int(motion), num1el=n, num2el=n, length=1.0, enr=1e-10)
    o, xo, co = fem.build(po, pull_speed=speed)
    m = o.mesh
    el = els[0]
    its_o = []
    for k in range(steps):
        if els[k] != el:
            old_nodes = set(int(v) for v in m.IX[:, el - 1])
            om.retarget_pull(m, po, els[k], speed)
            o = fem.OracleMesh(m, po)
            inh = {(int(nd), m.dofs[u]) for (u, nd, _v) in m.inh_dir_bcs}
            for node in old_nodes | set(int(v) for v in m.IX[:, els[k] - 1]):
                for d in range(1, m.ndf + 1):
                    if m.ID[d - 1, node - 1] == 0:
                        co[node - 1, d - 1] = speed if (node, d) in inh else 0.0
            el = els[k]
        fem.update_xms(po.motion, xo, co, DT, m.dofs)
        its_o.append(len(fem.time_step(o, xo, co, (k + 1) * DT, DT)))
    assert its == its_o
    assert np.linalg.norm(c - co) / np.linalg.norm(co) <= 1e-8
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8
