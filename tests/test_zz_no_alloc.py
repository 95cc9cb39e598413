"""mafem_pull_force_noalloc against mafem_pull_force, through the C ABI.

The allocation-free variant keeps its maps, element residuals and outputs in
a fixed array the library holds on the device instead of buffers it
allocates; both run the same kernels, so their results are equal bit for bit.
Two calls in a row on one device, the second on a smaller patch (fewer
residual rows per element), check that the shared array carries nothing
over from the previous call.
"""
import ctypes

import numpy as np
import pytest

import mafem
from mafem import _lib as L
from mafem.analysis import DeviceMesh

pytestmark = pytest.mark.gpu


def _device(motion, n, speed, seed):
    p = mafem.Params(motion=motion, num1el=n, num2el=n, length=1.0, enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=speed)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=speed)
    rng = np.random.default_rng(seed)
    xms[:, 2] += rng.normal(0, 1e-3, mesh.numnp)
    cps += np.where(mesh.ID.T != 0, rng.normal(0, 1e-2, cps.shape), 0.0)
    dev = DeviceMesh(mesh, p)
    dev.set_state(xms, cps)
    return dev


def _noalloc(dev):
    x = np.zeros(3)
    f = np.zeros(3)
    rc = dev.lib.mafem_pull_force_noalloc(dev.h, L.ptr(x, L.f64p), L.ptr(f, L.f64p))
    assert rc == L.OK, dev.lib.mafem_last_error(dev.h).decode()
    return x, f


def test_noalloc_pull_force_matches_allocating_entry_point():
    print("no-alloc pull force: ALEVB 16x16 then LAG 9x9", flush=True)
    big = _device(mafem.ALEVB, 16, 0.5, 1)       # 72 residual rows per element
    small = _device(mafem.LAG, 9, 0.3, 2)        # 36 rows, a centred pulled element
    xb, fb = _noalloc(big)
    xs, fs = _noalloc(small)
    xb_ref, fb_ref = big.pull_force()
    xs_ref, fs_ref = small.pull_force()
    assert np.array_equal(xb, xb_ref) and np.array_equal(fb, fb_ref)
    assert np.array_equal(xs, xs_ref) and np.array_equal(fs, fs_ref)
    assert np.linalg.norm(fb) > 0 and np.linalg.norm(fs) > 0
    # the big patch again, after the small one used the shared array
    xb2, fb2 = _noalloc(big)
    assert np.array_equal(xb2, xb_ref) and np.array_equal(fb2, fb_ref)
    big.close()
    small.close()


def test_noalloc_pull_force_rejects_what_the_allocating_one_rejects():
    print("no-alloc pull force: STATIC scenario rejected", flush=True)
    p = mafem.Params(motion=mafem.STATIC, scenario=mafem.F_COUE, num1el=4, num2el=4,
                     length=1.0, output=False)
    mesh = mafem.make_mesh(p)
    dev = DeviceMesh(mesh, p)
    x = np.zeros(3)
    f = np.zeros(3)
    rc = dev.lib.mafem_pull_force_noalloc(dev.h, L.ptr(x, L.f64p), L.ptr(f, L.f64p))
    assert rc == L.EINVAL
    rc = dev.lib.mafem_pull_force_noalloc(dev.h, None, None)
    assert rc == L.EINVAL
    dev.close()
