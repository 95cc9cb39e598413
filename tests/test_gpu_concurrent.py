"""Contexts advanced concurrently, each on its own host thread and stream.

The library shares nothing between contexts and every kernel's result is
independent of scheduling, so a patch's trajectory must be bit-identical
whether it runs alone or beside others.  This also guards the factorization's
in-launch ordering: under contention from other streams, a block that reads a
tile late must still see the value the algorithm defines (a tile permuted in
place by a later step while still an operand once broke this).
"""
import threading

import numpy as np
import pytest

import mafem
from mafem.patches import Patch, PatchSpec

pytestmark = pytest.mark.gpu

DT = 1e-3
SPECS = [PatchSpec(mafem.ALEVB, 40, 0.8, 0.5, 1.5, 11, 200),
         PatchSpec(mafem.LAG, 48, 2.0, 0.3, 2.0, 12, 200),
         PatchSpec(mafem.EUL, 32, 1.0, 1.0, 1.0, 13, 200),
         PatchSpec(mafem.ALEVB, 36, 0.4, 2.0, 0.8, 14, 200),
         PatchSpec(mafem.LAG, 56, 0.5, 0.6, 1.2, 15, 200),
         PatchSpec(mafem.ALEVB, 44, 3.0, 0.2, 2.5, 16, 200)]
STEPS = 3


def _run(spec, out, key):
    pt = Patch(spec)
    try:
        pt.dev.set_solver("direct", precision=64, rtol=1e-13)
        its = []
        for k in range(STEPS):
            pt.set_pull(spec.speed_at(k))
            pt.dev.predict(DT)
            rc, eps = pt.dev.time_step((k + 1) * DT, DT)
            assert rc == 0, (spec, k, eps)
            its.append(eps)
        out[key] = (its, pt.dev.get_state())
    except Exception as e:  # reported by the test thread
        out[key] = e
    finally:
        pt.close()


def test_concurrent_contexts_match_alone():
    print("contexts alone, then six at once on their own threads", flush=True)
    out = {}
    for i, sp in enumerate(SPECS):
        _run(sp, out, ("alone", i))
    ths = [threading.Thread(target=_run, args=(sp, out, ("conc", i))) for i, sp in enumerate(SPECS)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    for i in range(len(SPECS)):
        for tag in ("alone", "conc"):
            assert not isinstance(out[(tag, i)], Exception), out[(tag, i)]
        (ea, (xa, ca)), (ec, (xc, cc)) = out[("alone", i)], out[("conc", i)]
        assert ea == ec, (SPECS[i], ea, ec)
        assert np.array_equal(xa, xc) and np.array_equal(ca, cc), SPECS[i]
