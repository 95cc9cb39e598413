"""Serving loop (continuous batching over contexts admitted by HBM budget).

Bars:
  * every request's final state is bit-identical whether it was served alone
    or in a mix whose membership changed while it ran (arrivals and
    departures at different steps);
  * each served result agrees with the plain FP64 direct path on the same
    patch within 1e-8 relative, with the same Newton iteration counts;
  * a budget that holds two requests keeps at most two members on the device
    and still serves every request.
"""
import numpy as np
import pytest

import mafem
from mafem.patches import Patch, PatchSpec
from mafem.serving import Request, Server, device_memory

pytestmark = pytest.mark.gpu

SPECS = [PatchSpec(mafem.LAG, 12, 1.0, 1.0, 0.5, 1), PatchSpec(mafem.EUL, 10, 2.0, 0.5, 0.4, 2),
         PatchSpec(mafem.ALEVB, 16, 0.5, 2.0, 0.6, 3), PatchSpec(mafem.ALEVB, 9, 1.0, 1.0, 0.3, 4)]
STEPS = [5, 7, 4, 6]
ARRIVALS = [0.0, 0.0, 0.2, 0.4]


def requests():
    return [Request(k, s, steps=n, arrival=a)
            for k, (s, n, a) in enumerate(zip(SPECS, STEPS, ARRIVALS))]


def plain(spec, steps):
    pt = Patch(spec)
    pt.dev.set_solver("direct", precision=64, rtol=1e-13)
    pt.dev.set_ordering("nd")  # as the server's solo requests
    its = []
    for k in range(steps):
        pt.set_pull(spec.speed_at(k))
        pt.dev.predict(1e-3)
        rc, eps = pt.dev.time_step((k + 1) * 1e-3, 1e-3)
        assert rc == 0
        its.append(len(eps))
    x, c = pt.dev.get_state()
    pt.close()
    return its, x, c


def test_device_memory_query():
    free, total = device_memory()
    assert 0 < free <= total


def test_served_results_alone_and_in_mix():
    print("serving: 4 requests in a changing mix vs alone", flush=True)
    srv = Server(max_members=8)
    mixed = srv.run(requests())
    assert srv.rebuilds >= 3          # membership changed while requests ran
    for k, r in enumerate(mixed):
        assert r.failed < 0 and r.state is not None
        alone = Server(max_members=1).run([Request(k, SPECS[k], steps=STEPS[k])])[0]
        assert np.array_equal(alone.state[0], r.state[0]) and np.array_equal(alone.state[1], r.state[1]), k
        assert alone.iters == r.iters
        its, x, c = plain(SPECS[k], STEPS[k])
        assert r.iters == its, k
        assert np.linalg.norm(r.state[1] - c) / np.linalg.norm(c) <= 1e-8
        assert np.linalg.norm(r.state[0] - x) / np.linalg.norm(x) <= 1e-8


def test_budget_limits_members():
    print("serving: budget that fits the largest request alone", flush=True)
    probe = Server()
    sizes = sorted(probe._bytes(r) for r in requests())
    budget = sizes[-1]
    k_max = int(np.searchsorted(np.cumsum(sizes), budget, side="right"))
    assert k_max < len(sizes)         # the budget binds
    srv = Server(hbm_budget=budget, max_members=8)
    out = srv.run(requests())
    assert srv.peak_bytes <= budget
    assert max(srv.step_members) <= k_max
    assert all(r.failed < 0 and r.state is not None for r in out)


def test_solo_requests_match_plain_path():
    """Requests above the solo size run on their own thread and stream with
    the direct FP64 path, overlapping the batch of the others: their results
    are the plain path's, bit for bit; the batched ones still match alone."""
    print("serving: solo and batched requests together", flush=True)
    srv = Server(max_members=8, solo_min_elements=12 * 12, max_solo=2)
    out = srv.run(requests())
    assert srv.solo_count >= 1
    for k, r in enumerate(out):
        assert r.failed < 0 and r.state is not None
        its, x, c = plain(SPECS[k], STEPS[k])
        assert r.iters == its, k
        if SPECS[k].n ** 2 >= 12 * 12:
            assert np.array_equal(r.state[1], c) and np.array_equal(r.state[0], x), k
        else:
            assert np.linalg.norm(r.state[1] - c) / np.linalg.norm(c) <= 1e-8
