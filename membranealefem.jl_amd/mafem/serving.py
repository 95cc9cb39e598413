"""Serving loop: a stream of membrane simulations advanced by continuous
batching.

Each request is one patch (size, formulation and parameters of its own) to be
advanced a number of backward-Euler steps (the reference's `solve` for one
mesh, src/MembraneAleFem.jl:44-60, with its time loop src/Analysis.jl:17-102).
Requests are admitted while the device-memory estimate of the members fits the
HBM budget; a host thread builds the admitted requests' contexts (mesh tables,
state upload) while the device advances the current members, and new members
join at the next step boundary.  All members advance together through one
batch step (mafem_batch_time_step), each at its own time index; a member that
finished its steps leaves, its result (final xms, cps and the Newton history)
is kept and its device memory is released for the next admissions.  A
member's trajectory does not depend on the others in its batch (every batched
kernel works on one patch's data; factorizations and refactor decisions are
per patch), which the tests check bit for bit.
"""
import collections
import ctypes
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .analysis import Batch
from .patches import Patch, PatchSpec, device_bytes


@dataclass
class Request:
    rid: int
    spec: PatchSpec
    steps: int = 50
    arrival: float = 0.0       # seconds after the start of the run
    admitted: float = None
    started: float = None
    finished: float = None
    iters: list = field(default_factory=list)   # Newton iterations per step
    eps: list = field(default_factory=list)     # ε history per step
    state: tuple = None        # final (xms, cps)
    failed: int = -1           # step that did not converge, -1 = none
    error: object = None       # exception raised while serving it


def device_memory():
    """(free, total) bytes of the current device."""
    lib = L.load()
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    rc = lib.mafem_device_memory(ctypes.byref(f), ctypes.byref(t))
    if rc != L.OK:
        raise L.LibError(rc, "mafem_device_memory")
    return f.value, t.value


class Server:
    """Continuous-batching request loop over one device.

    Defaults from measurements on the serving mix (n from 32 to 128): batched
    members factor afresh in FP64 for every Newton iteration (so each GMRES
    solve takes one or two steps; the tangents of large centre-refined meshes
    are too ill-conditioned for an FP32 factorization to pay off), and
    requests of 96x96 elements or more run solo on their own streams."""

    def __init__(self, hbm_budget=None, precision=64, dt=1e-3, max_members=64, use_graphs=True,
                 keep_states=True, solo_min_elements=96 * 96, max_solo=3, refactor_iters=-1):
        if hbm_budget is None:
            free, _ = device_memory()
            hbm_budget = int(0.8 * free)
        self.budget = int(hbm_budget)
        self.precision = precision
        self.dt = dt
        self.max_members = max_members
        self.use_graphs = use_graphs
        self.keep_states = keep_states
        # Large requests run "solo": each on its own host thread and stream with
        # the direct FP64 path (fresh factorization every Newton iteration),
        # overlapping on the device with the batch of the others; the longest
        # factorization and sweep chains then no longer gate every batch step.
        self.refactor_iters = refactor_iters
        self.solo_min_elements = solo_min_elements
        self.max_solo = max_solo
        self.step_latency = []     # wall time of every batch step (s), one entry per member-step
        self.step_members = []     # members of every batch step
        self.newton_iters = 0
        self.rebuilds = 0
        self.rebuild_time = 0.0
        self.peak_bytes = 0
        self.solo_count = 0

    def _bytes(self, r):
        # FP64 band storage in the estimate (a member may fall back to it);
        # requests large enough to run solo factor with nested dissection
        solo = r.spec.n * r.spec.n >= self.solo_min_elements
        return int(1.15 * device_bytes(r.spec, precision=64, ordering="nd" if solo else "band"))

    def run(self, requests, time_limit=None, log=None):
        """Serve `requests` (sorted by arrival); returns them with results."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        todo = queue.Queue()
        ready = queue.Queue()

        def builder():
            while True:
                r = todo.get()
                if r is None:
                    return
                try:
                    ready.put((r, Patch(r.spec)))
                except Exception as e:  # surfaced on the serving thread
                    ready.put((r, e))

        worker = threading.Thread(target=builder, daemon=True)
        worker.start()
        used, building = 0, 0
        members = []               # [request, patch, step index]
        solos = []                 # [thread, request, patch]
        lock = threading.Lock()    # guards the statistics the solo threads update
        t0 = time.perf_counter()

        def solo(r, pt):
            try:
                pt.dev.set_solver("direct", precision=64)
                pt.dev.set_ordering("nd")
                for k in range(r.steps):
                    pt.set_pull(pt.spec.speed_at(k))
                    ts = time.perf_counter()
                    pt.dev.predict(self.dt)
                    rc, eps = pt.dev.time_step((k + 1) * self.dt, self.dt)
                    el = time.perf_counter() - ts
                    r.iters.append(len(eps))
                    r.eps.append(eps)
                    with lock:
                        self.step_latency.append(el)
                        if rc == 0:
                            self.newton_iters += len(eps)
                    if rc != 0:
                        r.failed = k
                        break
                if self.keep_states:
                    r.state = pt.dev.get_state()
            except Exception as e:  # reported by the serving thread
                r.error = e
            r.finished = time.perf_counter() - t0
        batch, changed = None, False
        try:
            while pending or building or members or solos:
                # solo requests that finished release their memory
                for sl in [x for x in solos if not x[0].is_alive()]:
                    th, r, pt = sl
                    solos.remove(sl)
                    if getattr(r, "error", None) is not None:
                        raise r.error
                    pt.close()
                    used -= self._bytes(r)
                now = time.perf_counter() - t0
                if time_limit is not None and now > time_limit:
                    break
                # admission: arrived requests that fit the budget
                while (pending and pending[0].arrival <= now
                       and len(members) + len(solos) + building < self.max_members
                       and used + self._bytes(pending[0]) <= self.budget):
                    r = pending.popleft()
                    r.admitted = now
                    used += self._bytes(r)
                    self.peak_bytes = max(self.peak_bytes, used)
                    building += 1
                    todo.put(r)
                # members whose contexts are ready join at this step boundary
                block = not members and building > 0
                while building:
                    try:
                        r, pt = ready.get(block=block, timeout=0.5 if block else None) \
                            if block else ready.get_nowait()
                    except queue.Empty:
                        break
                    building -= 1
                    block = False
                    if isinstance(pt, Exception):
                        raise pt
                    r.started = time.perf_counter() - t0
                    if (pt.mesh.numel >= self.solo_min_elements
                            and len(solos) < self.max_solo):
                        th = threading.Thread(target=solo, args=(r, pt), daemon=True)
                        solos.append([th, r, pt])
                        th.start()
                        self.solo_count += 1
                        continue
                    members.append([r, pt, 0])
                    changed = True
                if not members:
                    if building:
                        continue
                    if pending and pending[0].arrival > now:
                        time.sleep(min(0.01, pending[0].arrival - now))
                    elif pending and used + self._bytes(pending[0]) > self.budget and not solos:
                        raise RuntimeError(f"request {pending[0].rid} does not fit the budget")
                    elif solos:
                        time.sleep(0.005)  # solo requests hold the device memory meanwhile
                    continue
                if changed:
                    if batch is not None:
                        batch.close()
                    tb0 = time.perf_counter()
                    batch = Batch([m[1].dev for m in members], precision=self.precision,
                                  use_graphs=self.use_graphs, refactor_iters=self.refactor_iters)
                    self.rebuilds += 1
                    self.rebuild_time += time.perf_counter() - tb0
                    changed = False
                times = np.array([(m[2] + 1) * self.dt for m in members])
                for r, pt, k in members:
                    pt.set_pull(pt.spec.speed_at(k))
                ts = time.perf_counter()
                it, eps, st = batch.time_step(times, self.dt, predict=True)
                el = time.perf_counter() - ts
                with lock:
                    self.step_latency += [el] * len(members)
                self.step_members.append(len(members))
                if log is not None:
                    log(f"t={time.perf_counter() - t0:.2f}s step of {len(members)} members "
                        f"{el:.3f}s, {int(it.sum())} Newton iterations, pending {len(pending)}, "
                        f"building {building}")
                keep = []
                for j, m in enumerate(members):
                    r, pt, k = m
                    r.iters.append(int(it[j]))
                    r.eps.append(eps[j])
                    if st[j] == 0:
                        with lock:
                            self.newton_iters += int(it[j])
                    m[2] = k + 1
                    done = st[j] != 0 or m[2] >= r.steps
                    if st[j] != 0:
                        r.failed = k
                    if done:
                        if self.keep_states:
                            r.state = pt.dev.get_state()
                        r.finished = time.perf_counter() - t0
                        pt.close()
                        used -= self._bytes(r)
                        changed = True
                    else:
                        keep.append(m)
                members = keep
        finally:
            if batch is not None:
                batch.close()
            for r, pt, k in members:
                pt.close()
            for th, r, pt in solos:
                th.join()
                pt.close()
            todo.put(None)
            worker.join(timeout=60)
            while not ready.empty():
                r, pt = ready.get_nowait()
                if not isinstance(pt, Exception):
                    pt.close()
        self.wall = time.perf_counter() - t0
        return requests


def poisson_requests(specs, rate=200.0, steps=50, seed=0xC5):
    """Requests with exponential inter-arrival times (a Poisson stream)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for k, sp in enumerate(specs):
        t += float(rng.exponential(1.0 / rate))
        out.append(Request(k, sp, steps=steps, arrival=t))
    return out
