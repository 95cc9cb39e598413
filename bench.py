"""Benchmark of the per-time-step hot path on one MI355X.

Workload (the plain FP64 GPU configuration): flat square patch of side 1,
64 x 64 biquadratic B-spline elements, ALE (ALEVB) formulation, kb = 1,
ζv = 1, kg = -0.5, tether pull on the pulled element ramped linearly to 2.0
over 200 steps of dt = 1e-3, seeded N(0, 1e-6) out-of-plane perturbation of
the control points, Newton tolerance 1e-10.  (With N(0, 1e-4) the reference
algorithm's own Newton iteration does not converge from the perturbed state on
this centre-refined mesh, whose smallest elements are ~0.004 wide; 1e-6 is the
largest decade from which it converges, in 4 iterations.)  One bench step = one
backward-Euler time step (predictor + Newton to tolerance, every iteration
assembling r and K on the device and solving K du = -r).

Prints one JSON line.  `value` = Newton iterations per second (each one
assembly + direct solve + update); the line also carries element assemblies
per second, time steps per second, p50/p99 time-step latency, the roofline of
the dominant kernel and the CPU baseline (the oracle restatement timed on the
host).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "membranealefem.jl_amd"))
sys.path.insert(0, ROOT)

METRIC = "elem-assemblies/s + Newton steps/s at 1e-8 rel err vs FP64; p50/p99 step latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
TILE = 64


def build_case(n, motion_name, seed=0xC2, sigma=1e-6):
    import mafem
    motion = getattr(mafem, motion_name)
    p = mafem.Params(motion=motion, num1el=n, num2el=n, length=1.0, kb=1.0, zv=1.0, kg=-0.5,
                     enr=1e-10, output=False)
    mesh = mafem.make_mesh(p, pull_speed=0.0)
    xms, cps = mafem.prepare_input(p, mesh, pull_speed=0.0)
    rng = np.random.default_rng(seed)
    xms[:, 2] += rng.normal(0.0, sigma, mesh.numnp)
    return p, mesh, xms, cps


def pull_speed(k, total=200, final=2.0):
    return final * (k + 1) / total


# HBM counters per dispatch of the factorization kernels the rooflines report
# (FETCH_SIZE, WRITE_SIZE in KiB, mean over one time step's dispatches of the
# 64x64 ALEVB bench patch), by kernel; see the rooflines below.  None: not
# measured.
# profiles/k_step_la_hbm_traffic_pmc.txt and profiles/nd_hbm_traffic_pmc.txt
# hold the runs.
STEP_COUNTERS_KIB = {
    "k_step_la": (16013.467625473484, 19804.575757575756),
    "knd_step_la": (3675.1130514705883, 3150.5367647058824),
    "knd_schur": (157148.615234375, 93548.0),
}
# the same counters per dispatch of the area kernel of the bench patch, by motion
AREA_COUNTERS_KIB = {
    "ALEVB": (5800.09375, 168459.0),
}


def band_update_flops(dev):
    """Algorithmic FLOPs of the band-LU trailing updates of one factorization:
    2 T^3 per tile update, over the (tb x tb) window clipped at the end."""
    nt, tb = dev_band_shape(dev)
    tiles = 0
    for K in range(nt):
        w = min(tb, nt - 1 - K)
        tiles += w * w
    return 2.0 * TILE ** 3 * tiles, nt


def dev_band_shape(dev):
    m = dev.mesh
    nx = m.numnp * m.ndf
    nt = (nx + TILE - 1) // TILE
    kl = (2 * m.num1np + 3) * m.ndf - 1
    tb = min((kl + TILE - 1) // TILE, max(1, nt - 1))
    return nt, tb


def cpu_baseline(p_gpu, mesh, xms, cps, threads, dev=None):
    """One Newton iteration of the oracle (complex-step assembly with OpenMP
    over elements + SuperLU) on the same 64^2 state: the bounded sample.
    With `dev`, the device assembles and solves the same state (outside any
    timed region) and the relative differences of r, K and du from the
    oracle's are returned beside the baseline."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import scipy.sparse.linalg as spla

    from oracle import fem, mesh as om
    po = om.Params(motion=int(p_gpu.motion), num1el=p_gpu.num1el, num2el=p_gpu.num2el,
                   length=1.0, enr=p_gpu.enr)
    o, _, _ = fem.build(po, pull_speed=0.0)
    t0 = time.perf_counter()
    r, K = fem.calc_r_K(o, xms, cps, 1e-3, 1e-3)
    t1 = time.perf_counter()
    du = spla.spsolve(K.tocsc(), -r)
    t2 = time.perf_counter()
    base = {"value": 1.0 / (t2 - t0), "unit": "Newton steps/s", "cores": threads,
            "kind": "port",
            "sample": (f"1 Newton iteration at the 64x64 ALEVB bench state: oracle complex-step "
                       f"assembly {t1 - t0:.2f} s on {threads} OpenMP threads + scipy SuperLU "
                       f"solve {t2 - t1:.2f} s (1 thread); elem-assemblies/s "
                       f"{o.mesh.numel / (t1 - t0):.1f}")}
    vs = None
    if dev is not None:
        dev.set_state(xms, cps)
        r_g, K_g = dev.assemble(1e-3, 1e-3)
        du_g, relres = dev.solve(-r_g)
        dK = (K_g - K).tocsr()
        vs = {"state": "bench initial state, t = dt = 1e-3",
              "r_rel": float(np.linalg.norm(r_g - r) / np.linalg.norm(r)),
              "K_rel_fro": float(np.sqrt((dK.multiply(dK)).sum()) /
                                 np.sqrt((K.multiply(K)).sum())),
              "du_rel": float(np.linalg.norm(du_g - du) / np.linalg.norm(du)),
              "solve_relres": float(relres)}
    return base, vs


def batched_flops(patches):
    """Algorithmic FLOPs of one factorization of every patch (tile updates)."""
    tot = 0.0
    for pt in patches:
        nt, tb = dev_band_shape(pt.dev)
        tot += 2.0 * TILE ** 3 * sum(min(tb, nt - 1 - K) ** 2 for K in range(nt))
    return tot


def bench_batched_solo(args):
    """Non-lockstep form of the batched workload: the same patches, each
    advanced through all its steps (warm-up and timed alike) on the
    single-context direct FP64 path by one of N host threads, largest patches
    first.  Each context has its own stream and the library releases the
    interpreter lock inside its calls, so up to N patches' launches overlap on
    the device and no patch waits for another's Newton iteration.  `value` =
    Newton iterations of all patches / wall time; each patch's results are
    those of its one-at-a-time run (contexts share nothing; every sweep split
    gives bit-identical solutions).  A patch's context is closed when it is
    done, so at most N bands are held at once."""
    import queue
    import threading
    from mafem.patches import Patch, batched_specs

    specs = batched_specs(args.patches, seed=0xC4, sizes=(args.nmin, args.nmax))
    t0 = time.perf_counter()
    patches = []
    for k, sp in enumerate(specs):
        patches.append(Patch(sp))
        if (k + 1) % 64 == 0:
            print(f"bench: {k + 1}/{len(specs)} patches on the device", file=sys.stderr, flush=True)
    setup_s = time.perf_counter() - t0
    dt = 1e-3
    nsteps = args.warmup + args.steps
    todo = queue.Queue()
    for i in sorted(range(len(specs)), key=lambda i: -specs[i].n):
        todo.put(i)
    iters = np.zeros(len(specs), dtype=np.int64)
    lat, failed, errors = [], [], []
    lock = threading.Lock()
    done = [0]

    def worker():
        while True:
            try:
                i = todo.get_nowait()
            except queue.Empty:
                return
            pt = patches[i]
            try:
                pt.dev.set_solver("direct", precision=64, rtol=1e-13)
                pt.dev.set_ordering(args.ordering)
                for k in range(nsteps):
                    pt.set_pull(pt.spec.speed_at(k))
                    ts = time.perf_counter()
                    pt.dev.predict(dt)
                    rc, eps = pt.dev.time_step((k + 1) * dt, dt)
                    el = time.perf_counter() - ts
                    with lock:
                        lat.append(el)
                    if rc != 0:
                        with lock:
                            failed.append({"patch": int(i), "step": k, "n": specs[i].n,
                                           "motion": int(specs[i].motion), "kb": specs[i].kb,
                                           "zv": specs[i].zv, "speed": specs[i].speed})
                        break
                    iters[i] += len(eps)
            except Exception as e:  # reported below
                with lock:
                    errors.append((i, e))
            finally:
                pt.close()
            with lock:
                done[0] += 1
                if done[0] % 64 == 0:
                    print(f"bench: {done[0]}/{len(specs)} patches done", file=sys.stderr,
                          flush=True)

    threads = [threading.Thread(target=worker) for _ in range(args.solo_threads)]
    t_start = time.perf_counter()
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    t_total = time.perf_counter() - t_start
    if errors:
        raise RuntimeError(f"patch {errors[0][0]}: {errors[0][1]}")
    total = int(iters.sum())
    numel = np.array([sp.n * sp.n for sp in specs])
    res = {
        "metric": METRIC,
        "value": total / t_total,
        "unit": "Newton steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / nsteps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(specs)} independent F_PULL patches, n ~ U{{{args.nmin}.."
                               f"{args.nmax}}}, kb and zv log-uniform in [0.1, 10], pull ramped "
                               f"to U[0.2, 3] over 200 steps, motion U{{LAG, ALEVB}}, seed 0xC4, "
                               f"flat initial state, dt=1e-3; {args.solo_threads} host threads, "
                               f"one stream each, direct FP64 path per patch, {nsteps} steps "
                               f"per patch (all timed)",
                   "patches": len(specs), "elements": int(numel.sum()),
                   "threads": args.solo_threads, "time_steps": nsteps, "newton_iters": total},
        "elem_assemblies_per_s": float((iters * numel).sum()) / t_total,
        "step_latency_ms_p50": 1e3 * float(np.percentile(lat, 50)),
        "step_latency_ms_p99": 1e3 * float(np.percentile(lat, 99)),
        "failed_patches": failed,
        "setup_s": setup_s,
    }
    print(json.dumps(res), flush=True)


def bench_batched(args):
    """Batched form: many independent patches (sizes, formulations and
    parameters drawn per patch) advanced one time step at a time through one
    batch; FP64 assembly, residual, update and ε test, FP64 GMRES with FP32
    band-LU preconditioner.  Beside it, a stratified sample of the same
    patches runs alone on the plain FP64 direct path (one context, one patch
    at a time) for the same steps: its Newton throughput is the plain
    per-patch rate, and its final states give the error of the batched run."""
    import mafem
    from mafem.patches import Patch, batched_specs

    specs = batched_specs(args.patches, seed=0xC4, sizes=(args.nmin, args.nmax))
    t0 = time.perf_counter()
    patches = []
    for k, sp in enumerate(specs):
        patches.append(Patch(sp))
        if (k + 1) % 64 == 0:
            print(f"bench: {k + 1}/{len(specs)} patches on the device", file=sys.stderr, flush=True)
    if args.refactor_iters is None:  # fresh factors each Newton iteration (measured best on C4)
        args.refactor_iters = 0
    batch = mafem.Batch([pt.dev for pt in patches], precision=args.precision,
                        refactor_iters=args.refactor_iters, use_graphs=not args.no_graphs)
    setup_s = time.perf_counter() - t0
    dt = 1e-3

    # A patch whose Newton iteration does not converge (the reference raises
    # AssertionError for it) is a failed simulation: it is reported and left
    # out of later steps; only converged iterations are counted.
    alive = np.ones(len(patches), dtype=np.uint8)
    failed = []

    def step(k):
        for i, pt in enumerate(patches):
            if alive[i]:
                pt.set_pull(pt.spec.speed_at(k))
        it, eps, st = batch.time_step((k + 1) * dt, dt, predict=True, active=alive)
        for i in np.where(st == 1)[0]:
            failed.append({"patch": int(i), "step": k, "n": specs[i].n,
                           "motion": int(specs[i].motion), "kb": specs[i].kb, "zv": specs[i].zv,
                           "speed": specs[i].speed, "eps": [float(e) for e in eps[i][:4]]})
            alive[i] = 0
            print(f"bench: patch {i} did not converge at step {k} (reported, retired)",
                  file=sys.stderr, flush=True)
        return np.where(st == 0, it, 0), batch.stats()

    for k in range(args.warmup):
        _, st = step(k)
        print(f"bench: warm-up step {k} done ({st['patch_factors']} factorizations, "
              f"{st['fp64_patches']} FP64 patches)", file=sys.stderr, flush=True)
    lat, iters, elems = [], 0, 0
    agg = {"assemble_us": 0.0, "factor_us": 0.0, "solve_us": 0.0, "update_us": 0.0,
           "patch_factors": 0, "gmres_iters": 0, "factor_passes": 0, "gmres_cycles": 0,
           "hard_failures": 0}
    numel = np.array([pt.mesh.numel for pt in patches])
    t_start = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        ts = time.perf_counter()
        it, st = step(k)
        lat.append(time.perf_counter() - ts)
        iters += int(it.sum())
        elems += int((it * numel).sum())
        for key in agg:
            agg[key] += st[key]
        fp64_patches = st["fp64_patches"]
        print(f"bench: step {k} {lat[-1]:.3f} s, {int(it.sum())} Newton iterations, "
              f"{st['patch_factors']} factorizations, {st['gmres_iters']} GMRES steps, "
              f"{fp64_patches} FP64 patches", file=sys.stderr, flush=True)
    t_total = time.perf_counter() - t_start
    states = {}
    live = [i for i in np.argsort([sp.n for sp in specs], kind="stable") if alive[i]]
    sample = [int(live[i]) for i in np.linspace(0, len(live) - 1, args.sample).astype(int)]
    for i in sample:
        states[i] = patches[i].dev.get_state()
    fact_flops = batched_flops(patches)
    batch.close()
    for pt in patches:
        pt.close()

    # plain FP64 direct path, one patch at a time, same steps
    plain_iters, plain_time, errs = 0, 0.0, []
    for i in sample:
        pt = Patch(specs[i])
        pt.dev.set_solver("direct", precision=64, rtol=1e-13)
        pt.dev.set_ordering(args.ordering)
        for k in range(args.warmup + args.steps):
            pt.set_pull(pt.spec.speed_at(k))
            ts = time.perf_counter()
            pt.dev.predict(dt)
            rc, eps = pt.dev.time_step((k + 1) * dt, dt)
            el = time.perf_counter() - ts
            if rc != 0:
                raise RuntimeError(f"plain path: patch {i} did not converge at step {k}")
            if k >= args.warmup:
                plain_iters += len(eps)
                plain_time += el
        x, c = pt.dev.get_state()
        xb, cb = states[i]
        errs.append(max(np.linalg.norm(xb - x) / np.linalg.norm(x),
                        np.linalg.norm(cb - c) / np.linalg.norm(c)))
        pt.close()
    plain_rate = plain_iters / plain_time
    value = iters / t_total
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "Newton steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FP32 band-LU preconditioner)" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(specs)} independent F_PULL patches, n ~ U{{{args.nmin}.."
                               f"{args.nmax}}}, kb and zv log-uniform in [0.1, 10], pull ramped "
                               f"to U[0.2, 3] over 200 steps, motion U{{LAG, ALEVB}}, seed 0xC4, "
                               f"flat initial state, dt=1e-3, one batch"
                               f"{'' if args.no_graphs else ', HIP graph replay'}",
                   "patches": len(specs), "elements": int(numel.sum()),
                   "band_lu_precision": args.precision, "refactor_iters": args.refactor_iters,
                   "time_steps": args.steps, "newton_iters": iters},
        "elem_assemblies_per_s": elems / t_total,
        "time_steps_per_s": args.steps / t_total,
        "step_latency_ms_p50": 1e3 * float(np.percentile(lat, 50)),
        "step_latency_ms_p99": 1e3 * float(np.percentile(lat, 99)),
        "phase_ms_per_step": {k: agg[k] * 1e-3 / args.steps for k in
                              ("assemble_us", "factor_us", "solve_us", "update_us")},
        "patch_factorizations": agg["patch_factors"], "factor_passes": agg["factor_passes"],
        "gmres_iters": agg["gmres_iters"], "gmres_cycles": agg["gmres_cycles"],
        "hard_failures": agg["hard_failures"], "fp64_patches": fp64_patches,
        "failed_patches": failed,
        # a pass factors only the members whose preconditioner is refreshed:
        # the mean algorithmic FLOPs of a member's factorization times the
        # member factorizations, over the factorization time (FP32 or FP64
        # band LU per the member's storage)
        "factorization_tflops": (fact_flops / len(patches) * agg["patch_factors"] /
                                 (agg["factor_us"] * 1e-6) / 1e12) if agg["factor_us"] else None,
        "plain_fp64": {"newton_per_s": plain_rate, "sample": len(sample),
                       "path": "single-context direct FP64, one patch at a time",
                       "ordering": args.ordering},
        "speedup_vs_plain_fp64": value / plain_rate,
        "rel_err_vs_fp64_max": max(errs),
        "setup_s": setup_s,
    }
    print(json.dumps(res), flush=True)


def bench_serving(args):
    """Serving form: a Poisson stream of requests (each a patch advanced
    `--req-steps` steps; n from {32, 48, 64, 96, 128}, formulation uniform
    over LAG, EUL, ALEVB, parameters as in the batched form) served by
    continuous batching under an HBM budget.  Reports per-step latency
    percentiles over every member-step, request latency, aggregate Newton
    throughput, and the plain FP64 direct path's Newton throughput on a
    sample of the same requests (a bounded number of steps each)."""
    import mafem
    from mafem.patches import Patch, serving_specs
    from mafem.serving import Server, poisson_requests

    specs = serving_specs(args.requests, seed=0xC5)
    reqs = poisson_requests(specs, rate=args.rate, steps=args.req_steps, seed=0xC5)
    srv = Server(precision=args.precision if args.precision else 64, max_members=args.max_members,
                 solo_min_elements=args.solo_min_n ** 2, max_solo=args.max_solo,
                 refactor_iters=args.refactor_iters if args.refactor_iters is not None else -1)
    out = srv.run(reqs, log=lambda m: print("serve: " + m, file=sys.stderr, flush=True))
    done = [r for r in out if r.finished is not None and r.failed < 0]
    failed = [{"rid": r.rid, "n": r.spec.n, "motion": int(r.spec.motion), "step": r.failed}
              for r in out if r.failed >= 0]
    lat = np.array(srv.step_latency)
    req_lat = np.array([r.finished - r.arrival for r in done])
    queue_wait = np.array([r.started - r.arrival for r in done])
    value = srv.newton_iters / srv.wall

    # plain FP64 direct path on a sample of the requests, a few steps each
    plain_iters, plain_time = 0, 0.0
    idx = np.linspace(0, len(specs) - 1, min(args.sample, len(specs))).astype(int)
    for i in idx:
        pt = Patch(specs[i])
        pt.dev.set_solver("direct", precision=64, rtol=1e-13)
        pt.dev.set_ordering(args.ordering)
        for k in range(args.plain_steps):
            pt.set_pull(specs[i].speed_at(k))
            ts = time.perf_counter()
            pt.dev.predict(1e-3)
            rc, eps = pt.dev.time_step((k + 1) * 1e-3, 1e-3)
            plain_time += time.perf_counter() - ts
            plain_iters += len(eps)
            if rc != 0:
                break
        pt.close()
    plain_rate = plain_iters / plain_time
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "Newton steps/s",
        "n_gpus": 1,
        "steps": int(sum(len(r.iters) for r in out)),
        "warmup": 0,
        "ms_per_step": 1e3 * float(np.mean(lat)) if len(lat) else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"serving: {len(reqs)} requests, Poisson arrivals at {args.rate}/s, "
                               f"{args.req_steps} steps each, n in {{32,48,64,96,128}}, motion "
                               f"U{{LAG, EUL, ALEVB}}, kb and zv log-uniform in [0.1, 10], pull "
                               f"ramped to U[0.2, 3] over 200 steps, seed 0xC5, HBM-budget "
                               f"admission, continuous batching",
                   "requests": len(reqs), "served": len(done),
                   "band_lu_precision": srv.precision, "refactor_iters": srv.refactor_iters,
                   "solo_requests": srv.solo_count},
        "step_latency_ms_p50": 1e3 * float(np.percentile(lat, 50)),
        "step_latency_ms_p99": 1e3 * float(np.percentile(lat, 99)),
        "request_latency_s_p50": float(np.percentile(req_lat, 50)) if len(done) else None,
        "request_latency_s_p99": float(np.percentile(req_lat, 99)) if len(done) else None,
        "queue_wait_s_p50": float(np.percentile(queue_wait, 50)) if len(done) else None,
        "wall_s": srv.wall, "batch_rebuilds": srv.rebuilds, "rebuild_s": srv.rebuild_time,
        "max_members": int(max(srv.step_members)) if srv.step_members else 0,
        "hbm_budget_gb": srv.budget / 1e9, "peak_admitted_gb": srv.peak_bytes / 1e9,
        "failed_requests": failed,
        "plain_fp64": {"newton_per_s": plain_rate, "sample": int(len(idx)),
                       "steps_each": args.plain_steps,
                       "path": "single-context direct FP64, one request at a time",
                       "ordering": args.ordering},
        "speedup_vs_plain_fp64": value / plain_rate,
    }
    print(json.dumps(res), flush=True)


def bench_precision(args):
    """Precision form: one large patch (n = --n, default 256) whose tether moves
    across 40 % of the patch width over 500 steps (an extension: the reference
    pulls a fixed element; the tethered element steps toward decreasing
    zeta_1 from ceil(numel / 2), mafem.analysis.lateral_pull_element), pull
    speed ramped as in the plain workload.  It is advanced with the all-FP64
    direct path for --steps steps (the timed `value`) and, from the same
    initial state, with the mixed-precision solver (FP64 FGMRES
    right-preconditioned by an FP32 band LU of the current tangent, refactored
    for every Newton iteration; FP64 assembly, residual, update and epsilon
    test) for --mixed-steps steps; reports the per-step relative state error
    between the two and both throughputs.  Every step is timed (the first
    includes the sweep autotuner, well under a second at this size)."""
    import mafem
    from mafem.analysis import lateral_pull_element, move_tether
    from mafem.mesh import tether_ids
    n = args.n if args.n != 64 else 256
    dt = 1e-3
    total = 500
    out, fails = {}, {}
    for mode, nsteps in (("fp64", args.steps), ("mixed", min(args.steps, args.mixed_steps))):
        p, mesh, xms, cps = build_case(n, args.motion, sigma=0.0)
        dev = mafem.DeviceMesh(mesh, p)
        if mode == "mixed":
            dev.set_solver("krylov", precision=32, restart=30, max_iters=3000, refactor_iters=0,
                           rtol=1e-12)
        else:
            dev.set_solver("direct", precision=64, rtol=1e-13)
        dev.set_ordering(args.ordering)
        dev.set_state(xms, cps)
        el = lateral_pull_element(mesh, 0, total)
        _, dirs = tether_ids(mesh, p, el, 0.0)
        states, iters, times, stats, els = [], [], [], [], []
        for k in range(nsteps):
            new = lateral_pull_element(mesh, k, total)
            if new != el:
                _, dirs = move_tether(dev, mesh, p, el, new, pull_speed(k))
                el = new
            dev.set_prescribed([nd for (_u, nd, _v) in dirs], [mesh.dofs[u] for (u, _n, _v) in dirs],
                               [pull_speed(k)] * len(dirs))
            ts = time.perf_counter()
            dev.predict(dt)
            rc, eps = dev.time_step((k + 1) * dt, dt)
            te = time.perf_counter() - ts
            st = dev.stats()
            if rc != 0:  # the Newton iteration's own failure: reported, the run stops
                fails[mode] = {"step": k, "eps": [float(e) for e in eps]}
                print(f"bench: {mode} step {k}: Newton did not converge ({len(eps)} iterations)",
                      file=sys.stderr, flush=True)
                break
            print(f"bench: {mode} step {k} (element {el}): {te:.2f} s, {len(eps)} Newton "
                  f"iterations, factorizations {st['n_factor']}, gmres {st['gmres_iters']}, "
                  f"perturbed pivots {st['pivot_perturbations']}", file=sys.stderr, flush=True)
            states.append(dev.get_state())
            els.append(el)
            times.append(te)
            iters.append(len(eps))
            stats.append(st)
        out[mode] = (states, iters, times, stats, mesh, els)
        dev.close()
    sf, iF, tf, stf, mesh, els = out["fp64"]
    sm, im, tm, stm, _, _ = out["mixed"]
    errs = [max(np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0]),
                np.linalg.norm(a[1] - b[1]) / np.linalg.norm(b[1])) for a, b in zip(sm, sf)]
    fp64_rate = sum(iF) / sum(tf) if tf else None
    mixed_rate = sum(im) / sum(tm) if tm else None
    res = {
        "metric": METRIC,
        "value": fp64_rate,
        "unit": "Newton steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 0,
        "ms_per_step": 1e3 * sum(tf) / max(1, len(tf)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"F_PULL {args.motion} {n}x{n} patch, tether moved across 40% of "
                               f"the width over 500 steps (elements {els[0] if els else '-'}.."
                               f"{els[-1] if els else '-'} here), "
                               f"flat initial state, pull ramped to 2.0 over 200 steps, dt=1e-3; "
                               f"FP64 direct "
                               + ("band LU" if args.ordering == "band" else
                                  "nested-dissection multifrontal LU"),
                   "nmdf": mesh.nmdf, "numel": mesh.numel},
        "newton_iters_fp64": iF, "newton_iters_mixed": im,
        "per_step_rel_err_mixed_vs_fp64": errs,
        "mixed_newton_per_s": mixed_rate,
        "mixed_over_fp64": mixed_rate / fp64_rate if (mixed_rate and fp64_rate) else None,
        "newton_failures": fails,
        "step_s_fp64": tf, "step_s_mixed": tm,
        "ordering": args.ordering,
        "gmres_iters_mixed": [s["gmres_iters"] for s in stm],
        "pivot_perturbations_mixed": [s["pivot_perturbations"] for s in stm],
        "factor_us_fp64": [s["factor_us"] for s in stf],
        "factor_us_mixed": [s["factor_us"] for s in stm],
        "solve_us_fp64": [s["solve_us"] for s in stf],
        "solve_us_mixed": [s["solve_us"] for s in stm],
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--motion", default="ALEVB")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--solver", default="direct", choices=["direct", "krylov"])
    ap.add_argument("--ordering", default="nd", choices=["band", "nd"],
                    help="factorization ordering: band LU (natural node order) or nested "
                         "dissection (multifrontal)")
    ap.add_argument("--precision", type=int, default=None, choices=[64, 32])
    ap.add_argument("--workload", default="single",
                    choices=["single", "batched", "serving", "precision"])
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--req-steps", type=int, default=50)
    ap.add_argument("--solo-threads", type=int, default=0,
                    help="batched workload: run it non-lockstep on this many host threads "
                         "(direct FP64 path per patch) instead of one batch")
    ap.add_argument("--solo-min-n", type=int, default=32,
                    help="serving: requests of n x n elements or more run solo (each on its own "
                         "stream with the nested-dissection direct path; the default sends every "
                         "request of the stream, n >= 32, solo: 54 vs 29 Newton steps/s with "
                         "n >= 96 solo and the rest batched)")
    ap.add_argument("--max-solo", type=int, default=16, help="serving: concurrent solo requests")
    ap.add_argument("--mixed-steps", type=int, default=1,
                    help="precision workload: steps of the mixed-precision leg")
    ap.add_argument("--max-members", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-graphs", action="store_true",
                    help="batches: launch directly instead of replaying recorded graphs")
    ap.add_argument("--refactor-iters", type=int, default=None,
                    help="batches: refactor after a solve needing more Arnoldi steps (-1: every "
                         "Newton iteration)")
    ap.add_argument("--patches", type=int, default=512)
    ap.add_argument("--nmin", type=int, default=16)
    ap.add_argument("--nmax", type=int, default=96)
    ap.add_argument("--sample", type=int, default=8)
    ap.add_argument("--copy-probe", action="store_true",
                    help="one 256 MiB device-to-device copy and exit (calibrates HBM "
                         "counters: FETCH_SIZE / WRITE_SIZE of a copy of known size)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = 256 << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "ok": bool(dst[-1].item() == 1.0)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent patches shard across GPUs with no communication; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "serving":
        bench_serving(args)
        return
    if args.workload == "precision":
        bench_precision(args)
        return
    if args.workload == "batched":
        if args.solo_threads > 0:
            bench_batched_solo(args)
            return
        if args.precision is None:
            args.precision = 32
        bench_batched(args)
        return
    if args.precision is None:
        args.precision = 64

    import mafem
    p, mesh, xms, cps = build_case(args.n, args.motion)
    dev = mafem.DeviceMesh(mesh, p)
    dev.set_solver(args.solver, precision=args.precision, rtol=1e-13)
    dev.set_ordering(args.ordering)
    dev.set_state(xms, cps)
    pres = [(node, mesh.dofs[u]) for (u, node, _v) in mesh.inh_dir_bcs]
    nodes = [a for a, _ in pres]
    dofs = [b for _, b in pres]
    dt = 1e-3

    def step(k):
        dev.set_prescribed(nodes, dofs, [pull_speed(k)] * len(pres))
        dev.predict(dt)
        rc, eps = dev.time_step((k + 1) * dt, dt)
        if rc != 0:
            raise RuntimeError(f"Newton did not converge at step {k}: {eps}")
        return eps, dev.stats()

    for k in range(args.warmup):
        step(k)
    lat, iters = [], 0
    agg = {"assemble_us": 0.0, "area_kernel_us": 0.0, "factor_us": 0.0, "solve_us": 0.0,
           "update_us": 0.0, "n_assemble": 0, "n_band_solve": 0, "n_factor": 0,
           "gmres_iters": 0}
    step_us = []
    step_stats = []
    relres = []
    t_start = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        t0 = time.perf_counter()
        eps, st = step(k)
        lat.append(time.perf_counter() - t0)
        iters += st["newton_iters"]
        relres.append(st["last_relres"])
        for key in agg:
            agg[key] += st[key]
        if st["n_factor"] > 0 and st["step_kernel_us"] > 0:
            step_us.append(st["step_kernel_us"])
        step_stats.append(st)
    t_total = time.perf_counter() - t_start

    numel = mesh.numel
    flops_fact, nt = band_update_flops(dev)
    res = {
        "metric": METRIC,
        "value": iters / t_total,
        "unit": "Newton steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 64 else "f64 (FP32 preconditioner)",
        "data": "synthetic",
        "config": {"workload": f"F_PULL {args.motion} {args.n}x{args.n} patch, pull ramped to 2.0 "
                               f"over 200 steps, dt=1e-3, seeded z-perturbation N(0,1e-6), "
                               f"FP64 assembly; solver {args.solver}, "
                               + ("band-LU" if args.ordering == "band" else
                                  "nested-dissection multifrontal LU")
                               + f" factor FP{args.precision}",
                   "nmdf": mesh.nmdf, "numel": numel, "time_steps": args.steps,
                   "newton_iters": iters},
        "elem_assemblies_per_s": numel * agg["n_assemble"] / (agg["assemble_us"] * 1e-6),
        "area_kernel_elem_per_s": numel * agg["n_assemble"] / (agg["area_kernel_us"] * 1e-6),
        "time_steps_per_s": args.steps / t_total,
        "step_latency_ms_p50": 1e3 * float(np.percentile(lat, 50)),
        "step_latency_ms_p99": 1e3 * float(np.percentile(lat, 99)),
        "phase_us_per_newton": {k: agg[k] / max(1, iters) for k in
                                ("assemble_us", "area_kernel_us", "factor_us", "solve_us",
                                 "update_us")},
        "max_solve_relres": max(relres),
        "factorizations": agg["n_factor"], "gmres_iters": agg["gmres_iters"],
        "band_solves": agg["n_band_solve"],
    }
    # dominant kernel of the direct FP64 path: k_step_la, one launch per tile
    # step K of every chain of every factorization: the trailing update of
    # step K (tb x tb tile products A_IJ -= L_IK U_KJ on the matrix cores,
    # 2 * 64^3 FLOP each), the 64x64 diagonal LU of step K + 1 (2/3 * 64^3)
    # and its panel (2 tb triangular solves with 64 right-hand sides, 64^3
    # each); the timed launch is a middle step, where every tile is in range.
    # The two-ended factorization advances both chains (the top part downward,
    # the bottom part upward) in each launch; the separator's tb steps follow
    # as single-chain launches.
    if args.ordering == "nd":
        res["roofline"] = nd_schur_roofline(step_stats, agg)
        res["step_roofline"] = nd_roofline(step_stats, agg)
        finish_single(args, res, agg, mesh, numel, p, xms, cps, dev)
        return
    _, tb = dev_band_shape(dev)
    chains = st["factor_chains"]
    if chains == 2:
        a = (nt - tb) // 2
        launches = max(a, nt - tb - a) + tb
    else:
        launches = nt
    n_tiles = tb * tb
    flops_step = chains * (n_tiles * 2.0 * TILE ** 3 + 2.0 / 3.0 * TILE ** 3 + 2 * tb * TILE ** 3)
    s_us = float(np.mean(step_us)) if step_us else float("nan")
    achieved = flops_step / (s_us * 1e-6) / 1e12
    # HBM bytes per k_step_la launch, from two rocprofv3 --pmc passes over
    # `bench.py --steps 1 --warmup 0 --no-cpu-baseline` (FETCH_SIZE, then
    # WRITE_SIZE, KiB per dispatch averaged over its two-chain k_step_la
    # dispatches), each scaled by the same counter's reading for the 256 MiB
    # copy of `bench.py --copy-probe` (per 128 MiB copyBuffer dispatch:
    # FETCH_SIZE 65545 KiB, WRITE_SIZE 131072 KiB for 131072 KiB moved).
    traffic = None
    counters = STEP_COUNTERS_KIB["k_step_la"]
    if counters is not None:
        traffic = (counters[0] * (131072.0 / 65545.0) + counters[1] * 1.0) * 1024.0
    s_bytes = TILE * TILE * 8
    # algorithmic tiles moved per chain: the window read and written (2 tb^2),
    # step K's L and U panel tiles read (2 tb), the two next-panel tiles outside
    # the window read and written (4), the L tiles of row K+1 permuted (2 tb)
    res["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                       "traffic": traffic, "traffic_unit": "bytes per launch",
                       "algorithmic_bytes": chains * (2 * n_tiles + 4 * tb + 4) * s_bytes,
                       "kernel": ("k_step_la (one tile step of each chain of the band LU: trailing "
                                  "update on the matrix cores, the next 64x64 diagonal LU and "
                                  "the next panel in the same launch)"),
                       "flops_per_launch": flops_step, "avg_us": s_us,
                       "launches_per_factorization": launches,
                       "factorization_tflops": flops_fact / (agg["factor_us"] * 1e-6 /
                                                             max(1, agg["n_factor"])) / 1e12
                       if agg["factor_us"] > 0 else None}
    finish_single(args, res, agg, mesh, numel, p, xms, cps, dev)


def nd_schur_roofline(st_list, agg):
    """The nested-dissection factorization's matrix-core kernel, knd_schur
    (one launch per tree level: every front's boundary block minus the
    products of its fully-summed tile rows and columns, accumulated on the
    matrix cores and written once), which performs about three quarters of
    the factorization's FLOPs.  The library times every launch on the device's
    constant clock and counts its algorithmic FLOPs ((FT - fT)^2 fT tile
    products of 2*64^3 per front) and bytes; achieved = mean FLOPs per launch /
    mean launch duration over the timed steps."""
    w = [s for s in st_list if s["schur_kernel_us"] > 0]
    us = float(np.mean([s["schur_kernel_us"] for s in w])) if w else float("nan")
    fl = float(np.mean([s["schur_flops"] for s in w])) if w else float("nan")
    by = float(np.mean([s["schur_bytes"] for s in w])) if w else float("nan")
    achieved = fl / (us * 1e-6) / 1e12
    traffic = None
    counters = STEP_COUNTERS_KIB.get("knd_schur")
    if counters is not None:
        traffic = (counters[0] * (131072.0 / 65545.0) + counters[1] * 1.0) * 1024.0
    st = st_list[-1]
    per_fact_us = agg["factor_us"] / max(1, agg["n_factor"])
    return {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
            "traffic_unit": "bytes per launch", "algorithmic_bytes": by,
            "kernel": ("knd_schur (the Schur blocks of every front of a nested-dissection "
                       "level: boundary tiles minus the fully-summed rows' and columns' "
                       "products, FP64 on the matrix cores)"),
            "flops_per_launch": fl, "avg_us": us,
            "factorization_tflops": st["front_flops"] / (per_fact_us * 1e-6) / 1e12,
            "flops_per_factorization": st["front_flops"], "fronts": st["fronts"],
            "dependent_tile_steps": st["front_steps"]}


def nd_roofline(st_list, agg):
    """Dominant kernel of the nested-dissection factorization, knd_step_la (one
    launch per tile step of each tree level: for every front of the level the
    trailing update on the matrix cores, the next 64x64 diagonal LU and the
    next panel): the library times every launch on the device's constant clock
    and counts its algorithmic FLOPs (per front (FT-K-1)^2 tile products of
    2*64^3, the next diagonal LU 2/3*64^3 and 2 (FT-K-2) panel solves of
    64^3); achieved = mean FLOPs per launch / mean launch duration, averaged
    over the timed steps.  The whole factorization's rate is beside it."""
    w = [s for s in st_list if s["step_kernel_us"] > 0]
    us = float(np.mean([s["step_kernel_us"] for s in w])) if w else float("nan")
    fl = float(np.mean([s["step_flops"] for s in w])) if w else float("nan")
    by = float(np.mean([s["step_bytes"] for s in w])) if w else float("nan")
    achieved = fl / (us * 1e-6) / 1e12
    # HBM bytes per launch from FETCH_SIZE / WRITE_SIZE per dispatch (mean
    # over one time step's knd_step_la dispatches), scaled as for k_step_la
    traffic = None
    counters = STEP_COUNTERS_KIB["knd_step_la"]
    if counters is not None:
        traffic = (counters[0] * (131072.0 / 65545.0) + counters[1] * 1.0) * 1024.0
    st = st_list[-1]
    per_fact_us = agg["factor_us"] / max(1, agg["n_factor"])
    return {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
            "traffic_unit": "bytes per launch", "algorithmic_bytes": by,
            "kernel": ("knd_step_la (one tile step of every front of a nested-dissection "
                       "level: trailing update on the matrix cores, the next 64x64 diagonal "
                       "LU and the next panel in the same launch)"),
            "flops_per_launch": fl, "avg_us": us,
            "factorization_tflops": st["front_flops"] / (per_fact_us * 1e-6) / 1e12,
            "flops_per_factorization": st["front_flops"], "fronts": st["fronts"],
            "dependent_tile_steps": st["front_steps"]}


def finish_single(args, res, agg, mesh, numel, p, xms, cps, dev):
    # the area kernel (residual + all tangent columns of every area element,
    # FP64 forward mode on the vector ALUs): algorithmic FLOPs of the counted
    # forward-mode evaluation per element (F_res + F_dir per active direction)
    # over its measured time per assembly; algorithmic bytes per elem-assembly
    # (state, r_el, K_el read-modify-write, node indices)
    from mafem.opcount import assembly_flops, elem_bytes
    area_us = agg["area_kernel_us"] / max(1, agg["n_assemble"])
    a_flops = assembly_flops(args.motion, mesh.ID, mesh.IX)
    a_bytes = elem_bytes(args.motion) * numel
    a_tf = a_flops / (area_us * 1e-6) / 1e12
    a_traffic = None
    if AREA_COUNTERS_KIB.get(args.motion) is not None:
        fk, wk = AREA_COUNTERS_KIB[args.motion]
        a_traffic = (fk * (131072.0 / 65545.0) + wk * 1.0) * 1024.0
    res["area_roofline"] = {"bound": "valu", "achieved": a_tf, "peak": FP64_PEAK_TFLOPS,
                            "unit": "TFLOP/s", "frac": a_tf / FP64_PEAK_TFLOPS,
                            "flops_per_launch": a_flops, "avg_us": area_us,
                            "algorithmic_bytes": a_bytes, "traffic": a_traffic,
                            "achieved_gbs": a_bytes / (area_us * 1e-6) / 1e9,
                            "kernel": f"k_area ({args.motion} residual + tangent columns of "
                                      f"{numel} area elements, FP64 forward mode)"}
    if not args.no_cpu_baseline:
        res["cpu_baseline"], res["vs_oracle"] = cpu_baseline(p, mesh, xms, cps, args.cpu_threads,
                                                             dev)
    else:
        res["cpu_baseline"] = None
    dev.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
