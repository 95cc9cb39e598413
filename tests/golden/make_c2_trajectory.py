"""Writes tests/golden/c2_alevb64_ramp.npz: the oracle's trajectory of the first
steps of the 64 x 64 ALEVB tether-pull ramp (the benchmark workload).

Initial state: the F_PULL flat patch of side 1 (src/Input.jl:77-108) with the
control points' z perturbed by N(0, 1e-6) from numpy's default_rng(0xC2);
kb = 1, zv = 1, kg = -0.5, eps_nr = 1e-10.  Step k (0-based) prescribes the
pull speed 2.0 (k + 1) / 200 on the pulled element's v_z and v^m_z
(src/input/Bc.jl:227-243, re-applied as prepare_input does, src/Input.jl:
103-108), applies the predictor update_xms! and runs time_step!
(src/Analysis.jl:60-73) at time (k + 1) dt, dt = 1e-3.

Stored: the Newton epsilon history of every step and the state after the
last step.  Run: python tests/golden/make_c2_trajectory.py [steps]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fem, mesh as om  # noqa: E402

N, SEED, SIGMA, DT = 64, 0xC2, 1e-6, 1e-3


def initial_state():
    p = om.Params(motion=om.ALEVB, num1el=N, num2el=N, length=1.0, kb=1.0, zv=1.0, kg=-0.5,
                  enr=1e-10)
    o, xms, cps = fem.build(p, pull_speed=0.0)
    rng = np.random.default_rng(SEED)
    xms[:, 2] += rng.normal(0.0, SIGMA, o.mesh.numnp)
    return o, xms, cps


def main(steps=3):
    o, xms, cps = initial_state()
    m = o.mesh
    x0_sum, c0_sum = float(xms.sum()), float(cps.sum())
    eps_all = []
    for k in range(steps):
        speed = 2.0 * (k + 1) / 200
        for (u, node, _v) in m.inh_dir_bcs:
            cps[node - 1, m.dofs[u] - 1] = speed
        fem.update_xms(o.p.motion, xms, cps, DT, m.dofs)
        t0 = time.time()
        eps = fem.time_step(o, xms, cps, (k + 1) * DT, DT)
        print(f"step {k}: {len(eps)} Newton iterations, eps {eps}, {time.time() - t0:.1f} s",
              flush=True)
        eps_all.append(eps)
    L = max(len(e) for e in eps_all)
    hist = np.zeros((steps, L))
    for k, e in enumerate(eps_all):
        hist[k, :len(e)] = e
    np.savez_compressed(os.path.join(HERE, "c2_alevb64_ramp.npz"),
                        iters=np.array([len(e) for e in eps_all]), eps=hist,
                        xms=xms, cps=cps, x0_sum=x0_sum, c0_sum=c0_sum)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 3)
