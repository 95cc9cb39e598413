"""The per-element operation count of the element assembly.

oracle/flops.py evaluates one area element's residual and tangent columns
with a counting dual scalar (GeoDynStress.jl:111-174,
FiniteElement.jl:98-126, 284-327).  Its r_el and every tangent column must
equal the complex-step oracle's (csrc/oracle.c) to rounding on a perturbed
state, so the counted evaluation is the reference's computation; the counts
must equal the constants the benchmark prices the area kernel with.
"""
import numpy as np
import pytest

from mafem.opcount import ELEM_FLOPS
from oracle import fem, flops, mesh as om

MOTIONS = [("LAG", om.LAG), ("EUL", om.EUL), ("ALEV", om.ALEV), ("ALEVB", om.ALEVB)]


def _element(motion, n=8, seed=7):
    p = om.Params(motion=motion, num1el=n, num2el=n, length=1.0, enr=1e-10)
    o, xms, cps = fem.build(p, pull_speed=0.5)
    m = o.mesh
    rng = np.random.default_rng(seed)
    xms = xms + rng.normal(0.0, 1e-2, xms.shape)
    cps = cps + rng.normal(0.0, 1e-1, cps.shape)
    el = 2 * n + 2                       # interior, away from the tether and the rings
    nodes = m.IX[:, el] - 1
    recs = np.asarray(m.area_ufns).reshape(-1, 9, 55)[m.area_uel_ids[el] - 1]
    act = [d + m.ndf * a for a in range(9) for d in range(m.ndf) if m.ID[d, nodes[a]] != 0]
    return p, o, m, xms, cps, el, nodes, recs, act


@pytest.mark.parametrize("name,motion", MOTIONS)
def test_counted_element_is_the_oracle_element(name, motion):
    p, o, m, xms, cps, el, nodes, recs, act = _element(motion)
    f_res, f_dir, r, K = flops.count(m, p, xms[nodes], cps[nodes], recs, 1e-3, act)
    r0, K0 = fem.area_elements(o, xms, cps, 1e-3, el, el + 1)
    assert np.abs(r - r0[0]).max() <= 1e-13 * np.abs(r0[0]).max()
    assert np.abs(K - K0[0]).max() <= 1e-13 * np.abs(K0[0]).max()
    assert len(act) == 9 * m.ndf
    assert len(set(f_dir)) == 1                 # the same work in every direction
    assert (f_res, f_dir[0], len(act)) == ELEM_FLOPS[name]


def test_count_sees_a_dropped_term():
    """Planting a mistake in the restatement (the bending moment term of the
    ALEVB mesh-motion row dropped) breaks the oracle comparison."""
    p, o, m, xms, cps, el, nodes, recs, act = _element(om.ALEVB)
    p_alev = om.Params(motion=om.ALEV, num1el=8, num2el=8, length=1.0, enr=1e-10)
    _, _, r, _ = flops.count(m, p_alev, xms[nodes], cps[nodes], recs, 1e-3, act[:1])
    r0, _ = fem.area_elements(o, xms, cps, 1e-3, el, el + 1)
    assert np.abs(r - r0[0]).max() > 1e-6 * np.abs(r0[0]).max()
