"""Oracle pull force (PullForce.jl:8-80), pinned three ways, and the oracle's
generate_output (Output.jl:30-118), pinned by exact sampling of the linear
and constant fields of the flat state.

* Closed form on the flat F_PULL state: with x = L (ζ - 1/2), uniform tension
  λ = kb/4 and velocities normal to the plane, the velocity residual of a node
  is λ L ∫ ∇_ζ N over its support (the viscous and bending stresses vanish),
  so the force on the pulled nodes is λ L ∮ (Σ N) ν.  For an interior pulled
  element (odd n) it is zero; for the right-edge element of even n
  (ceil(numel/2) is then the last element of row n/2) only the edge ζ1 = 1
  contributes, where the last ζ1 basis function is 1:
  f = (λ L Σ_j ∫ N_j(ζ2) dζ2, 0, 0), with ∫ N_j = (t_{j+3} - t_j) / 3 over the
  pulled element's three ζ2 basis functions j.
* The 25 adjacent elements cover the supports of the pulled nodes: the force
  equals the sum of every element's velocity residual rows at those nodes.
* The reference's documented tether example (docs/src/index.md:142-176: LAG,
  19 x 19, length 16, pull speed 0.5, 32 steps of 0.5) reproduces the force
  curve of its figure (tests/golden/docs_pull_force.json, digitized from
  docs/src/assets/example-pull.png; one pixel is ~0.028).
"""
import json
import os

import numpy as np
import pytest

from oracle import fem, mesh as om

HERE = os.path.dirname(os.path.abspath(__file__))


def _build(motion, n, length=1.0, kb=1.0, speed=0.0):
    p = om.Params(motion=motion, scenario=om.F_PULL, num1el=n, num2el=n, length=length, kb=kb,
                  output=False)
    o, xms, cps = fem.build(p, pull_speed=speed)
    return p, o, xms, cps


@pytest.mark.parametrize("motion", [om.LAG, om.ALEVB])
@pytest.mark.parametrize("n,length,kb,speed", [(8, 1.0, 1.0, 0.0), (8, 16.0, 2.0, 0.5),
                                               (20, 1.0, 1.0, 0.5), (20, 4.0, 0.3, 0.0),
                                               (9, 1.0, 1.0, 0.5), (17, 2.0, 1.5, 0.0)])
def test_flat_state_closed_form(motion, n, length, kb, speed):
    p, o, xms, cps = _build(motion, n, length, kb, speed)
    m = o.mesh
    ids, maps = fem.get_adj_maps(n, m.numel, m.IX)
    f = fem.calc_pull_force(o, xms, cps, ids, maps)
    e = om.get_pull_el_id(m.numel)
    el1, el2 = (e - 1) % n + 1, (e - 1) // n + 1
    if n % 2:
        assert (el1, el2) == ((n + 1) // 2, (n + 1) // 2)
        expect = np.zeros(3)
    else:
        assert (el1, el2) == (n, n // 2)
        t = m.kv2.zs  # clamped knots, 0-based list
        integral = sum((t[j + 2] - t[j - 1]) / 3.0 for j in (el2, el2 + 1, el2 + 2))
        expect = np.array([kb / 4.0 * length * integral, 0.0, 0.0])
    assert np.allclose(f, expect, rtol=1e-13, atol=1e-13 * kb * length), (f, expect)
    if n % 2 == 0:
        assert f[0] > 0.0


@pytest.mark.parametrize("motion,n", [(om.LAG, 8), (om.ALEVB, 8), (om.EUL, 20), (om.ALEVB, 9)])
def test_adjacent_elements_cover_the_pulled_nodes(motion, n):
    p, o, xms, cps = _build(motion, n, speed=0.3)
    m = o.mesh
    rng = np.random.default_rng(17 + n)
    xms = xms + 1e-2 * rng.standard_normal(xms.shape)
    cps = cps + 1e-1 * rng.standard_normal(cps.shape)
    ids, maps = fem.get_adj_maps(n, m.numel, m.IX)
    f = fem.calc_pull_force(o, xms, cps, ids, maps)
    pulled = [int(v) for v in m.IX[:, om.get_pull_el_id(m.numel) - 1]]
    brute = np.zeros(3)
    for el in range(m.numel):
        nodes = [int(v) for v in m.IX[:, el]]
        if not set(nodes) & set(pulled):
            continue
        rv = fem.elem_rv(o, el, xms, cps)
        for b, node in enumerate(nodes):
            if node in pulled:
                brute += rv[3 * b: 3 * b + 3]
    assert np.abs(brute).max() > 1e-3
    assert np.allclose(f, brute, rtol=1e-12, atol=1e-14 * np.abs(brute).max())


def test_adjacency_wraps_rows_like_the_reference():
    """For even n the pulled element is the last of its row: ids pull + 1 and
    pull + 2 are the first elements of the next row and share no node."""
    p, o, xms, cps = _build(om.LAG, 8)
    m = o.mesh
    ids, maps = fem.get_adj_maps(8, m.numel, m.IX)
    pull = om.get_pull_el_id(m.numel)
    assert len(ids) == 25 and ids[12] == pull == 32
    assert ids[:5] == [14, 15, 16, 17, 18] and ids[-5:] == [46, 47, 48, 49, 50]
    for k, e in enumerate(ids):
        shared = len(maps[k][0]) // 3
        assert len(maps[k][0]) == len(maps[k][1])
        col = (e - 1) % 8 + 1
        if col <= 2:   # wrapped into the next row, left edge
            assert shared == 0
        else:
            assert shared > 0
    assert len(maps[12][0]) == 27 and maps[12][0] == maps[12][1]


def test_documented_tether_force_curve():
    fig = json.load(open(os.path.join(HERE, "golden", "docs_pull_force.json")))
    ex = fig["example"]
    p = om.Params(motion=om.LAG, scenario=om.F_PULL, num1el=ex["num1el"], num2el=ex["num2el"],
                  length=ex["length"], output=False)
    o, xms, cps = fem.build(p, pull_speed=ex["pull_speed"])
    m = o.mesh
    ids, maps = fem.get_adj_maps(ex["num1el"], m.numel, m.IX)
    dt = ex["dt"]
    for k, (zf, ff) in enumerate(fig["zp_fz"]):
        fem.update_xms(p.motion, xms, cps, dt, m.dofs)
        fem.time_step(o, xms, cps, dt * (k + 1), dt)
        zp = fem.pull_position(m, xms)[2]
        fz = fem.calc_pull_force(o, xms, cps, ids, maps)[2]
        assert abs(zp - zf) < 1e-9
        assert abs(fz - ff) < 0.03, (k, zp, fz, ff)


def _samples(kv):
    """ζ samples of generate_output along one direction: 0, the 3 Gauss
    abscissae of every element, 1."""
    t = kv.zs
    xi = (-np.sqrt(0.6), 0.0, np.sqrt(0.6))
    out = [0.0]
    for e in range(kv.nel):
        lo, hi = t[e + 2], t[e + 3]
        out += [0.5 * (lo + hi) + 0.5 * (hi - lo) * x for x in xi]
    return np.array(out + [1.0])


@pytest.mark.parametrize("motion,n1,n2,length", [(om.LAG, 4, 3, 1.0), (om.ALEVB, 20, 19, 2.0),
                                                 (om.EUL, 1, 2, 1.0)])
def test_generate_output_samples_linear_fields_exactly(motion, n1, n2, length):
    """generate_output (Output.jl:30-118) on the flat state: B-splines
    reproduce x = L (ζ - 1/2) and the constant λ, so every sample equals the
    field at its ζ (Gauss points of each element, edges, corners)."""
    p = om.Params(motion=motion, scenario=om.F_PULL, num1el=n1, num2el=n2, length=length,
                  kb=0.8, output=False)
    o, xms, cps = fem.build(p, pull_speed=0.0)
    m = o.mesh
    xout, uout = fem.generate_output(m, xms, cps)
    z1, z2 = _samples(m.kv1), _samples(m.kv2)
    assert xout.shape == (3 * n1 + 2, 3 * n2 + 2, 3)
    assert uout.shape == (3 * n1 + 2, 3 * n2 + 2, m.ndf)
    assert np.allclose(xout[:, :, 0], length * (z1[:, None] - 0.5), rtol=0, atol=1e-14 * length)
    assert np.allclose(xout[:, :, 1], length * (z2[None, :] - 0.5), rtol=0, atol=1e-14 * length)
    assert np.all(xout[:, :, 2] == 0.0)
    assert np.allclose(uout[:, :, m.dofs[om.LAM] - 1], 0.2, rtol=0, atol=1e-15)
