"""Oracle basis machinery against the reference's own unit-test vectors.

Vectors: tests/golden/reference_unit_vectors.json (transcribed from
test/input/spline_tests.jl, gauss_pt_tests.jl, gp_basis_fn_tests.jl).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import spline as sp

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                "reference_unit_vectors.json")))
EPS = np.finfo(float).eps


def _kv(e):
    curve = sp.CLOSED if e.get("curve") == "CLOSED" else sp.CLAMPED
    if "knots" in e:
        return sp.KnotVector([float(z) for z in e["knots"]], e["poly"], curve)
    return sp.KnotVector(e["nel"], e["poly"], curve)


def test_knot_span():
    for e in G["knot_span"]:
        kv = _kv(e)
        for z, idx in e["cases"]:
            assert sp.get_knot_span_index(kv, z) == idx, (e["line"], z)
        for z in e.get("raises", []):
            with pytest.raises(AssertionError):
                sp.get_knot_span_index(kv, z)


def test_bad_knot_vectors():
    for e in G["bad_knot_vectors"]:
        with pytest.raises(AssertionError):
            _kv(e)


def test_knot_vector_equality():
    # spline_tests.jl:22-27, 59
    a = sp.KnotVector([0., 0., 0., 1., 2., 3., 4., 4., 4.], 2, sp.CLAMPED)
    assert a == sp.KnotVector([0., 0., 0., 1., 2., 3., 4., 4., 4.], 2, sp.CLAMPED)
    assert not (a == sp.KnotVector([0., 0., 0., 1.0001, 2., 3., 4., 4., 4.], 2, sp.CLAMPED))
    assert sp.KnotVector([0., 0., 0., 1 / 3, 2 / 3, 1., 1., 1.], 2, sp.CLAMPED) == \
        sp.KnotVector(3, 2, sp.CLAMPED)


def test_bspline_vals():
    for e in G["bspline_vals"]:
        got = sp.get_bspline_vals(_kv(e), e["z"])
        assert np.linalg.norm(got - np.array(e["expect"])) < e["tol_eps"] * EPS, e["line"]
    e = G["bspline_vals_below_knot"]
    z = e["z_minus_eps_of"] - EPS
    got = sp.get_bspline_vals(_kv(e), z)
    assert np.linalg.norm(got - np.array(e["expect"])) < e["tol_eps"] * EPS


def test_bspline_ders():
    for e in G["bspline_ders"]:
        d = sp.get_bspline_ders(_kv(e), e["z"], e["nd"])
        got = d if e["col"] == "all" else d[:, e["col"]]
        exp = np.array(e["expect"], dtype=float)
        assert np.linalg.norm(got - exp) < e["tol_eps"] * EPS, e["line"]
    e = G["bspline_ders_raises"]
    kv = _kv(e)
    for z, nd in e["cases"]:
        with pytest.raises(AssertionError):
            sp.get_bspline_ders(kv, z, nd)


def test_bspline_indices():
    for e in G["bspline_indices"]:
        kv = _kv(e)
        for z, ids in e["cases"]:
            assert sp.get_bspline_indices(kv, z) == ids, (e["line"], z)


def test_collocation_reproduces_polynomials():
    # spline_tests.jl:146-157: x, 3x+1 and a parabola are reproduced
    kvs = [sp.KnotVector([0., 0., 0., 1 / 3, 2 / 3, 1., 1., 1.], 2, sp.CLAMPED),
           sp.KnotVector([0., 0., 0., 0., 1 / 3, 2 / 3, 1., 1., 1., 1.], 3, sp.CLAMPED)]
    for kv in kvs:
        for f in (lambda x: x, lambda x: 3 * x + 1, lambda x: 1 - 4 * (x - 0.5) ** 2):
            cps = sp.get_1d_bspline_cps(kv, f)
            for z in np.arange(0.0, 1.0 + 1e-12, 0.1):
                z = min(z, 1.0)
                ids = [i - 1 for i in sp.get_bspline_indices(kv, z)]
                assert math.isclose(float(sp.get_bspline_vals(kv, z) @ cps[ids]), f(z),
                                    rel_tol=1e-8, abs_tol=1e-12)
    for e in G["collocation_bad"]:
        with pytest.raises(AssertionError):
            sp.get_1d_bspline_cps(_kv(e), lambda x: x)


def test_unique_1d_elements():
    for e in G["unique_1d"]:
        num, num_el, ids, lst = sp.get_unique_1d_elements(_kv(e))
        assert num == e["num"] and num_el == e["num_el"], e["line"]
        assert ids == e["ids"], e["line"]
        assert np.allclose(np.array(lst, dtype=float), np.array(e["list"], dtype=float),
                           rtol=0, atol=1e-15), e["line"]


def test_collocate_closed():
    for e in G["collocate_closed"]:
        kv = sp.KnotVector(e["nel"], e["poly"], sp.CLOSED)
        assert np.allclose(sp.collocate_z(kv), e["expect"])


def test_gauss_points():
    g = G["gauss"]
    assert [x * x for x in sp.gauss_xis(3)] == pytest.approx(g["xi3_sq"], abs=4e-16)
    assert sp.gauss_weights(3) == [5 / 9, 8 / 9, 5 / 9]
    for key, n in (("map3", 3), ("map4", 4)):
        m = g[key]
        zs, ws = sp.gauss_zeta(n, m["lo"], m["hi"])
        assert np.linalg.norm(zs - (np.array(sp.gauss_xis(n)) * m["scale"] + m["shift"])) < EPS
        assert np.linalg.norm(ws - np.array(sp.gauss_weights(n)) * m["scale"]) < EPS
    # exactness: 3-point rule integrates x^4 exactly, 4-point x^6
    assert sum(w * x ** 4 for x, w in zip(sp.gauss_xis(3), sp.gauss_weights(3))) == \
        pytest.approx(2 / 5, rel=1e-15)
    assert sum(w * x ** 6 for x, w in zip(sp.gauss_xis(4), sp.gauss_weights(4))) == \
        pytest.approx(2 / 7, rel=1e-14)


def test_tensor_product_records():
    # gp_basis_fn_tests.jl:41-61
    kv3 = sp.KnotVector([0., 0., 0., 1., 3., 3.5, 6., 6., 6.], 2, sp.CLAMPED)
    kv4 = sp.KnotVector([2., 2., 2., 4., 5., 5., 5.], 2, sp.CLAMPED)
    rng = np.random.default_rng(3)
    for z1 in np.arange(0.0, 6.0 + 1e-12, 1.5):
        for z2 in np.arange(2.0, 5.0 + 1e-12, 1.5):
            f1 = sp.gp_basis_1d(rng.random(), z1, kv3)
            f2 = sp.gp_basis_1d(rng.random(), z2, kv4)
            rec = sp.gp_basis_2d(f1, f2)
            assert rec[0] == f1[0] * f2[0]
            for i1 in range(3):
                for i2 in range(3):
                    k = i1 + 3 * i2
                    assert rec[1 + k] == f1[1][i1] * f2[1][i2]
                    assert rec[10 + k] == f1[2][i1] * f2[1][i2]
                    assert rec[19 + k] == f1[1][i1] * f2[2][i2]
                    assert rec[28 + k] == f1[3][i1] * f2[1][i2]
                    assert rec[37 + k] == f1[1][i1] * f2[3][i2]
                    assert rec[46 + k] == f1[2][i1] * f2[2][i2]


def test_line_basis_sharing():
    g = G["line_basis"]
    ks = sp.KnotVector([float(z) for z in g["short"]], 2, sp.CLAMPED)
    kl = sp.KnotVector([float(z) for z in g["long"]], 2, sp.CLAMPED)
    fs, fl = sp.LineBasis(ks, 3), sp.LineBasis(kl, 3)
    assert fs.nel == g["nel_short"] and fl.nel == g["nel_long"]
    assert max(fs.uel_ids) == g["max_uel_short"] and max(fl.uel_ids) == g["max_uel_long"]
    for i in range(5):
        for j in range(3):
            for a, b in zip(fs.ufns[i][j], fl.ufns[i][j]):
                assert np.array_equal(np.asarray(a), np.asarray(b))
    zs01, _ = sp.gauss_zeta(3, 0.0, 1.0)
    assert np.array_equal(fs.ufns[0][0][1], sp.get_bspline_ders(ks, zs01[0] + 1, 0)[:, 0])
    assert np.array_equal(fs.ufns[1][1][2], sp.get_bspline_ders(kl, zs01[1] + 2, 1)[:, 1])
    assert np.array_equal(fs.ufns[3][2][1], sp.get_bspline_ders(kl, zs01[2] + 8, 0)[:, 0])
    assert np.array_equal(fs.ufns[3][2][3], sp.get_bspline_ders(kl, zs01[2] + 8, 2)[:, 2])
    assert np.array_equal(fs.ufns[4][0][1], sp.get_bspline_ders(kl, zs01[0] + 11, 0)[:, 0])
    ka = sp.KnotVector([float(z) for z in g["asym"]], 2, sp.CLAMPED)
    fa = sp.LineBasis(ka, 3)
    assert max(fa.uel_ids) == g["max_uel_asym"]
    zs03, _ = sp.gauss_zeta(3, 0.0, 3.0)
    assert np.array_equal(fa.ufns[0][0][1], sp.get_bspline_ders(ka, zs03[0], 0)[:, 0])
    zs96, _ = sp.gauss_zeta(3, 9.0, 9.6)
    assert np.array_equal(fa.ufns[5][1][1], sp.get_bspline_ders(ka, zs96[1], 0)[:, 0])


@pytest.mark.parametrize("n", [8, 18, 19, 29, 33, 64])
def test_partition_of_unity(n):
    kv = sp.KnotVector(sp.get_fine_zs(n, 2) if n >= 18 else n, 2, sp.CLAMPED) \
        if n >= 18 else sp.KnotVector(n, 2, sp.CLAMPED)
    l = sp.LineBasis(kv, 3)
    uid, recs = sp.area_basis(l, l)
    assert np.allclose(recs[:, :, 1:10].sum(-1), 1.0, atol=1e-14)
    assert np.allclose(recs[:, :, 10:28].reshape(len(recs), 9, 2, 9).sum(-1), 0.0, atol=1e-9)
    assert np.allclose(recs[:, :, 28:55].reshape(len(recs), 9, 3, 9).sum(-1), 0.0, atol=1e-6)
    # Gauss weights sum to the parametric area of the patch
    w = np.array([recs[u - 1, :, 0].sum() for u in uid])
    assert w.sum() == pytest.approx(1.0, rel=1e-13)


def test_fine_knot_unique_counts():
    # unique 1-D elements of the centre-refined knots: 5 (n=18), 15 (19-28),
    # 23 (29-32), 25 (n >= 33)
    for n, u in ((18, 5), (19, 15), (28, 15), (29, 23), (32, 23), (33, 25), (64, 25), (256, 25)):
        kv = sp.KnotVector(sp.get_fine_zs(n, 2), 2, sp.CLAMPED)
        assert sp.get_unique_1d_elements(kv)[0] == u, n
