"""Product-side mesh tables (the host mirror of the reference's input layer)
against the oracle's independent restatement: integer tables bit-exact, basis
tables to rounding, initial state to rounding."""
import numpy as np
import pytest

import mafem
from mafem import mesh as mm
from oracle import mesh as om

CASES = [(mafem.LAG, 8), (mafem.EUL, 11), (mafem.ALEV, 12), (mafem.ALEVB, 16),
         (mafem.ALEVB, 20), (mafem.LAG, 19), (mafem.ALEVB, 33), (mafem.ALEVB, 64)]


@pytest.mark.parametrize("motion,n", CASES)
def test_mesh_tables_match_oracle(motion, n):
    p = mafem.Params(motion=motion, num1el=n, num2el=n, length=1.0, output=False)
    m = mafem.make_mesh(p, pull_speed=0.7)
    o = om.make_mesh(om.Params(motion=int(motion), num1el=n, num2el=n, length=1.0),
                     pull_speed=0.7)
    assert (m.IX == o.IX).all()
    assert (m.ID == o.ID).all() and m.nmdf == o.nmdf
    assert (m.LM == o.LM).all()
    assert [tuple(t) for t in m.ID_inv] == o.ID_inv
    assert list(m.area_uel_ids) == list(o.area_uel_ids)
    scale = np.abs(o.area_ufns).max(axis=(0, 1))
    assert np.all(np.abs(m.area_ufns.transpose(1, 0, 2) - o.area_ufns) <= 1e-14 * scale + 1e-300)
    for b in (mafem.BOTTOM, mafem.RIGHT, mafem.TOP, mafem.LEFT):
        uids, recs = o.bdry_fns[int(b)]
        assert list(m.bdry_uel_ids[b]) == list(uids)
        assert list(m.bdry_elems[b]) == list(o.bdry_elems[int(b)])
        sc = np.abs(recs).max(axis=(0, 1))
        assert np.all(np.abs(m.bdry_ufns[b].transpose(1, 0, 2) - recs) <= 1e-14 * sc + 1e-300)
    assert [(int(u), n_, v) for u, n_, v in m.inh_dir_bcs] == o.inh_dir_bcs
    assert [(int(b), int(t), v) for b, t, v in m.inh_neu_bcs] == o.inh_neu_bcs
    xm, cm = mafem.prepare_input(p, m, pull_speed=0.7)
    xo, co = om.prepare_input(om.Params(motion=int(motion), num1el=n, num2el=n, length=1.0), o,
                              pull_speed=0.7)
    assert np.abs(xm - xo).max() < 1e-14
    assert np.array_equal(cm, co)


@pytest.mark.parametrize("n", [18, 19, 28, 29, 32, 33, 64, 256])
def test_fine_knots_match_oracle(n):
    from oracle import spline as sp
    assert mm.centre_refined_knots(n) == sp.get_fine_zs(n, 2)


def test_scenario_bcs_other_scenarios():
    for sc, mo, kw in ((mafem.F_CAVI, mafem.STATIC, {}), (mafem.F_COUE, mafem.STATIC, {}),
                       (mafem.F_POIS, mafem.STATIC, {}),
                       (mafem.F_BEND, mafem.LAG, {"bend_mf": 0.5})):
        p = mafem.Params(motion=mo, scenario=sc, num1el=5, num2el=4, length=1.0, output=False)
        m = mafem.make_mesh(p, **kw)
        o = om.make_mesh(om.Params(motion=int(mo), scenario=int(sc), num1el=5, num2el=4,
                                   length=1.0), **kw)
        assert (m.ID == o.ID).all(), sc
        assert [(int(b), int(t), v) for b, t, v in m.inh_neu_bcs] == o.inh_neu_bcs


def test_pull_element_position():
    # ceil(numel/2): right-edge element for even n, centre for odd n
    assert mafem.get_pull_el_id(64) == 32
    assert mafem.get_pull_el_id(8 * 8) == 32 and (32 - 1) % 8 == 7     # el1 = n
    assert mafem.get_pull_el_id(19 * 19) == 181 and (181 - 1) % 19 == 9
