"""Host side of the output files and the pull-force maps (no GPU).

* Float64 text as Julia prints it: known answers of Julia's `print` for
  doubles (plain notation for -4 < pt <= 6, else "e" notation; shortest
  round-trip digits), and exact read-back of every written double.
* writedlm / readdlm of state matrices, the times.txt and f-pull.txt lines.
* The host mirror of get_adj_maps (PullForce.jl:25-53) agrees with the
  oracle's restatement for even, odd and centre-refined meshes.
"""
import numpy as np
import pytest

import mafem
from mafem import output as out
from oracle import mesh as om
from oracle import fem

KNOWN = [
    (1.0, "1.0"), (0.5, "0.5"), (-2.5, "-2.5"), (0.001, "0.001"), (0.0001, "0.0001"),
    (1e-5, "1.0e-5"), (0.00012, "0.00012"), (1.2e-5, "1.2e-5"), (100000.0, "100000.0"),
    (123456.7, "123456.7"), (1e6, "1.0e6"), (1234567.0, "1.234567e6"),
    (0.1 + 0.2, "0.30000000000000004"), (-0.0, "-0.0"), (0.0, "0.0"),
    (2.0 ** 53, "9.007199254740992e15"), (1e100, "1.0e100"), (5e-324, "5.0e-324"),
    (1.7976931348623157e308, "1.7976931348623157e308"), (3e-12, "3.0e-12"),
    (float("inf"), "Inf"), (float("-inf"), "-Inf"),
]


@pytest.mark.parametrize("x,text", KNOWN)
def test_julia_float_known_answers(x, text):
    assert out.julia_float(x) == text


def test_julia_float_reads_back_exactly():
    rng = np.random.default_rng(0x0F)
    vals = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-12, 12, 2000),
                           rng.uniform(-1, 1, 500)])
    for v in vals:
        assert float(out.julia_float(v)) == v


def test_writedlm_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((37, 8)) * 1e-3
    a[0, 0] = 0.0
    a[1, 1] = 1e-7
    path = tmp_path / "t3-cps.txt"
    out.writedlm(path, a)
    lines = path.read_text().splitlines()
    assert len(lines) == 37 and all(len(line.split("\t")) == 8 for line in lines)
    assert lines[0].split("\t")[0] == "0.0"
    assert np.array_equal(out.readdlm(path), a)


def test_times_and_pull_lines():
    assert out.times_text([0.5, 1.0], 2) == "4 times:\n3\t0.5\n4\t1.0\n"
    assert out.PULL_HEADER == "time\txp\typ\tzp\tfx\tfy\tfz\n"
    line = out.pull_line(0.5, [0.0, -1e-17, 0.25], [1e-16, 0.0, 0.3198])
    assert line == "0.5\t0.0\t-1.0e-17\t0.25\t1.0e-16\t0.0\t0.3198\n"
    assert out.iteration_lines([0.1, 2e-13]) == \
        "--> iteration 1: ε = 0.1\n--> iteration 2: ε = 2.0e-13\n"


@pytest.mark.parametrize("n", [8, 9, 19, 20])
def test_host_adjacency_matches_oracle(n):
    po = om.Params(motion=om.LAG, scenario=om.F_PULL, num1el=n, num2el=n, output=False)
    mo = om.make_mesh(po, pull_speed=0.0)
    ids_o, maps_o = fem.get_adj_maps(n, mo.numel, mo.IX)
    p = mafem.Params(motion=mafem.LAG, scenario=mafem.F_PULL, num1el=n, num2el=n, output=False)
    m = mafem.make_mesh(p, pull_speed=0.0)
    assert np.array_equal(np.asarray(m.IX), np.asarray(mo.IX))
    ids, maps = mafem.get_adj_maps(m.num1el, m.numel, m.IX)
    assert ids == ids_o
    assert [tuple(x) for x in maps] == [tuple(x) for x in maps_o]


def test_host_adjacency_rejects_tiny_meshes():
    p = mafem.Params(motion=mafem.LAG, scenario=mafem.F_PULL, num1el=2, num2el=2, output=False)
    m = mafem.make_mesh(p, pull_speed=0.0)
    with pytest.raises(IndexError):
        mafem.get_adj_maps(m.num1el, m.numel, m.IX)


@pytest.mark.parametrize("n1,n2", [(4, 3), (20, 19)])
def test_host_corner_records_match_oracle(n1, n2):
    po = om.Params(motion=om.ALEVB, scenario=om.F_PULL, num1el=n1, num2el=n2, output=False)
    mo = om.make_mesh(po, pull_speed=0.0)
    p = mafem.Params(motion=mafem.ALEVB, scenario=mafem.F_PULL, num1el=n1, num2el=n2,
                     output=False)
    m = mafem.make_mesh(p, pull_speed=0.0)
    assert m.crnr_ufns.shape == (4, 55)
    for k in range(4):   # N exactly; derivatives to rounding (second derivatives ~1e3)
        assert np.array_equal(m.crnr_ufns[k][:10], mo.crnr_fns[k][:10])
        assert np.allclose(m.crnr_ufns[k], mo.crnr_fns[k], rtol=1e-14, atol=0)
    # corner basis of clamped knots: 1 at the corner node, 0 elsewhere
    local = [0, 2, 6, 8]
    for k in range(4):
        assert m.crnr_ufns[k][1 + local[k]] == 1.0 and np.abs(m.crnr_ufns[k][1:10]).sum() == 1.0
