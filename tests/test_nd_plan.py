"""Host logic of the nested-dissection front plan (csrc/nd_plan.h).

tests/nd_check.cpp builds the plan and runs a dense CPU emulation of the
multifrontal factorization and solve the device performs over it (same front
layout, entry ownership, child gather order and tile-local pivoting) on a
random matrix with the tangent's node-block pattern, prescribed identity rows
and a zero diagonal entry per node; its solution must match a dense
partial-pivoting LU solve.  The plan's structural invariants (every coupled
entry in exactly one front, children at lower levels, separators with two
children) are checked there too.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "membranealefem.jl_amd", "csrc")


@pytest.fixture(scope="module")
def nd_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("nd") / "nd_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "nd_check.cpp"),
                    "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("n1,n2,ndf,leaf,seed", [
    (10, 10, 8, 4, 1),   # 8x8 elements, ALEVB-sized node blocks
    (12, 9, 4, 4, 2),    # rectangular, LAG
    (9, 14, 7, 4, 3),    # EUL
    (10, 10, 8, 6, 4),   # larger leaves
    (5, 5, 8, 4, 5),     # one separator level
    (3, 4, 8, 4, 6),     # a single leaf front
    (18, 18, 4, 4, 7),   # five levels
])
def test_multifrontal_emulation_solves(nd_check, n1, n2, ndf, leaf, seed):
    out = subprocess.run([nd_check, str(n1), str(n2), str(ndf), str(leaf), str(seed)],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    rel_nd, rel_dense, diff, fronts, levels, steps = out.stdout.split()
    assert float(rel_nd) <= 1e-10
    assert float(diff) <= 1e-9
    assert int(fronts) >= 1 and int(levels) >= 1 and int(steps) >= 1


def test_plan_chain_lengths(nd_check):
    """The dependent tile steps of the 64x64 ALEVB plan (66 x 66 nodes, 8
    unknowns): the level maxima of the fully-summed tile rows, 2, 1, 2, 2, 4,
    4, 8, 8, 17 from the leaves to the root."""
    src = r'''
#include <cstdio>
#include "nd_plan.h"
int main() {
  auto P = mafem::nd::build_plan(66, 66, 8);
  for (auto &l : P.by_level) printf("%d ", P.fr[l.front()].fT);
  printf("| %lld %zu\n", (long long)P.critical_steps(), P.fr.size());
}
'''
    d = os.path.dirname(nd_check)
    with open(os.path.join(d, "steps.cpp"), "w") as f:
        f.write(src)
    exe = os.path.join(d, "steps")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(d, "steps.cpp"), "-o",
                    exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    levels, rest = out.split("|")
    assert [int(x) for x in levels.split()] == [2, 1, 2, 2, 4, 4, 8, 8, 17]
    steps, fronts = rest.split()
    assert int(steps) == 48 and int(fronts) == 511
