"""Host-only check of the launch-size guard that batched segment maps use.

A dispatch counts its work-items in 32 bits; a batch whose segment needs more
workgroups than fit (e.g. more than 2^24 workgroups of 256) must be refused
with an error instead of skipping the tail of the work.  The guard is a plain
C++ header, compiled here with the host compiler.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "membranealefem.jl_amd", "csrc")

PROG = r"""
#include <cstdio>
#include "launch_limits.h"
using mafem::launch_fits;
int main() {
  int bad = 0;
  auto check = [&](bool got, bool want, const char *what) {
    if (got != want) { std::printf("FAIL %s\n", what); ++bad; }
  };
  check(launch_fits(0, 256), true, "empty launch");
  check(launch_fits((1ull << 24) - 1, 256), true, "2^24 - 1 blocks of 256");
  check(launch_fits(1ull << 24, 256), false, "2^24 blocks of 256 (2^32 work-items)");
  check(launch_fits(8000000, 256), true, "512-patch bench BSR segment");
  check(launch_fits(17000000, 256), false, "~1100-patch BSR segment");
  check(launch_fits(1ull << 22, 1024), false, "2^22 sweep workgroups of 1024");
  check(launch_fits((1ull << 22) - 1, 1024), true, "2^22 - 1 sweep workgroups");
  check(launch_fits((1ull << 26) - 1, 64), true, "2^26 - 1 blocks of 64");
  check(launch_fits(1ull << 26, 64), false, "2^26 blocks of 64");
  check(launch_fits(1, 0), false, "zero-thread workgroup");
  std::printf("%s\n", bad ? "BAD" : "OK");
  return bad;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_launch_fits_boundaries(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(PROG)
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + CSRC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout
