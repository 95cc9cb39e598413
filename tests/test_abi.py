"""The C ABI: every symbol include/mafem.h declares is exported by the built
library and bound with a signature; the ctypes struct layouts agree with the
C compiler's (sizeof/offsetof of a test program built from the header); the
create path fails cleanly on invalid input.  No GPU compute here."""
import ctypes
import os
import re
import subprocess

import pytest

from mafem import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "mafem.h")


def declared():
    text = open(HDR).read()
    return sorted(set(re.findall(r"\b(mafem_[a-z_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared()
    for n in ("mafem_create", "mafem_destroy", "mafem_assemble", "mafem_time_step",
              "mafem_solve", "mafem_predict", "mafem_set_state", "mafem_get_state"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in declared():
        assert hasattr(lib, n), n
    assert set(declared()) == set(_lib.SIGNATURES)


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "lay.c"
    fields = {"mafem_mesh_desc": _lib.MeshDesc, "mafem_params": _lib.ParamsDesc,
              "mafem_stats": _lib.Stats, "mafem_solver_opts": _lib.SolverOpts,
              "mafem_batch_opts": _lib.BatchOpts, "mafem_batch_stats": _lib.BatchStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mafem.h"', "int main(void){"]
    for cname, cls in fields.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.split() for l in out if l)
    for cname, cls in fields.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


def test_create_rejects_null_arguments():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.mafem_create(None, None, None) == _lib.EINVAL
    rc = lib.mafem_create(ctypes.byref(h), None, None)
    assert rc < 0 and not h.value
    assert isinstance(lib.mafem_last_error(None), bytes)
    # calls on a null context fail without touching a device
    assert lib.mafem_predict(None, 1.0) == _lib.EINVAL
    assert lib.mafem_time_step(None, 0.0, 1.0, 14, None, None) == _lib.EINVAL


def test_batch_create_rejects_bad_arguments():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.mafem_batch_create(None, 1, None, None) == _lib.EINVAL
    assert lib.mafem_batch_create(ctypes.byref(h), 0, None, None) == _lib.EINVAL
    assert not h.value
    assert b"member" in lib.mafem_batch_last_error(None)
    assert lib.mafem_batch_time_step(None, None, None, None, 0, 14, None, None, None) == _lib.EINVAL
    st = _lib.BatchStats()
    assert lib.mafem_batch_get_stats(None, ctypes.byref(st)) == _lib.EINVAL
