"""ctypes binding of libmafem (the C ABI declared in include/mafem.h).

The library is loaded from the package directory; there is no fallback: a
missing or unloadable library raises at first use.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libmafem.so")

OK, NOT_CONVERGED, EINVAL, EHIP, ENOMEM, EUNSUPPORTED = 0, 1, -1, -2, -3, -4

i64p = ctypes.POINTER(ctypes.c_int64)
i32p = ctypes.POINTER(ctypes.c_int32)
f64p = ctypes.POINTER(ctypes.c_double)
u8p = ctypes.POINTER(ctypes.c_uint8)


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("num1el", ctypes.c_int64), ("num2el", ctypes.c_int64), ("numel", ctypes.c_int64),
        ("num1np", ctypes.c_int64), ("num2np", ctypes.c_int64), ("numnp", ctypes.c_int64),
        ("ndf", ctypes.c_int64), ("nmdf", ctypes.c_int64),
        ("IX", i64p), ("ID", i64p),
        ("v_order", ctypes.c_int32 * 3), ("m_order", ctypes.c_int32 * 3),
        ("lam_order", ctypes.c_int32), ("p_order", ctypes.c_int32),
        ("area_n_uel", ctypes.c_int64), ("area_uel_ids", i64p), ("area_ufns", f64p),
        ("bdry_n_el", ctypes.c_int64 * 4), ("bdry_elems", i64p * 4),
        ("bdry_n_uel", ctypes.c_int64 * 4), ("bdry_uel_ids", i64p * 4),
        ("bdry_ufns", f64p * 4),
        ("n_neu", ctypes.c_int64), ("neu_bdry", i32p), ("neu_type", i32p), ("neu_val", f64p),
        ("crnr_ufns", f64p),
    ]


class ParamsDesc(ctypes.Structure):
    _fields_ = [
        ("motion", ctypes.c_int32), ("scenario", ctypes.c_int32),
        ("kb", ctypes.c_double), ("kg", ctypes.c_double), ("zv", ctypes.c_double),
        ("pn", ctypes.c_double), ("adb", ctypes.c_double), ("am", ctypes.c_double),
        ("ek", ctypes.c_double), ("enr", ctypes.c_double), ("bend_tm", ctypes.c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("assemble_us", ctypes.c_double), ("area_kernel_us", ctypes.c_double),
        ("factor_us", ctypes.c_double), ("solve_us", ctypes.c_double),
        ("update_us", ctypes.c_double), ("total_us", ctypes.c_double),
        ("newton_iters", ctypes.c_int32), ("refine_sweeps", ctypes.c_int32),
        ("n_assemble", ctypes.c_int32), ("n_band_solve", ctypes.c_int32),
        ("n_factor", ctypes.c_int32), ("gmres_iters", ctypes.c_int32),
        ("step_kernel_us", ctypes.c_double),
        ("last_relres", ctypes.c_double),
        ("sweep_groups", ctypes.c_int32), ("pivot_perturbations", ctypes.c_int32),
        ("factor_chains", ctypes.c_int32), ("separator_rows", ctypes.c_int32),
        ("fronts", ctypes.c_int32), ("front_steps", ctypes.c_int32),
        ("front_flops", ctypes.c_double), ("step_flops", ctypes.c_double),
        ("step_bytes", ctypes.c_double), ("schur_kernel_us", ctypes.c_double),
        ("schur_flops", ctypes.c_double), ("schur_bytes", ctypes.c_double),
    ]


SOLVER_DIRECT, SOLVER_KRYLOV = 0, 1
ORDER_BAND, ORDER_ND = 0, 1


class SolverOpts(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32), ("factor_precision", ctypes.c_int32),
        ("restart", ctypes.c_int32), ("max_iters", ctypes.c_int32),
        ("refactor_iters", ctypes.c_int32), ("refine_sweeps", ctypes.c_int32),
        ("rtol", ctypes.c_double),
    ]


class BatchOpts(ctypes.Structure):
    _fields_ = [
        ("factor_precision", ctypes.c_int32), ("restart", ctypes.c_int32),
        ("max_cycles", ctypes.c_int32), ("refactor_iters", ctypes.c_int32),
        ("use_graphs", ctypes.c_int32), ("escalate_iters", ctypes.c_int32),
        ("rtol", ctypes.c_double),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("total_us", ctypes.c_double), ("assemble_us", ctypes.c_double),
        ("factor_us", ctypes.c_double), ("solve_us", ctypes.c_double),
        ("update_us", ctypes.c_double),
        ("newton_iters", ctypes.c_int32), ("factor_passes", ctypes.c_int32),
        ("gmres_cycles", ctypes.c_int32), ("hard_failures", ctypes.c_int32),
        ("patch_iters", ctypes.c_int64), ("patch_factors", ctypes.c_int64),
        ("gmres_iters", ctypes.c_int64), ("elements", ctypes.c_int64),
        ("fp64_patches", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


# every symbol include/mafem.h declares, with its ctypes signature
SIGNATURES = {
    "mafem_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(MeshDesc),
                                    ctypes.POINTER(ParamsDesc)]),
    "mafem_destroy": (None, [ctypes.c_void_p]),
    "mafem_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "mafem_set_state": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p]),
    "mafem_get_state": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p]),
    "mafem_set_prescribed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, i64p, i32p, f64p]),
    "mafem_predict": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "mafem_set_dof_ids": (ctypes.c_int, [ctypes.c_void_p, i64p, ctypes.c_int64]),
    "mafem_nnz": (ctypes.c_int, [ctypes.c_void_p, i64p]),
    "mafem_assemble": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, f64p,
                                      i64p, i64p, f64p]),
    "mafem_solve": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, f64p]),
    "mafem_time_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int32, i32p, f64p]),
    "mafem_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    "mafem_set_solver": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolverOpts)]),
    "mafem_set_ordering": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "mafem_get_solver": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolverOpts)]),
    "mafem_device_memory": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t),
                                           ctypes.POINTER(ctypes.c_size_t)]),
    "mafem_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_void_p),
                                          ctypes.POINTER(BatchOpts)]),
    "mafem_batch_destroy": (None, [ctypes.c_void_p]),
    "mafem_batch_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "mafem_batch_time_step": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p, u8p, ctypes.c_int32,
                                             ctypes.c_int32, i32p, f64p, i32p]),
    "mafem_batch_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BatchStats)]),
    "mafem_pull_force": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p]),
    "mafem_pull_force_noalloc": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p]),
    "mafem_generate_output": (ctypes.c_int, [ctypes.c_void_p, f64p, f64p]),
    "mafem_batch_pull_force": (ctypes.c_int, [ctypes.c_void_p, f64p]),
}

_lib = None


def load():
    """Load libmafem.so and bind every entry point (raises if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmafem.so not built ({LIB_PATH}); run __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def ptr(a, kind):
    return a.ctypes.data_as(kind)


class LibError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmafem error {code}: {msg}")
        self.code = code
