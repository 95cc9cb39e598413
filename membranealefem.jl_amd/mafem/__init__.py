"""MI355X (gfx950) implementation of MembraneAleFem.jl's per-time-step hot
path: element residual/tangent assembly and the Newton linear solve, behind
the C ABI of include/mafem.h (libmafem.so), with a host mirror of the
reference's entry points (time_step!, calc_r_K, update_xms!, run_analysis!).
"""
from .enums import (ALEV, ALEVB, BOTTOM, EUL, F_BEND, F_CAVI, F_COUE, F_POIS, F_PULL, LAG, LEFT,
                    MOMENT, RIGHT, SHEAR, STATIC, STRETCH, TOP, Boundary, Motion, Neumann,
                    Scenario, Unknown)
from .params import Params, check_params
from .mesh import Mesh, make_mesh, prepare_input, get_pull_el_id
from .analysis import (Batch, DeviceMesh, calc_pull_force, calc_r_K, generate_output,
                       get_adj_maps, restart, run_analysis, solve, time_step, update_xms)

__all__ = [
    "Params", "check_params", "Mesh", "make_mesh", "prepare_input", "get_pull_el_id",
    "Batch", "DeviceMesh", "calc_r_K", "run_analysis", "time_step", "update_xms",
    "get_adj_maps", "calc_pull_force", "generate_output", "solve", "restart",
    "Scenario", "Motion", "Boundary", "Neumann", "Unknown",
    "F_CAVI", "F_COUE", "F_POIS", "F_PULL", "F_BEND",
    "STATIC", "EUL", "LAG", "ALEV", "ALEVB",
    "BOTTOM", "RIGHT", "TOP", "LEFT", "SHEAR", "STRETCH", "MOMENT",
]
