"""Enumerations of the reference (src/input/Enums.jl:24-154, src/input/Dof.jl:29-51).

Values are the reference's own, so they cross the C ABI unchanged.
"""
import enum


class Scenario(enum.IntEnum):
    F_CAVI = 1
    F_COUE = 2
    F_POIS = 3
    F_PULL = 4
    F_BEND = 5


class Motion(enum.IntEnum):
    STATIC = 1
    EUL = 2
    LAG = 3
    ALEV = 4
    ALEVB = 5


class Boundary(enum.IntEnum):
    BOTTOM = 1
    RIGHT = 2
    TOP = 3
    LEFT = 4


class Neumann(enum.IntEnum):
    SHEAR = 1
    STRETCH = 2
    MOMENT = 3


class Unknown(enum.IntEnum):
    """Dof.Unknown (Dof.jl:29-34)."""
    vx = 1
    vy = 2
    vz = 3
    vmx = 4
    vmy = 5
    vmz = 6
    lam = 7
    pm = 8


F_CAVI, F_COUE, F_POIS, F_PULL, F_BEND = list(Scenario)
STATIC, EUL, LAG, ALEV, ALEVB = list(Motion)
BOTTOM, RIGHT, TOP, LEFT = list(Boundary)
SHEAR, STRETCH, MOMENT = list(Neumann)
