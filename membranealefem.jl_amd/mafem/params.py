"""`Params` of the reference (src/input/Params.jl:36-55) and its checks (:70-125)."""
from dataclasses import dataclass, field

from .enums import ALEVB, F_BEND, F_CAVI, F_COUE, F_POIS, F_PULL, Motion, Scenario, STATIC

POLY = 2
GP1D = 3
NDERS = 2
NEN = (POLY + 1) ** 2


@dataclass
class Params:
    """Simulation parameters; field names follow the reference (ζv -> zv,
    αdb -> adb, αm -> am, εk -> ek, εnr -> enr)."""
    motion: Motion = ALEVB
    scenario: Scenario = F_PULL
    num1el: int = 17
    num2el: int = 17
    output: bool = True
    length: float = 64.0
    kb: float = 1.0
    kg: float = -0.5
    zv: float = 1.0
    pn: float = 0.0
    poly: int = POLY
    gp1d: int = GP1D
    nders: int = NDERS
    nen: int = NEN
    adb: float = field(default=None)
    am: float = 1.0
    ek: float = 1.0e-15
    enr: float = 1.0e-12

    def __post_init__(self):
        self.motion = Motion(self.motion)
        self.scenario = Scenario(self.scenario)
        if self.adb is None:
            self.adb = self.length ** 2       # αdb = length^2 (Params.jl:51)


def check_params(p, **args):
    """check_params (Params.jl:70-125): raises AssertionError like the reference."""
    for k in ("dts", "t0", "t0_id"):
        assert k in args, f"\nneed '{k}'"
    if p.output:
        assert "out_file" in args, "\nneed output file 'out_file'"
        assert "out_path" in args, "\nneed output path 'out_path'"
    assert p.nen == (p.poly + 1) ** 2, "\nincorrect number of element nodes"
    assert p.nen == NEN and p.poly == POLY and p.nders == NDERS
    if p.scenario == F_BEND:
        assert "bend_tm" in args, "\nneed final ramp-up time 'bend_tm'"
        assert "bend_mf" in args, "\nneed final applied bending moment 'bend_mf'"
        assert args["bend_mf"] == p.kb / 2 / p.length, "\nα = 1/2 is the half-angle"
        assert p.motion != STATIC
        assert p.motion != ALEVB
        assert p.pn == 0.0
    elif p.scenario == F_PULL:
        assert "pull_speed" in args, "\nneed pull speed 'pull_speed'"
    elif p.scenario in (F_CAVI, F_COUE, F_POIS):
        assert p.motion == STATIC
