"""Synthetic tether-pull patches for the batched and serving workloads.

Every patch is the reference's F_PULL scenario (src/input/Bc.jl:188-269,
src/input/Input.jl:77-108) on a flat square of side 1 with its own size,
formulation and parameters; the pull is the prescribed velocity of the pulled
element's control points, ramped linearly in time to the drawn magnitude.
Mesh tables depend only on (motion, n), so they are built once per pair and
shared; each patch's copy carries its own edge traction kb / 4.
"""
import copy
from dataclasses import dataclass

import numpy as np

from .analysis import DeviceMesh
from .enums import ALEVB, EUL, LAG, Unknown
from .mesh import make_mesh, prepare_input
from .params import Params


@dataclass
class PatchSpec:
    motion: int
    n: int
    kb: float
    zv: float
    speed: float      # final pull speed of the ramp
    seed: int
    ramp_steps: int = 200

    def speed_at(self, k):
        """Prescribed pull speed of time step k (0-based)."""
        return self.speed * min(k + 1, self.ramp_steps) / self.ramp_steps


def batched_specs(count=512, seed=0xC4, sizes=(16, 96), motions=(LAG, ALEVB)):
    """Patches with n ~ U{sizes}, kb and zv log-uniform in [0.1, 10], tether
    magnitude ~ U[0.2, 3], motion uniform over `motions`; one sub-seed each."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.integers(sizes[0], sizes[1] + 1))
        kb = float(10.0 ** rng.uniform(-1.0, 1.0))
        zv = float(10.0 ** rng.uniform(-1.0, 1.0))
        speed = float(rng.uniform(0.2, 3.0))
        motion = motions[int(rng.integers(0, len(motions)))]
        out.append(PatchSpec(motion, n, kb, zv, speed, int(rng.integers(0, 2 ** 31))))
    return out


def serving_specs(count, seed=0xC5, sizes=(32, 48, 64, 96, 128), motions=(LAG, EUL, ALEVB)):
    """Request parameters of the serving stream: n from `sizes`, formulation
    uniform over `motions`, parameters drawn as in the batched workload."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(sizes[int(rng.integers(0, len(sizes)))])
        kb = float(10.0 ** rng.uniform(-1.0, 1.0))
        zv = float(10.0 ** rng.uniform(-1.0, 1.0))
        speed = float(rng.uniform(0.2, 3.0))
        motion = motions[int(rng.integers(0, len(motions)))]
        out.append(PatchSpec(motion, n, kb, zv, speed, int(rng.integers(0, 2 ** 31))))
    return out


_MESHES = {}


def mesh_for(motion, n, kb=1.0):
    """The F_PULL mesh of (motion, n) for bending modulus kb.  The mesh tables
    depend only on (motion, n) and are built once per pair; the Neumann edge
    traction is the patch's own kb / 4 (src/input/Bc.jl:259-264), as
    generate_mesh would give it for the patch's Params."""
    key = (int(motion), int(n))
    if key not in _MESHES:
        p = Params(motion=motion, num1el=n, num2el=n, length=1.0, enr=1e-10, output=False)
        _MESHES[key] = make_mesh(p, pull_speed=0.0)
    m = copy.copy(_MESHES[key])
    m.inh_neu_bcs = [(b, t, kb / 4) for (b, t, _v) in m.inh_neu_bcs]
    m._ctx = None
    return m


class Patch:
    """One patch: its parameters, host mesh and device context."""

    def __init__(self, spec: PatchSpec, sigma=0.0):
        self.spec = spec
        self.p = Params(motion=spec.motion, num1el=spec.n, num2el=spec.n, length=1.0,
                        kb=spec.kb, zv=spec.zv, kg=-0.5, enr=1e-10, output=False)
        self.mesh = mesh_for(spec.motion, spec.n, spec.kb)
        xms, cps = prepare_input(self.p, self.mesh, pull_speed=0.0)
        if sigma > 0.0:  # optional seeded out-of-plane perturbation
            rng = np.random.default_rng(spec.seed)
            xms[:, 2] += rng.normal(0.0, sigma, self.mesh.numnp)
        self.dev = DeviceMesh(self.mesh, self.p)
        self.dev.set_state(xms, cps)
        m = self.mesh
        pres = [(node, m.dofs[u]) for (u, node, _v) in m.inh_dir_bcs]
        if int(spec.motion) == int(EUL):  # EUL also prescribes v^m_z at the pulled nodes
            pres += [(node, m.dofs[Unknown.vmz]) for (u, node, _v) in m.inh_dir_bcs
                     if u == Unknown.vz]
        self._nodes = [n for n, _ in pres]
        self._dofs = [d for _, d in pres]

    def set_pull(self, speed):
        self.dev.set_prescribed(self._nodes, self._dofs, [float(speed)] * len(self._nodes))

    def close(self):
        self.dev.close()


def device_bytes(spec: PatchSpec, precision=32, restart=20, ordering="band"):
    """Device memory of one patch context plus its batch storage (bytes).
    ordering "nd": its factorization held as nested-dissection fronts instead
    of a band, bounded by five times the FP64 band (fronts of 32x32 to 128x128
    patches measure 4.9 to 2.7 times it)."""
    ndf = {int(LAG): 4, int(EUL): 7}.get(int(spec.motion), 8)
    n = spec.n
    numel = n * n
    n1 = n + 2
    numnp = n1 * n1
    nx = numnp * ndf
    nt = -(-nx // 64)
    tb = min(-(-((2 * n1 + 3) * ndf - 1) // 64), max(1, nt - 1))
    ne = 9 * ndf
    es = 4 if precision == 32 else 8
    rows = nt + tb if nt >= 8 * tb else nt  # the two-ended layout keeps tb more tile rows
    band = rows * (2 * tb + 1) * 64 * 64 * es + 2 * rows * 64 * 64 * es
    if ordering == "nd":
        band = 5 * (rows * (2 * tb + 1) * 64 * 64 * 8 + 2 * rows * 64 * 64 * 8)
    nnzb = numnp * 25
    kel = (numel + 4 * n) * ne * ne * 8
    bsr = nnzb * ndf * ndf * 8
    gather = 81 * (numel + 4 * n) * 8 * 2
    vecs = (7 + restart + 1) * nt * 64 * 8
    return band + kel + bsr + gather + vecs + numel * 9 * 56 * 8
