"""Algorithmic operation and byte counts of one area-element assembly.

One elem-assembly is r_el plus the full K_el of one area element
(src/analysis/FiniteElement.jl:98-126 with GeoDynStress.jl:111-174).  Its
algorithmic FLOPs are those of the forward-mode evaluation:

    F_elem = F_res + n_dir * F_dir,

F_res the real floating-point operations of the residual's values, F_dir the
tangent operations of one seeded direction (a tangent column), n_dir the
element's active directions (9 ndf on an interior element).  Counted with a
dual scalar that tallies +, -, *, /, sqrt (an FMA is two), on an interior
element of the 8 x 8 F_PULL patch; the count has no data-dependent branch
(for pn = 0) and is the same for every direction.

Algorithmic bytes: the element state 9 (3 + ndf) doubles read, r_el 9 ndf
doubles and K_el (9 ndf)^2 doubles written once, and the 9 node indices read
twice (4 bytes each); basis tables are cache-resident and not counted.
"""

# motion -> (F_res, F_dir, n_dir of an interior element)
ELEM_FLOPS = {
    "LAG": (9720, 14220, 36),
    "EUL": (10791, 15453, 63),
    "ALEV": (14580, 21654, 72),
    "ALEVB": (15066, 22626, 72),
}

NDF = {"LAG": 4, "EUL": 7, "ALEV": 8, "ALEVB": 8}


def elem_flops(motion):
    """F_elem of one interior element, in FLOP."""
    f_res, f_dir, n_dir = ELEM_FLOPS[motion]
    return f_res + n_dir * f_dir


def assembly_flops(motion, ID, IX):
    """Algorithmic FLOPs of one assembly of every area element: F_res per
    element plus F_dir per active direction (ID != 0) of its 9 nodes, as the
    reference evaluates them (columns of prescribed DOFs are skipped).
    ID: ndf x numnp (0 = prescribed), IX: 9 x numel (1-based node ids)."""
    import numpy as np
    f_res, f_dir, _ = ELEM_FLOPS[motion]
    ID = np.asarray(ID)
    IX = np.asarray(IX)
    n_dir = (ID[:, IX - 1] != 0).sum()
    return f_res * IX.shape[1] + f_dir * int(n_dir)


def elem_bytes(motion):
    """Algorithmic HBM bytes of one elem-assembly."""
    ndf = NDF[motion]
    ne = 9 * ndf
    return 9 * (3 + ndf) * 8 + ne * 8 + ne * ne * 8 + 2 * 9 * 4
