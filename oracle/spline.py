"""Oracle restatement of the reference's 1-D/2-D B-spline machinery.

Follows (reference paths):
  src/input/Spline.jl     knot vectors :20-96, fine knots :122-183, span index
                          (Piegl--Tiller A2.1) :198-250, values (A2.2) :265-296,
                          derivatives (A2.3) :318-416, indices :434-468,
                          collocation :496-626, unique elements :640-671
  src/input/GaussPoint.jl Gauss--Legendre points/weights :17-171
  src/input/GpBasisFn.jl  1-D/2-D tabulated basis functions :20-356

Index conventions are the reference's (1-based) wherever a value is returned
to a caller, so the golden vectors of the reference's own tests apply verbatim.

Pinned by: test/input/spline_tests.jl, gauss_pt_tests.jl, gp_basis_fn_tests.jl
(transcribed into tests/golden/reference_unit_vectors.json), partition of unity,
and reproduction of polynomials of degree <= 2 by collocation.
"""
import math

import numpy as np

CLAMPED = 1
CLOSED = 2

POLY = 2
GP1D = 3
NDERS = 2
NEN = (POLY + 1) ** 2
# size of one tabulated 2-D basis record: w, N[9], dN[9x2], ddN[9x3] (col-major)
REC = 1 + NEN + 2 * NEN + 3 * NEN


class KnotVector:
    """KnotVector (Spline.jl:19-31): knots, element count, degree, curve."""

    def __init__(self, arg, poly, curve):
        if isinstance(arg, (int, np.integer)):
            zs = _uniform_knots(int(arg), poly, curve)
        else:
            zs = [float(z) for z in arg]
        self.zs, self.nel, self.poly, self.curve = _check_knots(zs, poly, curve)

    def __eq__(self, other):
        return (self.zs == other.zs and self.poly == other.poly
                and self.curve == other.curve)


def _check_knots(zs, poly, curve):
    """knot_vector(ζs, poly, curve), Spline.jl:42-57."""
    if curve == CLAMPED:
        assert all(zs[i] <= zs[i + 1] for i in range(len(zs) - 1)), \
            "knots must be non-decreasing"
        assert all(z == zs[0] for z in zs[:poly + 1]), \
            "need poly+1 repeated knots at start"
        assert all(z == zs[-1] for z in zs[len(zs) - poly - 1:]), \
            "need poly+1 repeated knots at end"
    elif curve == CLOSED:
        d = np.diff(np.asarray(zs))
        ref = (zs[1] - zs[0]) * np.ones(len(zs) - 1)
        # Julia `≈` on vectors: norm(a-b) <= sqrt(eps)*max(norm(a),norm(b))
        assert np.linalg.norm(d - ref) <= math.sqrt(np.finfo(float).eps) * max(
            np.linalg.norm(d), np.linalg.norm(ref)), "closed knots must be uniform"
    else:
        raise AssertionError("curve not implemented")
    nel = len(zs) - 2 * poly - 1
    return list(zs), nel, poly, curve


def _uniform_knots(nel, poly, curve):
    """knot_vector(nel, poly, curve), Spline.jl:70-89."""
    num = nel + 2 * poly + 1
    zs = [0.0] * num
    if curve == CLAMPED:
        for idx in range(poly + 2, num - poly):          # 1-based idx
            zs[idx - 1] = (idx - poly - 1) / nel
        for idx in range(num - poly, num + 1):
            zs[idx - 1] = 1.0
    elif curve == CLOSED:
        for idx in range(1, num + 1):
            zs[idx - 1] = (idx - poly - 1) / nel
    return zs


def get_fine_zs(nel, poly):
    """get_fine_ζs (Spline.jl:122-183): knots refined towards the centre."""
    assert nel >= 18, "fine mesh requires at least 18 1-D elements"
    num = nel + 2 * poly + 1
    zs = [0.0] * (num + 1)       # 1-based storage, zs[0] unused
    nw1 = 6
    nf1 = nel - 2 * nw1 - 1
    zw1 = 1 / 3
    zf1 = 1.0 - 2 * zw1
    for idx in range(poly + 2, poly + nw1 + 2):
        dz = zw1 / nw1
        zs[idx] = (idx - poly - 1) * dz
    for idx in range(num - poly - nw1, num - poly):
        dz = zw1 / nw1
        zs[idx] = 1.0 - zw1 + (idx - num + poly + nw1) * dz
    nw2 = nf1 // 4
    nf2 = nf1 - 2 * nw2
    zw2 = 1 / 9
    zf2 = zf1 - 2 * zw2
    if 2 * nw2 <= nw1:
        dz = zf1 / (nf1 + 1)
        for idx in range(poly + nw1 + 2, num - poly - nw1):
            zs[idx] = zw1 + (idx - poly - nw1 - 1) * dz
    else:
        dz = zw2 / nw2
        for idx in range(nw1 + poly + 2, nw1 + poly + 2 + nw2):
            zs[idx] = zw1 + (idx - nw1 - poly - 1) * dz
        for idx in range(num - poly - nw1 - nw2, num - poly - nw1):
            zs[idx] = 1.0 - zw1 - zw2 + (idx - num + poly + nw1 + nw2) * dz
        dz = zf2 / (nf2 + 1)
        for idx in range(nw1 + poly + 2 + nw2, num - poly - nw1 - nw2):
            zs[idx] = zw1 + zw2 + (idx - nw1 - poly - 1 - nw2) * dz
    for idx in range(num - poly, num + 1):
        zs[idx] = 1.0
    return zs[1:]


def get_knot_span_index(kv, z):
    """Piegl--Tiller A2.1 extended to CLOSED knots (Spline.jl:198-250); 1-based."""
    zs = list(kv.zs)
    if kv.curve == CLOSED:
        for i in range(kv.poly):
            zs[i] = zs[kv.poly]
        for i in range(len(zs) - kv.poly, len(zs)):
            zs[i] = zs[len(zs) - kv.poly - 1]
    assert z >= zs[0], "ζ smaller than first active knot"
    assert z <= zs[-1], "ζ larger than last active knot"
    nk = len(zs)
    m = 1
    while zs[m - 1] == zs[0]:
        m += 1
    n = nk
    while zs[n - 1] == zs[-1]:
        n -= 1
    if z == zs[n]:          # ζs[n+1] in 1-based indexing
        return n
    low, high = m - 1, n + 1
    mid = (low + high) // 2
    while z < zs[mid - 1] or z >= zs[mid]:
        if z < zs[mid - 1]:
            high = mid
        else:
            low = mid
        mid = (low + high) // 2
    return mid


def get_bspline_vals(kv, z):
    """Piegl--Tiller A2.2 (Spline.jl:265-296)."""
    assert z >= kv.zs[0] and z <= kv.zs[-1]
    p = kv.poly
    zs = kv.zs
    left = [0.0] * (p + 1)
    right = [0.0] * (p + 1)
    s = get_knot_span_index(kv, z)
    bf = [0.0] * (p + 1)
    bf[0] = 1.0
    for j in range(1, p + 1):
        left[j] = z - zs[s - j]           # ζs[span+1-j] (1-based)
        right[j] = zs[s + j - 1] - z      # ζs[span+j]
        saved = 0.0
        for r in range(1, j + 1):
            temp = bf[r - 1] / (right[r] + left[j + 1 - r])
            bf[r - 1] = saved + right[r] * temp
            saved = left[j + 1 - r] * temp
        bf[j] = saved
    return np.array(bf)


def get_bspline_ders(kv, z, num_ders):
    """Piegl--Tiller A2.3 (Spline.jl:318-416); returns (poly+1) x (num_ders+1)."""
    assert z >= kv.zs[0] and z <= kv.zs[-1]
    p = kv.poly
    zs = kv.zs
    assert num_ders >= 0, "cannot have fewer than zero derivatives"
    assert num_ders <= p, "basis functions have only `poly` derivatives"
    left = [0.0] * (p + 2)
    right = [0.0] * (p + 2)
    s = get_knot_span_index(kv, z)
    ders = np.zeros((p + 1, num_ders + 1))
    # ndu and a use 1-based indices internally (index 0 unused)
    ndu = [[0.0] * (p + 2) for _ in range(p + 2)]
    a = [[0.0] * (p + 2) for _ in range(3)]
    ndu[1][1] = 1.0
    for j in range(1, p + 1):
        left[j + 1] = z - zs[s - j]
        right[j + 1] = zs[s + j - 1] - z
        saved = 0.0
        for r in range(1, j + 1):
            ndu[j + 1][r] = right[r + 1] + left[j + 2 - r]
            temp = ndu[r][j] / ndu[j + 1][r]
            ndu[r][j + 1] = saved + right[r + 1] * temp
            saved = left[j + 2 - r] * temp
        ndu[j + 1][j + 1] = saved
    for j in range(1, p + 2):
        ders[j - 1, 0] = ndu[j][p + 1]
    for r in range(1, p + 2):
        s1, s2 = 1, 2
        a[1][1] = 1.0
        for k in range(1, num_ders + 1):
            d = 0.0
            rk = r - k
            pk = p - k
            if r > k:
                a[s2][1] = a[s1][1] / ndu[pk + 2][rk]
                d = a[s2][1] * ndu[rk][pk + 1]
            j1 = 1 if rk >= 0 else -rk + 1
            j2 = k - 1 if r - 2 <= pk else p - r + 1
            for j in range(j1, j2 + 1):
                a[s2][j + 1] = (a[s1][j + 1] - a[s1][j]) / ndu[pk + 2][rk + j]
                d += a[s2][j + 1] * ndu[rk + j][pk + 1]
            if r - 1 <= pk:
                a[s2][k + 1] = -a[s1][k] / ndu[pk + 2][r]
                d += a[s2][k + 1] * ndu[r][pk + 1]
            ders[r - 1, k] = d
            s1, s2 = s2, s1
    r = p
    for k in range(1, num_ders + 1):
        ders[:, k] *= r
        r *= (p - k)
    return ders


def get_bspline_indices(kv, z=None, ks_id=None):
    """Global 1-based indices of the nonzero B-splines (Spline.jl:434-468)."""
    if ks_id is None:
        ks_id = get_knot_span_index(kv, z)
    p = kv.poly
    ids = [i + (ks_id - p - 1) for i in range(1, p + 2)]
    if kv.curve == CLAMPED:
        return ids
    return [(i - 1) % kv.nel + 1 for i in ids]


def collocate_z(kv):
    """collocate_ζ (Spline.jl:581-626)."""
    if kv.curve == CLOSED:
        z = np.array(kv.zs)
        return (z[kv.poly:len(z) - kv.poly - 1] + z[kv.poly + 1:len(z) - kv.poly]) / 2
    p = kv.poly
    uz = sorted(set(kv.zs))
    assert p > 1, "interpolation for poly ≥ 2 only when clamped"
    assert p < 4, "interpolation for poly ≤ 3 only when clamped"
    assert len(uz) == len(kv.zs) - 2 * p, "no repeated interior knots"
    nb = len(kv.zs) - p - 1
    zl = np.zeros(nb)
    zl[0] = uz[0]
    zl[-1] = uz[-1]
    if p == 2:
        for i in range(2, nb):
            zl[i - 1] = (uz[i - 2] + uz[i - 1]) / 2
    elif p == 3:
        zl[1] = (uz[0] + uz[1]) / 2
        zl[-2] = (uz[-2] + uz[-1]) / 2
        for i in range(3, nb - 1):
            zl[i - 1] = uz[i - 2]
    return zl


def _collocation_matrix(kv, zl):
    nb = len(zl)
    m = np.zeros((nb, nb))
    for j in range(nb):
        ids = get_bspline_indices(kv, zl[j])
        m[j, [i - 1 for i in ids]] = get_bspline_vals(kv, zl[j])
    return m


def get_1d_bspline_cps(kv, f):
    """get_1d_bspline_cps (Spline.jl:522-538): collocation solve."""
    zl = collocate_z(kv)
    m = _collocation_matrix(kv, zl)
    return np.linalg.solve(m, np.array([f(z) for z in zl]))


def get_2d_bspline_cps(kv1, kv2, f):
    """get_2d_bspline_cps (Spline.jl:540-567).

    The reference assembles the dense (n1 n2)^2 tensor-product collocation
    matrix and solves it with dense LU.  The matrix is kron(B2, B1); this
    restatement solves the same system through that factorisation,
    X = B1^-1 F B2^-T, which gives the same control points up to rounding
    and stays feasible for large meshes.  Output ordering: i1 + n1 (i2-1).
    """
    z1 = collocate_z(kv1)
    z2 = collocate_z(kv2)
    b1 = _collocation_matrix(kv1, z1)
    b2 = _collocation_matrix(kv2, z2)
    fv = np.array([[f(a, b) for b in z2] for a in z1])      # n1 x n2
    x = np.linalg.solve(b1, fv)
    x = np.linalg.solve(b2, x.T).T
    return x.reshape(-1, order="F")


def get_unique_1d_elements(kv):
    """get_unique_1d_elements (Spline.jl:640-671)."""
    p = kv.poly
    nk = len(kv.zs)
    prev = tuple([0.0] * (2 * p + 1))
    num_el = nk - 2 * p - 1
    uel_ids = [0] * num_el
    uel_num = 0
    uel_list = []
    zs = kv.zs
    for k in range(p + 1, nk - p):                  # 1-based k_idx
        el = k - p
        # ζs[k-p+1 : k+p+1] - ζs[k-p : k+p] with 1-based knot indices
        ctx = tuple(zs[i - 1] - zs[i - 2] for i in range(k - p + 1, k + p + 2))
        # Julia `≈` on scalars: |a-b| <= sqrt(eps)*max(|a|,|b|) (or a == b)
        same = all(a == b or abs(a - b) <= math.sqrt(np.finfo(float).eps) * max(abs(a), abs(b))
                   for a, b in zip(ctx, prev))
        if not same:
            uel_num += 1
            uel_list.append((zs[k - 1], zs[k]))
        uel_ids[el - 1] = uel_num
        prev = ctx
    return uel_num, num_el, uel_ids, uel_list


# ---------------------------------------------------------------- Gauss points

def gauss_xis(ngp):
    """get_ξs (GaussPoint.jl:106-133)."""
    assert ngp <= 4
    if ngp == 1:
        return [-1.0]
    if ngp == 3:
        return [-math.sqrt(3 / 5), 0.0, math.sqrt(3 / 5)]
    if ngp == 4:
        return [-math.sqrt(3 / 7 + math.sqrt(6 / 5) * 2 / 7),
                -math.sqrt(3 / 7 - math.sqrt(6 / 5) * 2 / 7),
                math.sqrt(3 / 7 - math.sqrt(6 / 5) * 2 / 7),
                math.sqrt(3 / 7 + math.sqrt(6 / 5) * 2 / 7)]
    raise AssertionError("not implemented")


def gauss_weights(ngp):
    """get_ξ_weights (GaussPoint.jl:142-171)."""
    assert ngp <= 4
    if ngp == 1:
        return [0.0]
    if ngp == 3:
        return [5 / 9, 8 / 9, 5 / 9]
    if ngp == 4:
        return [(18 - math.sqrt(30)) / 36, (18 + math.sqrt(30)) / 36,
                (18 + math.sqrt(30)) / 36, (18 - math.sqrt(30)) / 36]
    raise AssertionError("not implemented")


def gauss_zeta(ngp, zlo, zhi):
    """GaussPointsζ (GaussPoint.jl:64-77): mapped points and weights."""
    xs = np.array(gauss_xis(ngp))
    ws = np.array(gauss_weights(ngp))
    return xs * (zhi - zlo) / 2 + (zhi + zlo) / 2, ws * (zhi - zlo) / 2


# ------------------------------------------------------ tabulated basis records

def gp_basis_1d(w, z, kv):
    """GpBasisFnsζ (GpBasisFn.jl:20-55): (w, N[3], dN[3], ddN[3])."""
    d = get_bspline_ders(kv, z, NDERS)
    return (w, d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy())


def gp_basis_2d(f1, f2):
    """GpBasisFnsζα (GpBasisFn.jl:96-112) as a flat 55-double record.

    Layout (the isbits layout of the reference struct): w, N[9], dN[:,1],
    dN[:,2], ddN[:,1] (11), ddN[:,2] (22), ddN[:,3] (12); node a = i1 + 3(i2-1).
    """
    w1, n1, d1, dd1 = f1
    w2, n2, d2, dd2 = f2
    rec = np.zeros(REC)
    rec[0] = w1 * w2
    k = 0
    for i2 in range(3):
        for i1 in range(3):
            rec[1 + k] = n1[i1] * n2[i2]
            rec[10 + k] = d1[i1] * n2[i2]
            rec[19 + k] = n1[i1] * d2[i2]
            rec[28 + k] = dd1[i1] * n2[i2]
            rec[37 + k] = n1[i1] * dd2[i2]
            rec[46 + k] = d1[i1] * d2[i2]
            k += 1
    return rec


class LineBasis:
    """LineGpBasisFns (GpBasisFn.jl:143-202)."""

    def __init__(self, kv, ngp):
        num_uel, num_el, uel_ids, uel_list = get_unique_1d_elements(kv)
        self.nel = num_el
        self.uel_ids = uel_ids
        self.ufns = [[None] * ngp for _ in range(num_uel)]
        for u in range(num_uel):
            zs, ws = gauss_zeta(ngp, uel_list[u][0], uel_list[u][1])
            for g in range(ngp):
                self.ufns[u][g] = gp_basis_1d(ws[g], zs[g], kv)
        if kv.curve == CLAMPED:
            self.zmin = gp_basis_1d(1.0, kv.zs[0], kv)
            self.zmax = gp_basis_1d(1.0, kv.zs[-1], kv)
        else:
            self.zmin = gp_basis_1d(1.0, kv.zs[kv.poly], kv)
            self.zmax = gp_basis_1d(1.0, kv.zs[-kv.poly - 1], kv)


def bdry_basis(line, perp, along_z1):
    """BdryGpBasisFns (GpBasisFn.jl:258-277): (uel_ids, ufns[n_uel, ngp, 55])."""
    nu = len(line.ufns)
    ng = len(line.ufns[0])
    out = np.zeros((nu, ng, REC))
    for u in range(nu):
        for g in range(ng):
            out[u, g] = (gp_basis_2d(line.ufns[u][g], perp) if along_z1
                         else gp_basis_2d(perp, line.ufns[u][g]))
    return list(line.uel_ids), out


def area_basis(l1, l2):
    """AreaGpBasisFns (GpBasisFn.jl:325-356): (uel_ids, ufns[n_uel, 9, 55])."""
    nu1, nu2 = len(l1.ufns), len(l2.ufns)
    ng1, ng2 = len(l1.ufns[0]), len(l2.ufns[0])
    uel_ids = [0] * (l1.nel * l2.nel)
    for i1 in range(l1.nel):
        for i2 in range(l2.nel):
            uel_ids[i1 + i2 * l1.nel] = l1.uel_ids[i1] + (l2.uel_ids[i2] - 1) * nu1
    out = np.zeros((nu1 * nu2, ng1 * ng2, REC))
    for u1 in range(nu1):
        for u2 in range(nu2):
            for g1 in range(ng1):
                for g2 in range(ng2):
                    out[u1 + u2 * nu1, g1 + g2 * ng1] = gp_basis_2d(
                        l1.ufns[u1][g1], l2.ufns[u2][g2])
    return uel_ids, out
