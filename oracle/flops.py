"""Operation count of the forward-mode element assembly (test infrastructure).

A second restatement of one area element's residual and tangent columns --
GeoDynStress (src/analysis/GeoDynStress.jl:111-174) and
calc_elem_dof_residuals (src/analysis/FiniteElement.jl:284-327) -- evaluated
with a dual-number scalar that counts its real floating-point operations, the
values once (F_res) and the tangent parts per seeded direction (F_dir).  A
tangent column is the directional derivative with seed (e_col, dt e_(node,
comp)) for a motion DOF (FiniteElement.jl:110-126), which is what the device's
forward-mode element kernel evaluates per lane.

Counting: +, -, *, / and sqrt are one operation each, a sign change none; a
dual product costs one value and three tangent operations (two products and a
sum), a reciprocal one and two, a square root one and two; a scalar (basis,
parameter) times a dual costs one and one, a scalar plus a dual one and none.
Accumulations start from their first term (no additions of zero).  The
Dohrmann--Bochev matrix G^T H^-1 G depends only on the basis and is taken as
given (per unique element), as on the device.  F_elem = F_res + sum over the
element's active directions of F_dir.

Pinned by: its residual and every tangent column equal the complex-step
oracle's (csrc/oracle.c) to rounding (tests/test_oracle_flops.py).
"""
import numpy as np

from . import mesh as om

COUNT = {"v": 0, "t": 0}


class Dual:
    __slots__ = ("v", "t")

    def __init__(self, v, t=0.0):
        self.v, self.t = float(v), float(t)

    def __add__(self, o):
        if isinstance(o, Dual):
            COUNT["v"] += 1
            COUNT["t"] += 1
            return Dual(self.v + o.v, self.t + o.t)
        COUNT["v"] += 1
        return Dual(self.v + o, self.t)

    __radd__ = __add__

    def __sub__(self, o):
        if isinstance(o, Dual):
            COUNT["v"] += 1
            COUNT["t"] += 1
            return Dual(self.v - o.v, self.t - o.t)
        COUNT["v"] += 1
        return Dual(self.v - o, self.t)

    def __rsub__(self, o):
        COUNT["v"] += 1
        return Dual(o - self.v, -self.t)

    def __neg__(self):
        return Dual(-self.v, -self.t)

    def __mul__(self, o):
        if isinstance(o, Dual):
            COUNT["v"] += 1
            COUNT["t"] += 3
            return Dual(self.v * o.v, self.v * o.t + self.t * o.v)
        COUNT["v"] += 1
        COUNT["t"] += 1
        return Dual(self.v * o, self.t * o)

    __rmul__ = __mul__

    def recip(self):
        r = 1.0 / self.v
        COUNT["v"] += 1
        COUNT["t"] += 2
        return Dual(r, -self.t * r * r)

    def sqrt(self):
        s = np.sqrt(self.v)
        COUNT["v"] += 1
        COUNT["t"] += 2
        return Dual(s, self.t / (2.0 * s))


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _sum(terms):
    s = terms[0]
    for t in terms[1:]:
        s = s + t
    return s


def _acc(s, t):
    return t if s is None else s + t


def element(mesh, p, X, C, recs, dbt, col=None, dt=0.0):
    """r_el (len 9 ndf, layout dof + ndf (node - 1)) of one area element as
    duals seeded with tangent column `col` (None: no seed).  X: 9 x 3, C: 9 x
    ndf, recs: 9 Gauss-point records of 55 doubles, dbt: G^T H^-1 G (9 x 9)."""
    dofs, ndf = mesh.dofs, mesh.ndf
    mo = p.motion
    vo = [dofs.get(u, 0) for u in (om.VX, om.VY, om.VZ)]
    mord = [dofs.get(u, 0) for u in (om.VMX, om.VMY, om.VMZ)]
    lo, po = dofs.get(om.LAM, 0), dofs.get(om.PM, 0)
    mm = om.m_motion_order(mo, dofs)
    Xd = [[Dual(X[a, i]) for i in range(3)] for a in range(9)]
    Cd = [[Dual(C[a, d]) for d in range(ndf)] for a in range(9)]
    if col is not None:
        a0, d0 = col // ndf, col % ndf
        Cd[a0][d0].t = 1.0
        if d0 + 1 in mm:
            Xd[a0][mm.index(d0 + 1)].t = dt
    zero = Dual(0.0)

    def field(order, w):
        return [_sum([Cd[a][order[i] - 1] * w[a] for a in range(9)]) if order[i] else zero
                for i in range(3)]

    ale, eul = mo in (om.ALEV, om.ALEVB), mo == om.EUL
    rv = [None] * 27
    rm = [None] * 27
    rl = [None] * 9
    rp = [None] * 9
    for gp in range(9):
        rec = recs[gp]
        w, N, N1, N2 = rec[0], rec[1:10], rec[10:19], rec[19:28]
        B11, B22, B12 = rec[28:37], rec[37:46], rec[46:55]
        a1 = [_sum([Xd[a][i] * N1[a] for a in range(9)]) for i in range(3)]
        a2 = [_sum([Xd[a][i] * N2[a] for a in range(9)]) for i in range(3)]
        xx = [[_sum([Xd[a][i] * B[a] for a in range(9)]) for i in range(3)] for B in (B11, B22, B12)]
        a11, a12, a22 = _dot(a1, a1), _dot(a1, a2), _dot(a2, a2)
        det = a11 * a22 - a12 * a12
        idet = det.recip()
        g00, g01, g11 = a22 * idet, -(a12 * idet), a11 * idet
        au1 = [a1[i] * g00 + a2[i] * g01 for i in range(3)]
        au2 = [a1[i] * g01 + a2[i] * g11 for i in range(3)]
        J = det.sqrt()
        iJ = J.recip()
        n = [(a1[1] * a2[2] - a1[2] * a2[1]) * iJ, (a1[2] * a2[0] - a1[0] * a2[2]) * iJ,
             (a1[0] * a2[1] - a1[1] * a2[0]) * iJ]
        G1 = [_dot(xx[v], au1) for v in range(3)]   # Γ^1_v, v = 11, 22, 12
        G2 = [_dot(xx[v], au2) for v in range(3)]
        b11, b22, b12 = _dot(xx[0], n), _dot(xx[1], n), _dot(xx[2], n)
        t00, t01 = g00 * b11 + g01 * b12, g00 * b12 + g01 * b22
        t10, t11 = g01 * b11 + g11 * b12, g01 * b12 + g11 * b22
        bu00, bu01 = t00 * g00 + t01 * g01, t00 * g01 + t01 * g11
        bu10, bu11 = t10 * g00 + t11 * g01, t10 * g01 + t11 * g11
        H = (g00 * b11 + g01 * b12 + g01 * b12 + g11 * b22) * 0.5
        Kc = (b11 * b22 - b12 * b12) * idet
        v = field(vo, N)
        dv1 = field(vo, N1)
        dv2 = field(vo, N2)
        lam = _sum([Cd[a][lo - 1] * N[a] for a in range(9)])

        def visc(d1, d2):
            w0 = [d1[i] * g00 + d2[i] * g01 for i in range(3)]
            w1 = [d1[i] * g01 + d2[i] * g11 for i in range(3)]
            return _dot(au1, w0), _dot(au1, w1), _dot(au2, w0), _dot(au2, w1)

        p00, p01, p10, p11 = visc(dv1, dv2)
        c0 = H * H * p.kb - Kc * p.kg
        lc = c0 + lam
        h2 = H * (2.0 * p.kb)
        sg = [g00 * lc - bu00 * h2 + p00 * (2.0 * p.zv), g11 * lc - bu11 * h2 + p11 * (2.0 * p.zv),
              (g01 * lc - bu01 * h2 + (p01 + p10) * p.zv) * 2.0]
        kk = H * (p.kb + 2.0 * p.kg)
        Mv = [g00 * kk - bu00 * p.kg, g11 * kk - bu11 * p.kg,
              (g01 * kk - bu01 * p.kg) + (g01 * kk - bu10 * p.kg)]
        wJ = J * w
        tr = _dot(au1, dv1) + _dot(au2, dv2)
        ql = (J * tr - lam * (p.adb / p.zv)) * w
        if ale or eul:
            vm = field(mord, N)
        if ale:
            q00, q01, q10, q11 = visc(field(mord, N1), field(mord, N2))
            sgm = [g00 * c0 - bu00 * h2 + q00 * (2.0 * p.zv), g11 * c0 - bu11 * h2 + q11 * (2.0 * p.zv),
                   (g01 * c0 - bu01 * h2 + (q01 + q10) * p.zv) * 2.0]
            pm = _sum([Cd[a][po - 1] * N[a] for a in range(9)])
            nd = _dot(n, [vm[i] - v[i] for i in range(3)])
            qp = -(wJ * nd) - pm * (w * p.adb / p.zv)
            pmJ = pm * wJ
        if eul:
            nv = _dot(n, v)
            gm = [(vm[i] - n[i] * nv) * (p.am * 1.0) * wJ for i in range(3)]
        for a in range(9):
            s1 = sg[0] * N1[a] + sg[2] * (0.5 * N2[a])
            s2 = sg[1] * N2[a] + sg[2] * (0.5 * N1[a])
            nn = [B[a] - (G1[vv] * N1[a] + G2[vv] * N2[a])
                  for vv, B in enumerate((B11, B22, B12))]
            m = nn[0] * Mv[0] + nn[1] * Mv[1] + nn[2] * Mv[2]
            for i in range(3):
                rv[3 * a + i] = _acc(rv[3 * a + i], (a1[i] * s1 + a2[i] * s2 + n[i] * m) * wJ)
                if p.pn != 0.0:
                    rv[3 * a + i] = rv[3 * a + i] - (n[i] * N[a]) * (p.pn * 1.0) * wJ
            rl[a] = _acc(rl[a], ql * N[a])
            if ale:
                t1 = sgm[0] * N1[a] + sgm[2] * (0.5 * N2[a])
                t2 = sgm[1] * N2[a] + sgm[2] * (0.5 * N1[a])
                for i in range(3):
                    r_ = a1[i] * t1 + a2[i] * t2
                    if mo == om.ALEVB:
                        r_ = r_ + n[i] * m
                    rm[3 * a + i] = _acc(rm[3 * a + i], r_ * wJ - (pmJ * n[i]) * N[a])
                rp[a] = _acc(rp[a], qp * N[a])
            if eul:
                for i in range(3):
                    rm[3 * a + i] = _acc(rm[3 * a + i], gm[i] * N[a])
    azv = p.adb / p.zv
    for a in range(9):
        rl[a] = rl[a] + _sum([Cd[b][lo - 1] * dbt[a, b] for b in range(9)]) * azv
        if ale:
            rp[a] = rp[a] + _sum([Cd[b][po - 1] * dbt[a, b] for b in range(9)]) * azv
    r = [zero] * (9 * ndf)
    for a in range(9):
        for i in range(3):
            if vo[i]:
                r[vo[i] - 1 + ndf * a] = rv[3 * a + i]
            if mord[i]:
                r[mord[i] - 1 + ndf * a] = rm[3 * a + i]
        r[lo - 1 + ndf * a] = rl[a]
        if po:
            r[po - 1 + ndf * a] = rp[a]
    return r


def db_matrix(recs):
    """G^T H^-1 G of the Dohrmann--Bochev projection (FiniteElement.jl:315-323)
    with N_DB = (xi_1, xi_2, 1) at the raw Gauss abscissae."""
    from .spline import gauss_xis
    xi = gauss_xis(3)
    G = np.zeros((3, 9))
    Hm = np.zeros((3, 3))
    for gp in range(9):
        ndb = np.array([xi[gp % 3], xi[gp // 3], 1.0])
        G += np.outer(ndb, recs[gp][1:10]) * recs[gp][0]
        Hm += np.outer(ndb, ndb) * recs[gp][0]
    return G.T @ np.linalg.solve(Hm, G)


def count(mesh, p, X, C, recs, dt, active_cols):
    """(F_res, F_dir per active column, r_el, K_el columns) of one element."""
    dbt = db_matrix(recs)
    COUNT["v"] = COUNT["t"] = 0
    r = element(mesh, p, X, C, recs, dbt)
    f_res = COUNT["v"]
    f_dir = []
    K = np.zeros((9 * mesh.ndf, 9 * mesh.ndf))
    for col in active_cols:
        COUNT["v"] = COUNT["t"] = 0
        rc = element(mesh, p, X, C, recs, dbt, col=col, dt=dt)
        f_dir.append(COUNT["t"])
        K[:, col] = [x.t for x in rc]
    return f_res, f_dir, np.array([x.v for x in r]), K
