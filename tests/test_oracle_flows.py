"""Flat 2-D scenarios (STATIC motion; unknowns vx, vy, λ; Bc.jl:58-157,
286-316), pinned by their analytic steady flows.

* F_COUE: v_x = 0 on the bottom edge, 3 on the top, traction 4 ν on the left
  and right: plane Couette flow, v_x = 3 (y / L + 1/2), v_y = 0, λ = 4.
* F_POIS: walls at top and bottom, traction 4 ν on the left and 8 ν on the
  right: λ = 6 + 4 x / L and the parabolic profile
  v_x = (4 / L) / (2 ζv) (y + L/2)(L/2 - y) of the membrane Stokes equations.
* F_CAVI: lid-driven cavity: no closed form, but the flow is mirror-symmetric
  about x = 0 (v_x even, v_y odd).
Both profiles are quadratic at most, so the quadratic B-spline solution is
exact: every generate_output sample matches to rounding.
"""
import numpy as np
import pytest

from oracle import fem, mesh as om


def solve_flow(scen, L, zv, n1, n2):
    p = om.Params(motion=om.STATIC, scenario=scen, num1el=n1, num2el=n2, length=L, zv=zv,
                  output=False)
    o, xms, cps = fem.build(p)
    its = [len(fem.time_step(o, xms, cps, 0.1 * (k + 1), 0.1)) for k in range(2)]
    xout, uout = fem.generate_output(o.mesh, xms, cps)
    return its, xout, uout


@pytest.mark.parametrize("L,zv,n1,n2", [(1.0, 1.0, 6, 6), (2.0, 0.5, 5, 7)])
def test_couette(L, zv, n1, n2):
    its, xout, uout = solve_flow(om.F_COUE, L, zv, n1, n2)
    y = xout[..., 1]
    assert its[1] == 1
    assert np.abs(uout[..., 0] - 3 * (y / L + 0.5)).max() < 1e-12
    assert np.abs(uout[..., 1]).max() < 1e-12
    assert np.abs(uout[..., 2] - 4.0).max() < 1e-12


@pytest.mark.parametrize("L,zv,n1,n2", [(1.0, 1.0, 6, 6), (2.0, 0.5, 7, 5)])
def test_poiseuille(L, zv, n1, n2):
    its, xout, uout = solve_flow(om.F_POIS, L, zv, n1, n2)
    x, y = xout[..., 0], xout[..., 1]
    g = 4.0 / L
    assert np.abs(uout[..., 0] - g / (2 * zv) * (y + L / 2) * (L / 2 - y)).max() < 1e-12
    assert np.abs(uout[..., 1]).max() < 1e-12
    assert np.abs(uout[..., 2] - (6.0 + g * x)).max() < 1e-12


def test_cavity_symmetry():
    its, xout, uout = solve_flow(om.F_CAVI, 1.0, 1.0, 6, 6)
    vx, vy = uout[..., 0], uout[..., 1]
    assert np.abs(vx).max() == pytest.approx(1.0)
    assert np.abs(vx - vx[::-1, :]).max() < 1e-12
    assert np.abs(vy + vy[::-1, :]).max() < 1e-12
    assert np.abs(vy).max() > 1e-2
