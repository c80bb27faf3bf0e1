"""Oracle pins for the grid residuals, Stokes norms and the time-step limit.

- Grid residuals (StokesResiduals.jl:94-137): for the direct solution of the
  assembled system they vanish (to rounding) at every node, except where the
  assembled system replaces the momentum equation (internal velocity zone
  rows, which the build places at node (i+1, j+1) of the zone's cell indices)
  and, with interface stabilization, the y rows, where the residual equals
  exactly the stabilization terms the build added (the residual formula has
  none; the terms themselves are pinned in
  test_oracle_stokes_stabilization.py).  A dropped term or a
  wrong sign would leave O(1) residuals at interior nodes.
- Norms (MathTools.jl:64-84): against numpy norms of the difference, and the
  1e38 convention for a zero reference vector.
- Time step (Update.jl:43-98): closed form for known maxima.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from stokes_cases import CASES, random_case


def _pack(pr, vx1, vy1, pr1):
    """S in the reference's unknown order (StokesBuildManager.jl:339-431)."""
    nx, ny = pr.grid.xnum, pr.grid.ynum
    S = np.empty(pr.N)
    for j in range(1, nx):
        for i in range(1, ny):
            c = (j - 1) * (ny - 1) + i
            S[3 * c - 3], S[3 * c - 2], S[3 * c - 1] = vx1[i, j], vy1[i, j], pr1[i - 1, j - 1]
    return S


def _solved(case):
    pr = random_case(**case)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    S = oracle.direct_solve(pr.N, Li, Lj, Lv, R)
    return pr, oracle.stokes_unpack(pr, S)


@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[4], CASES[5]])
def test_residuals_of_direct_solution_vanish(case):
    pr, (vx, vy, p) = _solved(case)
    rx, ry, rc = oracle.stokes_grid_residuals(pr, vx, vy, p)
    if pr.iuse_interface_stabilization:
        # YStokes.jl has no stabilization terms, so at the y rows the grid
        # residual of the stabilized solution is what the build added: the
        # stabilization coefficients applied to S minus their right-hand-side
        # part (the folded boundary values)
        Li, Lj, Lv, R = oracle.stokes_build(pr)
        Lp = oracle.stokes_build(dataclasses.replace(pr, iuse_interface_stabilization=0))
        S = _pack(pr, vx, vy, p)
        T = oracle.coo_spmv(pr.N, Li, Lj, Lv - Lp[2], S) - (R - Lp[3])
        nx, ny = pr.grid.xnum, pr.grid.ynum
        added = np.zeros_like(ry)
        for j in range(1, nx):
            for i in range(1, ny):
                added[i, j] = T[3 * ((j - 1) * (ny - 1) + i) - 2]
        assert np.abs(added).max() > 1e-3 * np.abs(ry).max()
        ry = ry - added
    assert np.abs(rx).max() <= 1e-11 * np.abs(pr.RX1).max()
    assert np.abs(ry).max() <= 1e-11 * max(np.abs(pr.RY1).max(), 1.0)
    assert np.abs(rc).max() <= 1e-9


def test_zone_residuals_sit_on_the_built_zone_rows():
    pr, (vx, vy, p) = _solved(CASES[3])
    rx, ry, rc = oracle.stokes_grid_residuals(pr, vx, vy, p)
    z = pr.bcs.bintern_zone
    big_x = {tuple(t) for t in (np.argwhere(np.abs(rx) > 1e-8 * np.abs(pr.RX1).max()) + 1).tolist()}
    big_y = {tuple(t) for t in (np.argwhere(np.abs(ry) > 1e-8 * np.abs(pr.RY1).max()) + 1).tolist()}
    assert big_x and big_x <= {(i + 1, z[0] + 1) for i in range(z[1], z[2] + 1)}
    assert big_y and big_y <= {(i + 1, z[4] + 1) for i in range(z[5], z[6] + 1)}


def test_norms_against_numpy():
    rng = np.random.default_rng(3)
    v, vo = rng.standard_normal(1000), rng.standard_normal(1000)
    d, r = oracle.get_L2_norms(v, vo)
    assert d == pytest.approx(np.linalg.norm(v - vo), rel=1e-15)
    assert r == pytest.approx(np.linalg.norm(v - vo) / np.linalg.norm(v), rel=1e-15)
    d, r = oracle.get_inf_norms(v, vo)
    assert d == np.abs(v - vo).max() and r == d / np.abs(v).max()
    assert oracle.get_L2_norms(np.zeros(5), np.ones(5))[1] == 1e38
    assert oracle.get_inf_norms(np.zeros(5), np.ones(5))[1] == 1e38


def test_timestep_closed_form():
    vx = np.array([[1e-9, -3e-9], [2e-9, 0.0]])
    vy = np.array([[0.5e-9, 4e-9]])
    dt, up = oracle.calculate_new_timestep(1e15, 0.5, 1000.0, 2000.0, vx, vy)
    assert dt == min(0.5 * 1000.0 / 3e-9, 0.5 * 2000.0 / 4e-9) and up
    dt, up = oracle.calculate_new_timestep(1.0, 0.5, 1000.0, 2000.0, vx, vy)
    assert dt == 1.0 and not up
    dt, up = oracle.calculate_new_timestep(7.0, 0.5, 1.0, 1.0, np.zeros((2, 2)), np.zeros((1, 2)))
    assert dt == 7.0 and not up
