"""Closed-form pin of the y-Stokes interface-stabilization terms (CPU only).

With interface stabilization the build adds, at each internal vy row
(discretize_y_stokes_equation/internal/InternalCells.jl:104-168 and the
stencil-node files it calls),
  -drho/dy * g * dt                  to the central vy coefficient, and
  -drho/dx * g * dt / 4              to each of the four vx coefficients
                                     around the vy node,
with drho/dy, drho/dx the averages of the two one-sided difference quotients
of rho1_vy.  For a density linear in x and y those quotients are exact, so
(L_stab - L_plain) S at the vy row of cell (i, j) must equal
  -g dt (c vy(i+1, j+1) + b/4 (vx(i+1, j) + vx(i+1, j+1) + vx(i+2, j) + vx(i+2, j+1)))
for rho = a + b x + c y, and every other row must be unchanged.  A dropped
term, a wrong sign or a wrong vx corner fails this.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from earthbox_gpu.grid import graded_grid, uniform_grid
from stokes_cases import random_case


def _pack(g, vx1, vy1, pr1):
    """Solution vector in the reference's unknown order: per cell
    (j-1)(ny-1)+i the triple [vx(i+1,j+1), vy(i+1,j+1), p(i,j)]
    (StokesBuildManager.jl:339-431)."""
    nx, ny = g.xnum, g.ynum
    S = np.empty(3 * (nx - 1) * (ny - 1))
    for j in range(1, nx):
        for i in range(1, ny):
            c = (j - 1) * (ny - 1) + i
            S[3 * c - 3] = vx1[i, j]
            S[3 * c - 2] = vy1[i, j]
            S[3 * c - 1] = pr1[i - 1, j - 1]
    return S


@pytest.mark.parametrize("graded", [False, True])
def test_stabilization_terms_closed_form(graded):
    nx, ny = 17, 13
    base = random_case(nx=nx, ny=ny, seed=11, stab=1)
    if graded:
        g = graded_grid(nx, ny, base.grid.xsize, base.grid.ysize, seed=5)
        base = dataclasses.replace(base, grid=g)
    g = base.grid
    a, b, c = 3000.0, 2.5e-3, -4.0e-3            # kg/m^3, kg/m^4
    X, Y = np.meshgrid(g.gridx_vy, g.gridy_b)    # vy nodes
    rho = np.asfortranarray(a + b * X + c * Y)
    stab = dataclasses.replace(base, rho1_vy=rho, iuse_interface_stabilization=1)
    plain = dataclasses.replace(base, rho1_vy=rho, iuse_interface_stabilization=0)
    rng = np.random.default_rng(12)
    vx1 = rng.uniform(-1e-9, 1e-9, (ny + 1, nx))
    vy1 = rng.uniform(-1e-9, 1e-9, (ny, nx + 1))
    pr1 = rng.uniform(-1e7, 1e7, (ny - 1, nx - 1))
    S = _pack(g, vx1, vy1, pr1)
    Ls = oracle.stokes_build(stab)
    Lp = oracle.stokes_build(plain)
    N = S.size
    # the stabilization changes coefficient values only, not the triplet
    # order; applying the coefficient difference avoids cancelling two
    # products dominated by the pressure terms
    assert np.array_equal(Ls[0], Lp[0]) and np.array_equal(Ls[1], Lp[1])
    diff = oracle.coo_spmv(N, Ls[0], Ls[1], Ls[2] - Lp[2], S)
    gdt = base.gravity_y * base.timestep
    interior = np.zeros(N, bool)
    expected = np.zeros(N)
    for j in range(2, nx - 2):
        for i in range(2, ny - 2):
            cell = (j - 1) * (ny - 1) + i
            r = 3 * cell - 2                    # the vy row (0-based)
            vxs = vx1[i, j - 1] + vx1[i, j] + vx1[i + 1, j - 1] + vx1[i + 1, j]
            expected[r] = -gdt * (c * vy1[i, j] + b / 4.0 * vxs)
            interior[r] = True
    scale = np.abs(expected[interior]).max()
    assert scale > 0
    np.testing.assert_allclose(diff[interior], expected[interior], rtol=1e-9, atol=1e-10 * scale)
    # vx rows and continuity rows carry no stabilization term
    rows = np.arange(N)
    for off in (0, 2):
        assert np.all(diff[rows % 3 == off] == 0.0)
    # the right-hand side is unchanged at interior rows
    assert np.array_equal(Ls[3][interior], Lp[3][interior])


def test_stabilization_off_without_timestep():
    """dt = 0 makes the stabilization terms vanish exactly."""
    base = random_case(nx=11, ny=9, seed=13, stab=1)
    z = dataclasses.replace(base, timestep=0.0)
    p0 = dataclasses.replace(base, timestep=0.0, iuse_interface_stabilization=0)
    Lz, Lp = oracle.stokes_build(z), oracle.stokes_build(p0)
    for u, v in zip(Lz, Lp):
        assert np.array_equal(u, v)
