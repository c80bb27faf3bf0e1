"""Pins of the assembled-system restatement (CPU only)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from earthbox_gpu import workloads
from stokes_cases import CASES, random_case


@pytest.mark.parametrize("case", CASES[:6])
def test_build_solve_matches_dense(case):
    """SuperLU (+refinement) equals a dense LU solve of the same matrix, and
    the COO has the row structure of the reference's stencils (Figure 5 of
    StokesBuildManager.jl:434-473: at most 11 entries per momentum row, 4 per
    continuity row, Kbond identity rows for ghost unknowns)."""
    pr = random_case(**case)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    N = pr.N
    A = sp.csr_matrix((Lv, (Li - 1, Lj - 1)), shape=(N, N))
    counts = np.bincount(Li - 1, minlength=N)
    assert counts[0::3].max() <= 11 and counts[1::3].max() <= 11 and counts[2::3].max() <= 4
    kb = 2 * pr.pscale / (pr.grid.xstpavg + pr.grid.ystpavg)
    nx, ny = pr.grid.xnum, pr.grid.ynum
    ghost_vx = [3 * ((nx - 2) * (ny - 1) + i) for i in range(ny - 1)]   # right column
    for r in ghost_vx:
        assert counts[r] == 1 and A[r, r] == kb
    S = oracle.direct_solve(N, Li, Lj, Lv, R)
    Sd = np.linalg.solve(A.toarray(), R)
    assert np.linalg.norm(S - Sd) <= 1e-9 * np.linalg.norm(Sd)


def test_hydrostatic_free_slip():
    """Constant density, free slip: v = 0 and p increases linearly with depth,
    dp/dy = rho*g (closed form)."""
    pr = workloads.falling_block(21, 21, eta_block=1e21, rho_block=3200.0)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    S = oracle.direct_solve(pr.N, Li, Lj, Lv, R)
    vx1, vy1, pr1 = oracle.stokes_unpack(pr, S)
    vscale = 3200.0 * 10.0 * pr.grid.xsize ** 2 / 1e21   # Stokes velocity scale rho g L^2 / eta
    assert np.abs(vx1).max() < 1e-14 * vscale and np.abs(vy1).max() < 1e-14 * vscale
    dp = np.diff(pr1, axis=0) / pr.grid.ystpavg
    assert np.allclose(dp, 3200.0 * 10.0, rtol=1e-9)
    assert pr1[0, 0] == pytest.approx(0.0, abs=1e-6)   # pressure_bc in the corner cell


def test_falling_block_symmetry():
    """Centred block in a symmetric free-slip box: vy is mirror-symmetric and
    vx antisymmetric about the vertical centre line."""
    pr = workloads.falling_block(41, 41, eta_block=1e24)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    S = oracle.direct_solve(pr.N, Li, Lj, Lv, R)
    vx1, vy1, pr1 = oracle.stokes_unpack(pr, S)
    assert vy1[1:-1, 1:-1].max() > 0  # block sinks (y down)
    s = np.abs(vy1).max()
    assert np.allclose(vy1[:, ::-1], vy1, atol=1e-9 * s, rtol=0)
    assert np.allclose(vx1[:, ::-1], -vx1, atol=1e-9 * s, rtol=0)


def test_discrete_divergence_free():
    pr = random_case(nx=17, ny=13, seed=9)
    pr.RC1[...] = 0
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    S = oracle.direct_solve(pr.N, Li, Lj, Lv, R)
    vx1, vy1, _ = oracle.stokes_unpack(pr, S)
    g = pr.grid
    div = (np.diff(vx1[1:-1, :], axis=1) / g.xstp_b[None, :] +
           np.diff(vy1[:, 1:-1], axis=0) / g.ystp_b[:, None])
    scale = np.abs(vx1).max() / g.xstpavg
    assert np.abs(div.ravel()[1:]).max() < 1e-8 * scale   # cell (1,1) holds the pressure BC


def test_mg_oracle_fixed_point_equals_direct_solve():
    """With constant viscosity, free slip and no stabilization the reference
    multigrid discretizes the same equations as the assembled system, and its
    pressure gauge (p[1,1] = prnorm) matches pressure_bc in the corner cell:
    V-cycles converge to the SuperLU solution."""
    import oracle.mg as mg
    pr = workloads.falling_block(33, 33, eta_block=1e21)
    g = pr.grid
    H = mg.Hierarchy(33, 33, g.xsize, g.ysize)
    f = H.levels[0]
    f.etas[...] = pr.etas0
    f.etan[...] = pr.etan0
    f.RX[...] = pr.RX1
    f.RY[...] = pr.RY1
    f.RC[...] = pr.RC1
    H.restrict_viscosity()
    sm = mg.smoothing_counts(len(H.levels), 2, cap=48)
    for _ in range(40):
        H.vcycle(sm)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    vx1, vy1, pr1 = oracle.stokes_unpack(pr, oracle.direct_solve(pr.N, Li, Lj, Lv, R))
    s = np.abs(vy1).max()
    assert np.abs(f.vy[1:-1, 1:-1] - vy1[1:-1, 1:-1]).max() < 1e-9 * s
    assert np.abs(f.pr - pr1).max() < 1e-9 * np.abs(pr1).max()
