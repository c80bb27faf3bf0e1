"""Pins of the oracle's viscoelastic terms and right-hand sides (K7,
StokesViscoelasticTerms.jl:26-108, StokesContinuityRhs.jl:30-190) by closed
forms on a graded grid: the Maxwell series viscosity, fields whose
difference quotients are known exactly, and the zero entries the reference
loops skip."""
import numpy as np

import oracle
from ve_cases import ve_fields


def _run(g, f, dt=3.15e11, gx=0.0, gy=9.81, stab=0):
    return oracle.stokes_viscoelastic_rhs(g, f, dt, gx, gy, stab)


def test_maxwell_series_viscosity_and_stress_memory():
    g, f = ve_fields(13, 9, 1)
    dt = 3.15e11
    o = _run(g, f, dt)
    for eta, mu, s1, e0, s0 in ((f["etas1"], f["mus1"], f["sxy1"], o["etas0"], o["sxy0"]),
                                (f["etan1"], f["mun1"], f["sxx1"], o["etan0"], o["sxx0"])):
        # viscous and elastic elements in series: 1/eta0 = 1/eta + 1/(dt mu)
        np.testing.assert_allclose(1.0 / e0, 1.0 / eta + 1.0 / (dt * mu), rtol=1e-12)
        # the old stress kept in proportion to the elastic share: s0 = s1 eta0 / (dt mu)
        np.testing.assert_allclose(s0, s1 * e0 / (dt * mu), rtol=1e-12, atol=1e-300)


def test_linear_fields_give_exact_difference_quotients():
    """sxx1 = beta(x) * y on pressure nodes and sxy1 = gamma(y) * x on basic
    nodes with uniform viscosity and modulus (constant factor f): the y- and
    x-quotients of the stress terms are known per node, so a shifted row or
    column index, a dropped term or a wrong sign breaks the equality."""
    nx, ny = 15, 11
    g, f = ve_fields(nx, ny, 2, dilatation=False)
    rng = np.random.default_rng(7)
    eta, mu, dt = 1e21, 3e10, 3.15e11
    fac = eta / (eta + dt * mu)
    for k in ("etas1", "etan1"):
        f[k] = np.full_like(f[k], eta)
    for k in ("mus1", "mun1"):
        f[k] = np.full_like(f[k], mu)
    beta = rng.uniform(-50, 50, nx - 1)   # per pressure column
    gamma = rng.uniform(-50, 50, ny)      # per basic row
    xp, yp = g.gridx_pr, g.gridy_pr
    xb, yb = g.gridx_b, g.gridy_b
    f["sxx1"] = np.asfortranarray(np.outer(yp, np.ones(nx - 1)) * beta[None, :])
    f["sxy1"] = np.asfortranarray(gamma[:, None] * np.outer(np.ones(ny), xb))
    f["rho1"] = np.zeros((ny, nx), order="F")
    o = _run(g, f, dt, gx=0.0, gy=0.0)
    RY, RX = o["RY1"], o["RX1"]
    for i in range(2, ny):            # vy nodes i = 2..ynum-1, j = 2..xnum (1-based)
        for j in range(2, nx + 1):
            # d(sxx0)/dy at pressure column j-1, minus d(sxy0)/dx at basic row i
            exp = fac * beta[j - 2] - fac * gamma[i - 1]
            assert abs(RY[i - 1, j - 1] - exp) <= 1e-9 * (abs(fac * beta[j - 2]) + abs(fac * gamma[i - 1])), (i, j)
    for i in range(2, ny + 1):        # vx nodes i = 2..ynum, j = 2..xnum-1
        for j in range(2, nx):
            dsxx_dx = fac * yp[i - 2] * (beta[j - 1] - beta[j - 2]) / (xp[j - 1] - xp[j - 2])
            dsxy_dy = fac * xb[j - 1] * (gamma[i - 1] - gamma[i - 2]) / (yb[i - 1] - yb[i - 2])
            exp = -dsxx_dx - dsxy_dy
            assert abs(RX[i - 1, j - 1] - exp) <= 1e-9 * (abs(dsxx_dx) + abs(dsxy_dy)), (i, j)


def test_gravity_terms_and_skipped_entries():
    """No old stress: the momentum right-hand sides are minus gravity times the
    density at the velocity node (the mean of the two basic nodes beside it,
    exact for density linear along that direction, or rho1_vy with interface
    stabilization); entries outside the loops stay zero."""
    nx, ny = 12, 9
    g, f = ve_fields(nx, ny, 3, dilatation=False)
    f["sxy1"][:] = 0.0
    f["sxx1"][:] = 0.0
    rho = 3000.0 + 0.002 * g.gridx_b[None, :] + 0.001 * g.gridy_b[:, None]
    f["rho1"] = np.asfortranarray(rho)
    gx, gy = 2.5, 9.81
    o = _run(g, f, gx=gx, gy=gy)
    RX, RY, RC = o["RX1"], o["RY1"], o["RC1"]
    for i in range(2, ny):
        for j in range(2, nx + 1):
            r = 3000.0 + 0.002 * g.gridx_vy[j - 1] + 0.001 * g.gridy_b[i - 1]
            assert abs(RY[i - 1, j - 1] + gy * r) <= 1e-12 * gy * r
    for i in range(2, ny + 1):
        for j in range(2, nx):
            r = 3000.0 + 0.002 * g.gridx_b[j - 1] + 0.001 * g.gridy_vx[i - 1]
            assert abs(RX[i - 1, j - 1] + gx * r) <= 1e-12 * gx * r
    assert not RY[0, :].any() and not RY[ny - 1, :].any() and not RY[:, 0].any()
    assert not RX[0, :].any() and not RX[ny, :].any() and not RX[:, 0].any() and not RX[:, nx - 1].any()
    assert not RC.any()
    o2 = _run(g, f, gx=gx, gy=gy, stab=1)
    inner = (slice(1, ny - 1), slice(1, nx))
    np.testing.assert_array_equal(o2["RY1"][inner], -gy * f["rho1_vy"][inner])


def test_plastic_volumetric_term():
    """RC1 = 2 sin(dilatation) eii: 30 degrees gives eii, 90 degrees 2 eii."""
    g, f = ve_fields(10, 8, 4)
    eii = f["eii_plastic_pressure"]
    f["dilatation_grid"] = np.asfortranarray(np.where(np.arange(eii.size).reshape(eii.shape, order="F") % 2, 30.0, 90.0))
    o = _run(g, f)
    exp = np.where(f["dilatation_grid"] == 30.0, eii, 2.0 * eii)
    np.testing.assert_allclose(o["RC1"], exp, rtol=1e-15)
