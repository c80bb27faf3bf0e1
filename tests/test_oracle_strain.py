"""Pins of the oracle's grid strain rate and forecast stress
(GridStrainRate.jl:33-159, GridStress.jl) by closed forms on a graded grid:
a linear velocity field has exact difference quotients everywhere (ghost rows
included), a rigid rotation has no strain, and the forecast stress is the
Maxwell relation sigma = 2 eta_ve eps + chi sigma_old."""
import numpy as np

import oracle
from strain_cases import strain_fields

DT = 3.15e11


def test_linear_velocity_gives_uniform_strain_rate():
    a, b, c, d = 3e-15, -1e-15, 2e-15, 5e-16
    g, vx1, vy1, f = strain_fields(17, 11, 1, linear=(a, b, c, d))
    o = oracle.grid_strain_rate_and_stress(g, vx1, vy1, f, DT)
    np.testing.assert_allclose(o["exx"], 0.5 * (a - d), rtol=1e-9)
    np.testing.assert_allclose(o["exy"], 0.5 * (b + c), rtol=1e-9)
    np.testing.assert_allclose(o["eii"], np.hypot(0.5 * (a - d), 0.5 * (b + c)), rtol=1e-9)


def test_rigid_rotation_has_no_strain_and_keeps_only_the_old_stress():
    w = 1e-15
    g, vx1, vy1, f = strain_fields(13, 9, 2, linear=(0.0, -w, w, 0.0))
    o = oracle.grid_strain_rate_and_stress(g, vx1, vy1, f, DT)
    scale = w * 1e-6
    assert np.abs(o["exx"]).max() <= scale and np.abs(o["exy"]).max() <= scale
    chi_s = f["etas1"] / (f["etas1"] + DT * f["mus1"])
    chi_n = f["etan1"] / (f["etan1"] + DT * f["mun1"])
    np.testing.assert_allclose(o["sxy2"], chi_s * f["sxy1"], rtol=1e-6, atol=1e-9 * np.abs(f["sxy1"]).max())
    np.testing.assert_allclose(o["sxx2"], chi_n * f["sxx1"], rtol=1e-6, atol=1e-9 * np.abs(f["sxx1"]).max())


def test_forecast_stress_is_the_maxwell_relation():
    g, vx1, vy1, f = strain_fields(21, 15, 3)
    o = oracle.grid_strain_rate_and_stress(g, vx1, vy1, f, DT)
    for eta, mu, old, eps, new, ds in (("etas1", "mus1", "sxy1", "exy", "sxy2", "dsxy"),
                                       ("etan1", "mun1", "sxx1", "exx", "sxx2", "dsxx")):
        e, m = f[eta], f[mu]
        eta_ve = e * m * DT / (e + m * DT)        # viscous and elastic in series
        chi = eta_ve / (m * DT)                   # share of the old stress kept
        exp = 2.0 * eta_ve * o[eps] + chi * f[old]
        np.testing.assert_allclose(o[new], exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max())
        np.testing.assert_allclose(o[ds] + f[old], o[new], rtol=1e-12, atol=1e-6)
