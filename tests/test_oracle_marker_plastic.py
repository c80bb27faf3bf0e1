"""Pins of the marker plastic-yielding oracle
(MarkerPlasticity.update_plastic_yielding!, ViscoelasticForecast.jl:14-127,
ElasticForecast.jl:14-110) by what the Gerya (2010, 2019) corrections fix:

- viscoelastic forecast: a yielding marker's stress has invariant exactly the
  yield stress and keeps its direction, and its viscosity times twice the
  effective strain rate is the yield stress;
- elastic forecast: the new viscosity's viscoelastic factor times the elastic
  forecast is the yield stress, and stresses are left alone;
- the Mohr-Coulomb yield stress written with sqrt(1 - sin^2) (not asin/cos),
  the fluid-pressure variant, both limit pairs, the mobile-wall exclusion,
  ntimestep 0, zero strength, outside markers and the sticky pfailure flag.
"""
import numpy as np
import pytest

import oracle
from marker_plastic_cases import AIR, MAT_MU, PARAMS, PLATE, WALL, markers


def _sy(m, par, use_fluid=0):
    """Mohr-Coulomb yield stress (YieldStress.jl) with cos(asin(f)) = sqrt(1 - f^2)."""
    c, f, p = m["cohesion"], m["fric"], m["pr"]
    if use_fluid:
        pf = 1000.0 * 9.81 * np.maximum(0.0, m["y"] - par["y_sealevel"])
        p = np.where(p < pf, 0.0, p - pf)
    return np.maximum(c * np.sqrt(1.0 - f * f) + f * p, 0.0)


def _mu(m):
    return MAT_MU[m["matid"] - 1]


def _forecast(m, par, kind):
    mu, dt, r = _mu(m), par["timestep"], m["sr_ratio"]
    if kind == 0:
        z = m["eta"] / (m["eta"] + dt * mu)
        return np.hypot(m["sxx"] * z + 2 * m["eta"] * m["exx"] * r * (1 - z),
                        m["sxy"] * z + 2 * m["eta"] * m["exy"] * r * (1 - z))
    return np.hypot(m["sxx"] + 2 * mu * dt * m["exx"] * r, m["sxy"] + 2 * mu * dt * m["exy"] * r)


def _run(m, par, kind):
    p = dict(par, forecast=kind)
    return oracle.update_marker_plastic_yielding(m, p)


@pytest.mark.parametrize("kind", [0, 1])
def test_yielding_markers_sit_on_the_yield_surface(kind):
    m = markers(4000, 11)
    par = dict(PARAMS, viscosity_min=1e-30, viscosity_max=1e60)
    sxx, sxy, eta, pf = _run(m, par, kind)
    sy = _sy(m, par)
    applicable = (m["cohesion"] > 0) | (m["fric"] > 0)
    inside = m["flags"] == 1
    F = _forecast(m, par, kind)
    yld = inside & applicable & (F > sy * (1 + 1e-12))
    keep = inside & ~(applicable & (F > sy * (1 - 1e-12)))
    assert 300 < yld.sum() < 3700 and keep.sum() > 300
    mu, dt = _mu(m), par["timestep"]
    if kind == 0:
        nz = yld & (np.hypot(m["sxx"], m["sxy"]) > 0)
        np.testing.assert_allclose(np.hypot(sxx, sxy)[nz], sy[nz], rtol=1e-13)
        # direction kept: the cross product of old and new stress vanishes
        cross = (m["sxx"] * sxy - m["sxy"] * sxx)[nz] / (np.hypot(m["sxx"], m["sxy"]) * sy)[nz]
        assert np.abs(cross).max() < 1e-13
        assert np.all(np.sign(sxx[nz]) == np.sign(m["sxx"][nz]))
        z0 = yld & (np.hypot(m["sxx"], m["sxy"]) == 0)
        assert z0.sum() > 0
        np.testing.assert_allclose(sxx[z0], sy[z0], rtol=1e-14)
        np.testing.assert_array_equal(sxy[z0], sxx[z0])
        eii = m["sr_ratio"] * np.hypot(m["exx"], m["exy"])
        np.testing.assert_allclose((2 * eta * eii)[yld], sy[yld], rtol=1e-13)
    else:
        np.testing.assert_array_equal(sxx, m["sxx"])
        np.testing.assert_array_equal(sxy, m["sxy"])
        z = eta / (eta + dt * mu)
        np.testing.assert_allclose((F * z)[yld], sy[yld], rtol=1e-10)
    np.testing.assert_array_equal(eta[keep], m["eta"][keep])
    np.testing.assert_array_equal(sxx[keep], m["sxx"][keep])
    assert np.all(pf[yld] == 1.0)
    np.testing.assert_array_equal(pf[keep], np.where(m["pfailure"][keep] == 1.0, 1.0, 0.0).astype(np.float32))
    np.testing.assert_array_equal(pf[~inside], m["pfailure"][~inside])
    np.testing.assert_array_equal(eta[~inside], m["eta"][~inside])


def test_viscosity_and_yield_stress_limits():
    m = markers(3000, 12)
    par = dict(PARAMS, viscosity_min=1e20, viscosity_max=1e23, yield_stress_min=5e6, yield_stress_max=5e7)
    sxx, sxy, eta, pf = _run(m, par, 0)
    sy = np.clip(_sy(m, par), 5e6, 5e7)
    F = _forecast(m, par, 0)
    yld = (m["flags"] == 1) & ((m["cohesion"] > 0) | (m["fric"] > 0)) & (F > sy * (1 + 1e-12))
    nz = yld & (np.hypot(m["sxx"], m["sxy"]) > 0)
    np.testing.assert_allclose(np.hypot(sxx, sxy)[nz], sy[nz], rtol=1e-13)
    eii = m["sr_ratio"] * np.hypot(m["exx"], m["exy"])
    np.testing.assert_allclose(eta[yld], np.clip(sy / 2 / eii, 1e20, 1e23)[yld], rtol=1e-13)
    assert (eta[yld] == 1e20).any() and (eta[yld] == 1e23).any()


def test_fluid_pressure_yield_stress():
    m = markers(3000, 13)
    m["pr"] = np.random.default_rng(5).uniform(0, 2e9, m["pr"].size)
    par = dict(PARAMS, iuse_fluid_pressure_for_yield=1, viscosity_min=1e-30, viscosity_max=1e60)
    sxx, sxy, eta, pf = _run(m, par, 0)
    sy = _sy(m, par, use_fluid=1)
    F = _forecast(m, par, 0)
    yld = (m["flags"] == 1) & ((m["cohesion"] > 0) | (m["fric"] > 0)) & (F > sy * (1 + 1e-12))
    nz = yld & (np.hypot(m["sxx"], m["sxy"]) > 0)
    np.testing.assert_allclose(np.hypot(sxx, sxy)[nz], sy[nz], rtol=1e-13)
    # fluid pressure lowers the yield stress of deep markers only
    deep = m["y"] > par["y_sealevel"]
    assert (sy[deep] < _sy(m, par)[deep]).any() and np.all(sy[~deep] == _sy(m, par)[~deep])


def test_mobile_wall_and_sandbox_domains_do_not_yield():
    m = markers(3000, 14)
    par = dict(PARAMS, no_yielding_in_mobile_wall=1, matid_sticky_air=AIR, matid_mobile_wall=WALL,
               matid_plate_extension=PLATE)
    sxx, sxy, eta, pf = _run(m, par, 0)
    excl = np.isin(m["matid"], [AIR, WALL, PLATE]) & (m["flags"] == 1)
    np.testing.assert_array_equal(eta[excl], m["eta"][excl])
    np.testing.assert_array_equal(sxx[excl], m["sxx"][excl])
    np.testing.assert_array_equal(pf[excl], np.where(m["pfailure"][excl] == 1, 1, 0).astype(np.float32))
    free = sxx, sxy, eta, pf
    ref = _run(m, dict(PARAMS), 0)
    mat1 = (m["matid"] == 1)
    for a, b in zip(free, ref):
        np.testing.assert_array_equal(a[mat1], b[mat1])
    assert not np.array_equal(ref[2][excl], m["eta"][excl])


def test_first_timestep_and_zero_strength_do_not_yield():
    m = markers(2000, 15)
    sxx, sxy, eta, pf = _run(m, dict(PARAMS, ntimestep=0), 0)
    np.testing.assert_array_equal(eta, m["eta"])
    np.testing.assert_array_equal(sxx, m["sxx"])
    inside = m["flags"] == 1
    np.testing.assert_array_equal(pf[inside], np.where(m["pfailure"][inside] == 1, 1, 0).astype(np.float32))
    sxx, sxy, eta, pf = _run(m, PARAMS, 1)
    weak = (m["cohesion"] == 0) & (m["fric"] == 0)
    np.testing.assert_array_equal(eta[weak], m["eta"][weak])


def test_forecast_threshold_by_stress_only_markers():
    """With zero strain rate the viscoelastic forecast is |s|*z: markers at
    |s|*z = sy*(1 +- 1e-9) fall on either side of the yield test."""
    n = 200
    m = markers(n, 16)
    m["flags"][:] = 1
    m["exx"][:] = 0.0
    m["exy"][:] = 0.0
    m["pfailure"][:] = 0
    m["cohesion"][:] = 2e7
    m["fric"][:] = 0.0
    mu, dt = _mu(m), PARAMS["timestep"]
    z = m["eta"] / (m["eta"] + dt * mu)
    side = np.where(np.arange(n) % 2 == 0, 1 + 1e-9, 1 - 1e-9)
    ang = np.random.default_rng(3).uniform(0, 2 * np.pi, n)
    s = 2e7 / z * side
    m["sxx"], m["sxy"] = s * np.cos(ang), s * np.sin(ang)
    sxx, sxy, eta, pf = _run(m, PARAMS, 0)
    np.testing.assert_array_equal(pf, (side > 1).astype(np.float32))
