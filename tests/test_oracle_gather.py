"""Pins of the oracle's grid -> marker gathers after the solve
(MarkerStrainRate.jl, MarkerStrainRateRatio.jl, MarkerPressure.jl) by what the
mathematics fixes: bilinear interpolation reproduces affine fields, its error
for a quadratic field names the cell it used, and the strain-rate ratio has
closed forms for uniform fields."""
import numpy as np

import oracle
from gather_cases import BASIC, PRESSURE, gather_case, node_coords


def _affine(g, m, f, coef):
    a, b, c = coef
    out = dict(f)
    for kind, keys in (("pressure", PRESSURE), ("basic", BASIC)):
        Y, X = node_coords(g, kind)
        for k in keys:
            out[k] = np.asfortranarray(a + b * X + c * Y)
    return out


def test_affine_fields_reproduced_everywhere():
    """Every marker, edge half-cells (where the pressure cell is clamped and
    the value extrapolated) included, recovers a + b x + c y."""
    g, m, f = gather_case(23, 17, 4000, 1)
    coef = (3.0, 2.0e-5, -7.0e-5)
    r = oracle.marker_strain_rate_and_pressure(g, m, _affine(g, m, f, coef), 1.0)
    ins = m["inside"] == 1
    exp = coef[0] + coef[1] * m["x"] + coef[2] * m["y"]
    for k in ("marker_exx", "marker_exy", "marker_pr"):
        np.testing.assert_allclose(r[k][ins], exp[ins], rtol=0, atol=1e-13 * np.abs(exp).max(), err_msg=k)


def test_basic_distances_are_normalized_offsets():
    g, m, f = gather_case(19, 13, 2000, 2)
    r = oracle.marker_strain_rate_and_pressure(g, m, f, 1.0)
    ins = m["inside"] == 1
    xn, yn = m["xn"][ins] - 1, m["yn"][ins] - 1
    np.testing.assert_allclose(r["marker_dx"][ins] * g.xstp_b[xn] + g.gridx_b[xn], m["x"][ins], rtol=1e-15)
    np.testing.assert_allclose(r["marker_dy"][ins] * g.ystp_b[yn] + g.gridy_b[yn], m["y"][ins], rtol=1e-15)
    assert np.all((r["marker_dx"][ins] >= 0) & (r["marker_dx"][ins] <= 1))


def _interp_error(P, t):
    """(t - P[k]) (P[k+1] - t) for the upper-left node k = the number of nodes
    at or below t, clamped to [1, len(P)-1] (1-based): the exact error of
    linear interpolation of t^2 between P[k] and P[k+1]."""
    k = np.clip(np.searchsorted(P, t, side="right"), 1, P.size - 1) - 1
    return (t - P[k]) * (P[k + 1] - t)


def test_quadratic_field_names_the_pressure_cell():
    """For p = x^2 + y^2 on the pressure nodes the bilinear value exceeds
    x^2 + y^2 by exactly the two 1-D interpolation errors of the cell used, so
    a wrong pressure-cell choice (off by one, wrong clamp) shows."""
    g, m, f = gather_case(21, 15, 4000, 3)
    Y, X = node_coords(g, "pressure")
    h = dict(f, pr1=np.asfortranarray((X * 1e-4) ** 2 + (Y * 1e-4) ** 2))
    r = oracle.marker_strain_rate_and_pressure(g, m, h, 1.0)
    ins = m["inside"] == 1
    xs, ys = m["x"][ins] * 1e-4, m["y"][ins] * 1e-4
    exp = xs ** 2 + ys ** 2 + _interp_error(g.gridx_vy[1:-1] * 1e-4, xs) + \
        _interp_error(g.gridy_vx[1:-1] * 1e-4, ys)
    np.testing.assert_allclose(r["marker_pr"][ins], exp, rtol=1e-12, atol=1e-12)


def test_quadratic_field_names_the_basic_cell():
    g, m, f = gather_case(21, 15, 4000, 4)
    Y, X = node_coords(g, "basic")
    h = dict(f, exy=np.asfortranarray((X * 1e-4) ** 2 - 2.0 * (Y * 1e-4) ** 2))
    r = oracle.marker_strain_rate_and_pressure(g, m, h, 1.0)
    ins = m["inside"] == 1
    xs, ys = m["x"][ins] * 1e-4, m["y"][ins] * 1e-4
    exp = xs ** 2 - 2.0 * ys ** 2 + _interp_error(g.gridx_b * 1e-4, xs) - \
        2.0 * _interp_error(g.gridy_b * 1e-4, ys)
    np.testing.assert_allclose(r["marker_exy"][ins], exp, rtol=1e-12, atol=1e-12)


def _uniform(g, f, vals):
    out = dict(f)
    for k, v in vals.items():
        out[k] = np.asfortranarray(np.full(f[k].shape, v))
    return out


def test_ratio_is_one_for_purely_viscous_forecast():
    """sxx2 = 2 eta exx, sxy2 = 2 eta exy, no stress change: the stress-based
    strain rate equals the velocity-based one."""
    g, m, f = gather_case(17, 11, 3000, 5)
    eta = 3.0e21
    m = dict(m, eta=np.full(m["x"].size, eta))
    e1, e2 = 2.0e-15, -1.5e-15
    h = _uniform(g, f, {"exx": e1, "exy": e2, "sxx2": 2 * eta * e1, "sxy2": 2 * eta * e2,
                        "dsxx": 0.0, "dsxy": 0.0})
    r = oracle.marker_strain_rate_and_pressure(g, m, h, 3.15e12)
    ins = m["inside"] == 1
    np.testing.assert_allclose(r["marker_sr_ratio"][ins], 1.0, rtol=1e-13)


def test_ratio_elastic_closed_form_per_material():
    """No forecast stress, uniform stress change c: ratio = |c| / (2 dt mu_m
    eii) with mu_m the marker's material modulus (1-based material ids)."""
    g, m, f = gather_case(17, 11, 3000, 6)
    dt, c1, c2, e1, e2 = 2.0e12, 4.0e5, -3.0e5, 1.0e-15, 2.0e-15
    h = _uniform(g, f, {"exx": e1, "exy": e2, "sxx2": 0.0, "sxy2": 0.0, "dsxx": c1, "dsxy": c2})
    r = oracle.marker_strain_rate_and_pressure(g, m, h, dt)
    ins = m["inside"] == 1
    mu = m["mat_mu"][m["matid"][ins].astype(int) - 1]
    exp = 0.5 * np.hypot(c1, c2) / (dt * mu) / np.hypot(e1, e2)
    np.testing.assert_allclose(r["marker_sr_ratio"][ins], exp, rtol=1e-13)


def test_ratio_defaults_to_one_without_strain_and_outside_markers_keep_values():
    g, m, f = gather_case(17, 11, 500, 7)
    h = _uniform(g, f, {"exx": 0.0, "exy": 0.0})
    prev = {k: np.full(500, -9.0) for k in oracle.GATHER_OUT}
    r = oracle.marker_strain_rate_and_pressure(g, m, h, 1e12, prev=prev)
    ins = m["inside"] == 1
    assert np.all(r["marker_sr_ratio"][ins] == 1.0)
    for k in oracle.GATHER_OUT:
        assert np.all(r[k][~ins] == -9.0), k
    assert (~ins).sum() >= 2


def test_stress_correction_adds_affine_change_and_skips_outside_markers():
    """StressCorrection: marker_sxx += dsxx and marker_sxy += dsxy at the
    marker; affine stress changes on either grid are recovered exactly up to
    rounding, and markers outside the grid are left alone."""
    g, m, f = gather_case(23, 17, 4000, 8)
    rng = np.random.default_rng(8)
    sxx0, sxy0 = 1e6 * rng.standard_normal(4000), 1e6 * rng.standard_normal(4000)
    cx, cy = (2.0e5, 3.0, -1.0), (-1.0e5, -2.0, 4.0)
    h = dict(f)
    Y, X = node_coords(g, "pressure")
    h["dsxx"] = np.asfortranarray(cx[0] + cx[1] * X + cx[2] * Y)
    Y, X = node_coords(g, "basic")
    h["dsxy"] = np.asfortranarray(cy[0] + cy[1] * X + cy[2] * Y)
    r = oracle.marker_strain_rate_and_pressure(g, dict(m, marker_sxx=sxx0, marker_sxy=sxy0), h, 1.0)
    ins = m["inside"] == 1
    ex = cx[0] + cx[1] * m["x"] + cx[2] * m["y"]
    ey = cy[0] + cy[1] * m["x"] + cy[2] * m["y"]
    np.testing.assert_allclose(r["marker_sxx"][ins] - sxx0[ins], ex[ins], rtol=0, atol=1e-9 * 1e6)
    np.testing.assert_allclose(r["marker_sxy"][ins] - sxy0[ins], ey[ins], rtol=0, atol=1e-9 * 1e6)
    assert np.array_equal(r["marker_sxx"][~ins], sxx0[~ins])
    assert np.array_equal(r["marker_sxy"][~ins], sxy0[~ins])


def test_stress_correction_refused_with_subgrid_diffusion():
    """With a positive sub-grid stress diffusion coefficient the reference runs
    update_subgrid_stress! before the stress correction
    (StokesContinuitySolver.jl:365-381); the host wrapper refuses the fused
    correction then (checked before any device work)."""
    import types
    import pytest
    from earthbox_gpu.marker_gather import STRESS, update_markers_after_stokes
    with pytest.raises(ValueError, match="subgrid_diff_coef_stress"):
        update_markers_after_stokes(types.SimpleNamespace(grid=None), {}, {}, 1.0, out={}, outputs=STRESS,
                                    subgrid_diff_coef_stress=1.0)
