"""Pins of the oracle's nodal plastic-yield update
(PlasticFailureNodes.jl:24-319) by closed forms: the Drucker-Prager yield
stress from the friction angle itself, the Maxwell return map (the new
viscosity brings the elastic trial stress onto the yield stress), the
pressure-to-node interpolation on constant and linear fields, the rules that
switch yielding off, and the count and RMS of the yield errors."""
import numpy as np

import oracle
from plastic_cases import PARAMS, plastic_fields


def _run(g, f, ye, **kw):
    par = dict(PARAMS, **kw)
    return oracle.update_nodes_for_plastic_yielding(g, f, par, ye)


def test_return_map_and_drucker_prager_on_linear_fields():
    """Uniform grid, pressure and sxx2 linear in x and y: at interior basic
    nodes the four-cell averages are the fields' values at the node, so the
    yield stress, the elastic trial stress and the new viscosity have closed
    forms; the new viscosity must also return the trial stress to the yield
    stress through the Maxwell update."""
    nx, ny = 21, 13
    g, f, ye = plastic_fields(nx, ny, 1, graded=False)
    xp, yp = np.meshgrid(g.gridx_pr, g.gridy_pr)
    f["pr1"] = np.asfortranarray(1e8 + 300.0 * xp + 500.0 * yp)
    f["sxx2"] = np.asfortranarray(2e7 - 100.0 * xp + 150.0 * yp)
    dt = PARAMS["timestep"]
    ev, psr, pyn, ye2, n, err = _run(g, f, ye, viscosity_min=1e10, viscosity_max=1e30)
    checked = 0
    for j in range(2, nx):
        for i in range(2, ny):
            x, y = g.gridx_b[j - 1], g.gridy_b[i - 1]
            P = 1e8 + 300.0 * x + 500.0 * y
            sxx = 2e7 - 100.0 * x + 150.0 * y
            sxy = f["sxy2"][i - 1, j - 1]
            phi = np.deg2rad(f["fric_degrees_grid"][i - 1, j - 1])
            c = f["cohesion_grid"][i - 1, j - 1]
            mu, eta_old, eta_flow = (f[k][i - 1, j - 1] for k in ("mus1", "etas1", "eta_flow"))
            if c == 0.0 and phi == 0.0:
                continue
            sii = np.hypot(sxy, sxx)
            sy = max(c * np.cos(phi) + np.sin(phi) * P, 0.0)
            se = sii * (mu * dt + eta_old) / eta_old
            eta_vp = dt * mu * sy / (se - sy) if sy < se else np.inf
            if eta_vp < eta_flow:
                assert pyn[i - 1, j - 1] == 1.0
                np.testing.assert_allclose(ev[i - 1, j - 1], eta_vp, rtol=1e-10)
                # Maxwell return map: the trial stress relaxes onto the yield stress
                np.testing.assert_allclose(se * ev[i - 1, j - 1] / (ev[i - 1, j - 1] + mu * dt), sy, rtol=1e-10)
                np.testing.assert_allclose(psr[i - 1, j - 1], 0.5 * sii / eta_vp - 0.5 * sii / eta_flow,
                                           rtol=1e-9, atol=1e-30)
                checked += 1
            else:
                assert pyn[i - 1, j - 1] == 0.0 and ev[i - 1, j - 1] == eta_flow and psr[i - 1, j - 1] == 0.0
    assert checked > 20


def test_constant_pressure_reaches_every_node_unchanged():
    """A constant pressure and sxx2 field interpolates to the same constant at
    corners, edges and interior (a dropped or doubled term would not)."""
    nx, ny = 9, 7
    g, f, ye = plastic_fields(nx, ny, 2)
    f["pr1"][:] = 2e8
    f["sxx2"][:] = 3e7
    f["sxy2"][:] = 4e7                 # stress invariant 5e7 everywhere
    f["fric_degrees_grid"][:] = 30.0
    f["cohesion_grid"][:] = 1e6
    f["mus1"][:] = 1e10
    f["etas1"][:] = 1e21
    f["eta_flow"][:] = 1e30
    dt = PARAMS["timestep"]
    ev, psr, pyn, ye2, n, err = _run(g, f, ye, viscosity_max=1e40)
    sy = 1e6 * np.cos(np.pi / 6) + 0.5 * 2e8
    se = 5e7 * (1e10 * dt + 1e21) / 1e21
    eta = dt * 1e10 * sy / (se - sy)
    assert se > sy
    np.testing.assert_allclose(ev, eta, rtol=1e-12)
    assert (pyn == 1.0).all() and n == nx * ny
    np.testing.assert_allclose(ye2, 5e7 - sy, rtol=1e-12)
    np.testing.assert_allclose(err, abs(5e7 - sy), rtol=1e-12)


def test_switches_that_disable_yielding():
    nx, ny = 31, 17
    g, f, ye = plastic_fields(nx, ny, 3)
    f["sxy2"][:] = 5e8                 # far above any yield stress
    f["eta_flow"][:] = 1e30
    base = _run(g, f, ye, viscosity_max=1e40)
    applicable = (f["cohesion_grid"] > 0) | (f["fric_degrees_grid"] > 0)
    assert (base[2][applicable] == 1.0).all() and (base[2][~applicable] == 0.0).all()
    assert base[4] == applicable.sum()
    # no yielding before the first time step
    ev, psr, pyn, ye2, n, err = _run(g, f, ye, ntimestep=0, viscosity_max=1e40)
    assert n == 0 and err == 0.0 and not pyn.any() and not psr.any()
    np.testing.assert_array_equal(ev, f["eta_flow"])
    np.testing.assert_array_equal(ye2, ye)
    # mobile wall moved by velocity_internal_x * timesum, and the plate extension beside it
    P = PARAMS
    xl = P["x_left_mobile_wall"] + P["velocity_internal_x"] * P["timesum"]
    xr = P["x_right_mobile_wall"] + P["velocity_internal_x"] * P["timesum"]
    X, Y = np.meshgrid(g.gridx_b, g.gridy_b)
    wall = (xl <= X) & (X <= xr) & (P["y_top_mobile_wall"] <= Y) & (Y <= P["y_bottom_mobile_wall"])
    ext = ((xl - P["plate_extension_width"] <= X) & (X <= xl)
           & (P["y_bottom_mobile_wall"] - P["plate_extension_thickness"] <= Y) & (Y <= P["y_bottom_mobile_wall"]))
    assert wall.any() and ext.any()
    pw = _run(g, f, ye, no_yielding_in_mobile_wall=1, viscosity_max=1e40)[2]
    assert not pw[wall].any() and (pw[applicable & ~wall] == 1.0).all()
    pe = _run(g, f, ye, no_yielding_in_plate_extension=1, viscosity_max=1e40)[2]
    assert not pe[ext].any() and (pe[applicable & ~ext] == 1.0).all()


def test_fluid_pressure_and_tension_cutoff():
    """With fluid pressure, nodes below sea level whose pressure is under the
    hydrostatic fluid pressure yield at cohesion * cos(phi); elsewhere the
    friction term acts on the pressure above the fluid pressure."""
    nx, ny = 11, 21
    g, f, ye = plastic_fields(nx, ny, 4, graded=False)
    f["pr1"][:] = 3e8
    f["sxx2"][:] = 0.0
    f["sxy2"][:] = 1e9
    f["fric_degrees_grid"][:] = 30.0
    f["cohesion_grid"][:] = 2e7
    f["eta_flow"][:] = 1e30
    f["mus1"][:] = 1e10
    f["etas1"][:] = 1e21
    dt = PARAMS["timestep"]
    ev = _run(g, f, ye, iuse_fluid_pressure_for_yield=1, viscosity_max=1e40)[0]
    se = 1e9 * (1e10 * dt + 1e21) / 1e21
    for i in range(2, ny):
        y = g.gridy_b[i - 1]
        pf = 1000.0 * 9.81 * max(0.0, y - PARAMS["y_sealevel"])
        sy = 2e7 * np.cos(np.pi / 6) + (0.0 if 3e8 < pf else 0.5 * (3e8 - pf))
        np.testing.assert_allclose(ev[i - 1, 5], dt * 1e10 * sy / (se - sy), rtol=1e-12)


def test_yield_count_and_rms_error():
    g, f, ye = plastic_fields(41, 23, 5)
    ev, psr, pyn, ye2, n, err = _run(g, f, ye)
    changed = ye2 != ye
    counted = (pyn == 1.0) | ((f["plastic_yield"] > 0) & ((f["cohesion_grid"] > 0) | (f["fric_degrees_grid"] > 0)))
    assert n == counted.sum() and n > 0
    assert not changed[~counted].any()
    np.testing.assert_allclose(err, np.sqrt(np.mean(ye2[counted] ** 2)), rtol=1e-12)


def test_harmonic_normal_viscosity_where_nodes_yield():
    """Cells beside a yielding node take the harmonic mean of their four basic
    viscosities (1/eta = mean of 1/eta_k; four equal values give that value);
    the others keep etan1."""
    g, f, _ = plastic_fields(12, 9, 6)
    ny, nx = g.ynum, g.xnum
    rng = np.random.default_rng(3)
    plastics = np.asfortranarray((rng.random((ny, nx)) < 0.1).astype(float))
    plasticn = np.asfortranarray((rng.random((ny - 1, nx - 1)) < 0.1).astype(float))
    pyield = np.asfortranarray((rng.random((ny, nx)) < 0.1).astype(float))
    etan1 = np.asfortranarray(rng.uniform(1e20, 1e22, (ny - 1, nx - 1)))
    out = oracle.harmonic_normal_viscosity(g, f["etas1"], plastics, plasticn, pyield, etan1)
    e = f["etas1"]
    for i in range(ny - 1):
        for j in range(nx - 1):
            near = (plasticn[i, j] > 0 or plastics[i:i + 2, j:j + 2].any() or pyield[i:i + 2, j:j + 2].any())
            if near:
                np.testing.assert_allclose(1.0 / out[i, j], np.mean(1.0 / e[i:i + 2, j:j + 2]), rtol=1e-14)
            else:
                assert out[i, j] == etan1[i, j]
    flat = np.asfortranarray(np.full((ny, nx), 3e21))
    out2 = oracle.harmonic_normal_viscosity(g, flat, np.ones((ny, nx), order="F"), plasticn, pyield, etan1)
    np.testing.assert_allclose(out2, 3e21, rtol=1e-15)


def test_plastic_strain_rate_at_pressure_nodes_is_the_cell_centre_value():
    """A bilinear field on the basic nodes averages to its value at the cell
    centre (mean of the four corner coordinates)."""
    g, _, _ = plastic_fields(14, 10, 7)
    X, Y = np.meshgrid(g.gridx_b, g.gridy_b)
    eii = np.asfortranarray(1e-15 + 2e-20 * X - 3e-20 * Y + 1e-25 * X * Y)
    out = oracle.plastic_strain_rate_at_pressure_nodes(g, eii)
    xc = 0.5 * (g.gridx_b[:-1] + g.gridx_b[1:])
    yc = 0.5 * (g.gridy_b[:-1] + g.gridy_b[1:])
    Xc, Yc = np.meshgrid(xc, yc)
    np.testing.assert_allclose(out, 1e-15 + 2e-20 * Xc - 3e-20 * Yc + 1e-25 * Xc * Yc, rtol=1e-12)
