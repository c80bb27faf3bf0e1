"""Pins of the heat-path oracle by closed forms, invariants and cross-checks
with already-pinned oracle parts.

The system rows are pinned by fields whose discrete operator is known
exactly: a linear temperature has zero flux divergence, a quadratic one a
constant divergence 2k on a graded grid (only with the right control-volume
widths), a linear conductivity times a linear temperature gives the product
of the slopes; boundary rows reproduce T_b - c*T_neighbour.  The solve is
pinned by the discrete sine eigenmode (backward Euler multiplies it by
1/(1 + dt*lambda/rhocp)), the heat loop by the same mode's closed-form
sub-steps, the residual by the quadratic field, the right-side terms by
constant fields, the heat scatter against the pinned Stokes scatter's
inclusive rule, the marker updates by constant and affine fields.
"""
import numpy as np
import pytest

import oracle
from earthbox_gpu.grid import graded_grid, uniform_grid
from heat_cases import (MAT_HR, MAT_MU, dirichlet_zero_bc, fixed_top_bottom_bc, grid_fields, heat_markers,
                        random_bc)


def _mesh(g):
    X = np.asfortranarray(np.repeat(g.gridx_b[None, :], g.ynum, axis=0))
    Y = np.asfortranarray(np.repeat(g.gridy_b[:, None], g.xnum, axis=1))
    return X, Y


def _interior(a):
    return a[1:-1, 1:-1]


def _LT(g, bc, T, kt, rhocp, dt):
    """L*T of the oracle's COO, as a grid array."""
    z = np.zeros((g.ynum, g.xnum), order="F")
    N, Li, Lj, Lv, R = oracle.heat_build(g, bc, z, z, rhocp, kt, dt)
    return oracle.coo_spmv(N, Li, Lj, Lv, T.ravel(order="F")).reshape((g.ynum, g.xnum), order="F"), (Li, R)


@pytest.mark.parametrize("axis", [0, 1])
def test_linear_temperature_has_no_flux(axis):
    g = graded_grid(23, 17, 3e5, 1e5, seed=3)
    X, Y = _mesh(g)
    T = np.asfortranarray(400.0 + 2e-3 * (X if axis else Y))
    kt = np.full((g.ynum, g.xnum), 2.5, order="F")
    rc, dt = 3e6, 1e12
    LT, _ = _LT(g, fixed_top_bottom_bc(g), T, kt, np.full_like(kt, rc, order="F"), dt)
    np.testing.assert_allclose(_interior(LT), _interior(T) * rc / dt, rtol=1e-13)


@pytest.mark.parametrize("axis", [0, 1])
def test_quadratic_temperature_divergence_on_graded_grid(axis):
    g = graded_grid(29, 19, 3e5, 1e5, seed=5)
    X, Y = _mesh(g)
    C = X if axis else Y
    T = np.asfortranarray(C * C)
    k = 2.5
    kt = np.full((g.ynum, g.xnum), k, order="F")
    rc, dt = 3e6, 1e12
    LT, _ = _LT(g, fixed_top_bottom_bc(g), T, kt, np.full_like(kt, rc, order="F"), dt)
    # L*T = rhocp/dt*T - k*(x_R - x_L)/xstpc = rhocp/dt*T - 2k
    err = _interior(LT) - (_interior(T) * rc / dt - 2.0 * k)
    assert np.max(np.abs(err)) <= 1e-12 * np.max(np.abs(_interior(T) * rc / dt))


def test_linear_conductivity_times_linear_temperature():
    g = graded_grid(21, 25, 2e5, 2e5, seed=9)
    X, Y = _mesh(g)
    b, c = 3e-3, 1e-5
    T = np.asfortranarray(300.0 + b * X + 0.5 * b * Y)
    kt = np.asfortranarray(2.0 + c * X + 2 * c * Y)
    rc, dt = 2.5e6, 5e11
    LT, _ = _LT(g, fixed_top_bottom_bc(g), T, kt, np.full_like(kt, rc, order="F"), dt)
    # -d/dx(k dT/dx) - d/dy(k dT/dy) = -(b*c) - (0.5b*2c)
    err = _interior(LT) - (_interior(T) * rc / dt - (b * c + 0.5 * b * 2 * c))
    assert np.max(np.abs(err)) <= 1e-12 * np.max(np.abs(_interior(T) * rc / dt))


def test_boundary_rows_and_counts():
    g = graded_grid(13, 11, 1e5, 5e4, seed=1)
    bc = random_bc(g, 2)
    rng = np.random.default_rng(4)
    T = np.asfortranarray(rng.uniform(200, 1500, (g.ynum, g.xnum)))
    f = grid_fields(g, 3)
    LT, (Li, R) = _LT(g, bc, T, f["kt1"], f["rhocp1"], 1e12)
    ny, nx = g.ynum, g.xnum
    assert Li.size == 5 * (nx - 2) * (ny - 2) + 2 * (2 * nx + 2 * ny - 4)
    Rg = R.reshape((ny, nx), order="F")
    np.testing.assert_array_equal(LT[0, 1:-1], T[0, 1:-1] - bc["btopt"][1:-1, 1] * T[1, 1:-1])
    np.testing.assert_array_equal(LT[-1, 1:-1], T[-1, 1:-1] - bc["bbottomt"][1:-1, 1] * T[-2, 1:-1])
    np.testing.assert_array_equal(LT[:, 0], T[:, 0] - bc["bleftt"][:, 1] * T[:, 1])
    np.testing.assert_array_equal(LT[:, -1], T[:, -1] - bc["brightt"][:, 1] * T[:, -2])
    np.testing.assert_array_equal(Rg[0, 1:-1], bc["btopt"][1:-1, 0])
    np.testing.assert_array_equal(Rg[:, 0], bc["bleftt"][:, 0])
    np.testing.assert_array_equal(Rg[:, -1], bc["brightt"][:, 0])


def _sine_mode(g):
    X, Y = _mesh(g)
    return np.asfortranarray(np.sin(np.pi * X / g.xsize) * np.sin(np.pi * Y / g.ysize))


def _sine_lambda(g, k):
    hx, hy = g.xsize / (g.xnum - 1), g.ysize / (g.ynum - 1)
    return k * (4 / hx**2 * np.sin(np.pi * hx / (2 * g.xsize)) ** 2 +
                4 / hy**2 * np.sin(np.pi * hy / (2 * g.ysize)) ** 2)


def test_solve_decays_the_discrete_sine_mode():
    g = uniform_grid(33, 25, 2e5, 1e5)
    k, rc, dt = 3.0, 3e6, 3e15
    T0 = 800.0 * _sine_mode(g)
    kt = np.full((g.ynum, g.xnum), k, order="F")
    z = np.zeros_like(kt)
    T2 = oracle.heat_solve(g, dirichlet_zero_bc(g), z, T0, np.full_like(kt, rc, order="F"), kt, dt)
    factor = 1.0 / (1.0 + dt * _sine_lambda(g, k) / rc)
    np.testing.assert_allclose(T2, T0 * factor, atol=1e-11 * 800.0)
    assert 0.2 < factor < 0.9   # the step diffuses visibly


def test_residual_closed_form_and_zero_at_the_solution():
    g = graded_grid(19, 23, 1e5, 1e5, seed=7)
    X, _ = _mesh(g)
    k, rc, dt = 2.0, 3e6, 1e12
    kt = np.full((g.ynum, g.xnum), k, order="F")
    rhocp = np.full_like(kt, rc, order="F")
    rng = np.random.default_rng(1)
    RT = np.asfortranarray(rng.uniform(0, 1e-6, kt.shape))
    T0 = np.asfortranarray(rng.uniform(300, 400, kt.shape))
    T2 = np.asfortranarray(X * X * 1e-8)
    rest = oracle.heat_residuals(g, RT, rhocp, T0, T2, kt, dt)
    exp = RT - rhocp * (T2 - T0) / dt + 2 * k * 1e-8
    np.testing.assert_allclose(_interior(rest), _interior(exp), rtol=1e-9, atol=1e-12)
    assert np.all(rest[0, :] == 0) and np.all(rest[:, -1] == 0)
    f = grid_fields(g, 5)
    bc = fixed_top_bottom_bc(g)
    T2 = oracle.heat_solve(g, bc, RT, f["tk1"], f["rhocp1"], f["kt1"], dt)
    rest = oracle.heat_residuals(g, RT, f["rhocp1"], f["tk1"], T2, f["kt1"], dt)
    scale = np.max(np.abs(f["rhocp1"] * f["tk1"] / dt))
    assert np.max(np.abs(rest)) <= 1e-12 * scale


def test_heat_loop_sub_steps_of_the_sine_mode():
    g = uniform_grid(25, 21, 2e5, 1e5)
    k, rc = 3.0, 3e6
    lam = _sine_lambda(g, k)
    T0 = 500.0 * _sine_mode(g)
    kt = np.full((g.ynum, g.xnum), k, order="F")
    z = np.zeros_like(kt)
    timestep, limit = 1e15, 60.0
    tk2, rest, log = oracle.heat_loop(g, dirichlet_zero_bc(g), z, T0, np.full_like(kt, rc, order="F"), kt,
                                      timestep, limit)
    # closed form of the same sub-steps: amplitude a, change a*(1 - f(dt))
    amp = np.max(np.abs(T0))
    f = lambda dt: 1.0 / (1.0 + dt * lam / rc)
    dt_loop, total, prod, steps = timestep, 0.0, 1.0, []
    while total < timestep:
        change = amp * (1 - f(dt_loop))
        if change > limit:
            dt_loop = dt_loop * limit / change
        amp *= f(dt_loop)
        prod *= f(dt_loop)
        steps.append(dt_loop)
        total += dt_loop
        dt_loop = min(dt_loop, timestep - total)
    assert len(log) == len(steps) >= 3
    np.testing.assert_allclose([s[0] for s in log], steps, rtol=1e-9)
    assert abs(sum(s[0] for s in log) - timestep) <= 1e-9 * timestep
    np.testing.assert_allclose(tk2, T0 * prod, atol=1e-9 * 500)


def test_rhs_constant_fields():
    g = uniform_grid(9, 7, 1e4, 1e4)
    ny, nx = g.ynum, g.xnum
    c = lambda v, s: np.full(s, v, order="F")
    f = {"hr1": c(1e-6, (ny, nx)), "ha1": c(0.03, (ny, nx)), "rho1": c(3000.0, (ny, nx)),
         "vx1": c(2e-10, (ny + 1, nx)), "vy1": c(-1e-10, (ny, nx + 1)), "sxy2": c(1e7, (ny, nx)),
         "etas1": c(1e21, (ny, nx)), "sxx2": c(-2e7, (ny - 1, nx - 1)), "etan1": c(4e21, (ny - 1, nx - 1))}
    gx, gy = 0.5, 9.81
    for ia, ish in ((0, 0), (1, 0), (0, 1), (1, 1)):
        RT = oracle.heat_rhs(g, f, gx, gy, ia, ish)
        exp = 1e-6 + ia * 0.03 * 3000.0 * (gx * 2e-10 + gy * -1e-10) + ish * (1e14 / 1e21 + 4e14 / 4e21)
        np.testing.assert_allclose(_interior(RT), exp, rtol=1e-13)
        assert np.all(RT[0, :] == 1e-6) and np.all(RT[:, 0] == 1e-6)


def test_heat_scatter_matches_the_pinned_inclusive_rule():
    g = graded_grid(31, 17, 3e5, 1e5, seed=2)
    m = heat_markers(g, 4, 11)
    ny, nx = g.ynum, g.xnum
    rng = np.random.default_rng(3)
    prev = {k: np.asfortranarray(rng.uniform(1, 2, (ny, nx))) for k in ("tk2", "kt1", "rhocp1", "hr1", "ha1")}
    # a node whose four cells hold no inside marker keeps its backups
    empty = np.isin(m["marker_xn"], [5, 6]) & np.isin(m["marker_yn"], [7, 8])
    m["flags"][empty] = -1
    out = oracle.scatter_heat(g, m, prev, MAT_HR, m["serp"])
    hr = MAT_HR[m["matid"] - 1] + m["serp"]
    for key, val in (("tk1", m["TK"]), ("kt1", m["kt"]), ("rhocp1", m["rhocp"]), ("hr1", hr), ("ha1", m["ha"])):
        st = oracle.scatter_stokes(g, {"x": m["x"], "y": m["y"], "inside_flags": m["flags"], "rho": val},
                                   {"rho1": prev["tk2" if key == "tk1" else key]})
        np.testing.assert_array_equal(out[key], st["rho1"])
        np.testing.assert_array_equal(out["wtnodes"], st["wtnodes"])
    assert (out["wtnodes"] == 0).any()


def test_thermal_bc_closed_form():
    g = uniform_grid(7, 6, 1e4, 1e4)
    bc = random_bc(g, 5)
    rng = np.random.default_rng(1)
    t = np.asfortranarray(rng.uniform(200, 300, (g.ynum, g.xnum)))
    o = oracle.apply_thermal_bc(g, bc, t)
    np.testing.assert_array_equal(o[1:-1, 1:-1], t[1:-1, 1:-1])
    np.testing.assert_array_equal(o[0, 1:-1], bc["btopt"][1:-1, 0] + bc["btopt"][1:-1, 1] * t[1, 1:-1])
    np.testing.assert_array_equal(o[-1, 1:-1], bc["bbottomt"][1:-1, 0] + bc["bbottomt"][1:-1, 1] * t[-2, 1:-1])
    # the sides run after the top and bottom rows, over every row
    np.testing.assert_array_equal(o[:, 0], bc["bleftt"][:, 0] + bc["bleftt"][:, 1] * o[:, 1])
    np.testing.assert_array_equal(o[:, -1], bc["brightt"][:, 0] + bc["brightt"][:, 1] * o[:, -2])


def test_grid_to_marker_temperature_reproduces_affine_fields():
    g = graded_grid(17, 13, 1e5, 5e4, seed=8)
    m = heat_markers(g, 3, 4)
    X, Y = _mesh(g)
    field = np.asfortranarray(300.0 + 2e-3 * X - 5e-3 * Y)
    t = oracle.grid_to_marker_temperature(g, m, field, 0, m["TK"])
    ins = m["flags"] == 1
    np.testing.assert_allclose(t[ins], 300.0 + 2e-3 * m["x"][ins] - 5e-3 * m["y"][ins], rtol=1e-13)
    assert np.array_equal(t[~ins], m["TK"][~ins])
    t2 = oracle.grid_to_marker_temperature(g, m, field, 1, m["TK"])
    np.testing.assert_allclose(t2[ins], m["TK"][ins] + t[ins], rtol=1e-13)


def test_subgrid_temperature_closed_form_and_clamp():
    g = uniform_grid(11, 9, 1e5, 8e4)
    m = heat_markers(g, 3, 5)
    ny, nx = g.ynum, g.xnum
    k, rc, tn = 2.5, 3e6, 900.0
    c = lambda v: np.full((ny, nx), v, order="F")
    for dt, coef in ((1e12, 1.0), (1e16, 1.0)):
        t, dT = oracle.subgrid_temperature_markers(g, m, c(tn), c(k), c(rc), dt, coef)
        tsg = rc / k / (2 / g.xstpavg**2 + 2 / g.ystpavg**2)
        term = max(-coef * dt / tsg, -30.0)
        ins = m["flags"] == 1
        np.testing.assert_allclose(dT[ins], (tn - m["TK"][ins]) * (1 - np.exp(term)), rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(t[ins], m["TK"][ins] + dT[ins], rtol=1e-15)
        assert np.all(dT[~ins] == 0)
    assert term == -30.0


def test_update_marker_temperature_composition():
    """With tk1 reproduced by the markers and no sources, the marker update
    equals the grid change interpolated to markers (plus the sub-grid part)."""
    g = uniform_grid(21, 17, 2e5, 1e5)
    m = heat_markers(g, 3, 6, outside=0)
    f = grid_fields(g, 6)
    bc = fixed_top_bottom_bc(g)
    r = oracle.update_marker_temp_for_grid_and_subgrid_changes(g, bc, m, f, 3e12, 50.0, 0.0)
    exp = oracle.grid_to_marker_temperature(g, m, r["tk2"] - f["tk1"], 1, m["TK"])
    np.testing.assert_array_equal(r["TK"], exp)
    r2 = oracle.update_marker_temp_for_grid_and_subgrid_changes(g, bc, m, f, 3e12, 50.0, 1.0)
    assert r2["dtkn"] is not None and np.abs(r2["dtkn"]).max() > 0
    np.testing.assert_allclose(r2["dtk1"], r["dtk1"] - r2["dtkn"], rtol=0, atol=1e-9)


def test_sticky_correction():
    g = uniform_grid(5, 5, 1e4, 1e4)
    m = heat_markers(g, 2, 1)
    t = oracle.sticky_temperature_correction(m, 1, 3, 273.0)
    sticky = (m["flags"] == 1) & np.isin(m["matid"], [1, 3])
    assert np.all(t[sticky] == 273.0)
    assert np.array_equal(t[~sticky], m["TK"][~sticky])


def test_subgrid_stress_closed_form():
    g = graded_grid(15, 13, 1e5, 5e4, seed=6)
    m = heat_markers(g, 3, 7)
    m["mat_mu"] = MAT_MU
    ny, nx = g.ynum, g.xnum
    sxy1 = np.full((ny, nx), 2e7, order="F")
    sxx1 = np.full((ny - 1, nx - 1), -1e7, order="F")
    dt, coef = 3e12, 1.0
    rng = np.random.default_rng(2)
    f = {"sxy1": sxy1, "sxx1": sxx1, "dsxy": np.asfortranarray(rng.uniform(-1e6, 1e6, (ny, nx))),
         "dsxx": np.asfortranarray(rng.uniform(-1e6, 1e6, (ny - 1, nx - 1)))}
    r = oracle.update_subgrid_stress(g, m, f, dt, coef)
    ins = m["flags"] == 1
    term = 1 - np.exp(np.maximum(-coef * dt / (m["eta"] / MAT_MU[m["matid"] - 1]), -30.0))
    np.testing.assert_allclose(r["sxy"][ins], m["sxy"][ins] + (2e7 - m["sxy"][ins]) * term[ins], rtol=1e-12)
    np.testing.assert_allclose(r["sxx"][ins], m["sxx"][ins] + (-1e7 - m["sxx"][ins]) * term[ins], rtol=1e-12)
    assert np.array_equal(r["sxy"][~ins], m["sxy"][~ins])
    # remaining change and zero where no weight
    np.testing.assert_array_equal(r["dsxy"], f["dsxy"] - r["dsxyn"])
    np.testing.assert_array_equal(r["dsxx"], f["dsxx"] - r["dsxxn"])
    assert np.all(r["dsxyn"][r["wtetas"] == 0] == 0) and np.all(r["dsxxn"][r["wtetan"] == 0] == 0)
    # markers that all relax by the same amount give that amount on every weighted node
    m2 = dict(m, sxy=np.full(m["x"].size, 1e7), eta=np.full(m["x"].size, 1e21), matid=np.ones_like(m["matid"]))
    r2 = oracle.update_subgrid_stress(g, m2, f, dt, coef)
    d = (2e7 - 1e7) * (1 - np.exp(-coef * dt / (1e21 / MAT_MU[0])))
    w = r2["wtetas"] > 0
    np.testing.assert_allclose(r2["dsxyn"][w], d, rtol=1e-13)
