"""GPU parity of the heat path with the oracle, through the C ABI.

Bars: the heat scatter, thermal boundary conditions, right-hand side, the
system rows (L*T and R), residuals, the grid -> marker temperature updates
and the sticky correction are bit-exact; the solve stands in for a direct
solve and must land within 1e-10 of the oracle's direct solution (relative to
the largest temperature), with the FP64 residual ||R - L*T|| / ||R|| at most
1e-12; the heat loop takes the oracle's sub-steps (time steps within 1e-9
relative, which only the solve's last bits can move); the sub-grid updates
go through exp of the device library and agree to 1e-13 relative.
"""
import numpy as np
import pytest
import torch

import oracle
from earthbox_gpu import heat
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.grid import graded_grid, uniform_grid
from earthbox_gpu.marker_gather import update_subgrid_stress
from heat_cases import (MAT_HR, MAT_MU, dirichlet_zero_bc, fixed_top_bottom_bc, grid_fields, heat_markers,
                        random_bc)

pytestmark = pytest.mark.gpu


def _dev(d):
    return {k: to_device(np.asfortranarray(v) if np.ndim(v) == 2 else v) for k, v in d.items()}


def _air_layer(f, g, kair=20.0):
    """Sticky-air-like top layer: high conductivity, low rho*Cp."""
    f = dict(f)
    na = max(1, g.ynum // 8)
    f["kt1"] = np.asfortranarray(f["kt1"].copy())
    f["rhocp1"] = np.asfortranarray(f["rhocp1"].copy())
    f["kt1"][:na, :] = kair
    f["rhocp1"][:na, :] = 1e6
    return f


@pytest.mark.parametrize("nx,ny,per,seed,serp", [(9, 7, 2, 1, False), (41, 23, 4, 2, True), (101, 51, 5, 3, True)])
def test_heat_scatter_bit_exact(gpu_ctx, nx, ny, per, seed, serp):
    print(f"heat scatter {nx}x{ny} per-cell={per} serp={serp}", flush=True)
    g = graded_grid(nx, ny, 2e5, 1e5, seed=seed)
    m = heat_markers(g, per, seed)
    empty = np.isin(m["marker_xn"], [3, 4]) & np.isin(m["marker_yn"], [2, 3])
    m["flags"][empty] = -1
    rng = np.random.default_rng(seed)
    prev = {k: np.asfortranarray(rng.uniform(1, 2, (ny, nx))) for k in ("tk2", "kt1", "rhocp1", "hr1", "ha1")}
    bc = random_bc(g, seed)
    exp = oracle.scatter_heat(g, m, prev, MAT_HR, m["serp"] if serp else None, bc)
    gpu_ctx.set_grid(g)
    dm = {"x": to_device(m["x"]), "y": to_device(m["y"]), "inside_flags": to_device(m["flags"]),
          "TK": to_device(m["TK"]), "kt": to_device(m["kt"]), "rhocp": to_device(m["rhocp"]),
          "ha": to_device(m["ha"]), "matid": to_device(m["matid"])}
    if serp:
        dm["serpentinization_heat_production"] = to_device(m["serp"])
    fields = _dev({"tk1": prev["tk2"], "kt1": prev["kt1"], "rhocp1": prev["rhocp1"], "hr1": prev["hr1"],
                   "ha1": prev["ha1"], "wtnodes": np.zeros((ny, nx), order="F")})
    heat.interpolate_marker_heat_parameters_to_grid(gpu_ctx, dm, {"mat_hr": to_device(MAT_HR)}, fields, _dev(bc))
    gpu_ctx.synchronize()
    for k in ("tk1", "kt1", "rhocp1", "hr1", "ha1", "wtnodes"):
        assert np.array_equal(to_host(fields[k]), exp[k]), k
    assert (exp["wtnodes"] == 0).any()


@pytest.mark.parametrize("ia,ish", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_heat_rhs_bit_exact(gpu_ctx, ia, ish):
    g = graded_grid(33, 21, 1e5, 5e4, seed=4)
    f = grid_fields(g, 7)
    exp = oracle.heat_rhs(g, f, 0.3, 9.81, ia, ish)
    gpu_ctx.set_grid(g)
    RT = heat.calculate_heat_rhs_terms_grid(gpu_ctx, _dev(f), 0.3, 9.81, ia, ish)
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(RT), exp)


@pytest.mark.parametrize("nx,ny,seed", [(7, 5, 1), (37, 23, 2), (129, 65, 3)])
def test_heat_system_rows_bit_exact(gpu_ctx, nx, ny, seed):
    g = graded_grid(nx, ny, 2e5, 1e5, seed=seed)
    f = grid_fields(g, seed)
    bc = random_bc(g, seed + 1)
    rng = np.random.default_rng(seed)
    RT = np.asfortranarray(rng.uniform(0, 1e-6, (ny, nx)))
    T = np.asfortranarray(rng.uniform(200, 1600, (ny, nx)))
    dt = 3.15576e12
    N, Li, Lj, Lv, R = oracle.heat_build(g, bc, RT, f["tk1"], f["rhocp1"], f["kt1"], dt)
    exp_LT = oracle.coo_spmv(N, Li, Lj, Lv, T.ravel(order="F"))
    gpu_ctx.set_grid(g)
    LT, Rg = heat.heat_system_apply(gpu_ctx, to_device(RT), to_device(f["tk1"]), to_device(f["rhocp1"]),
                                    to_device(f["kt1"]), _dev(bc), dt, to_device(T))
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(LT).ravel(order="F"), exp_LT)
    assert np.array_equal(to_host(Rg).ravel(order="F"), R)


SOLVE_CASES = [
    # nx, ny, graded, bc kind, air layer, timestep
    (9, 7, False, "fixed", False, 3.15576e12),      # one level: dense inverse only
    (37, 23, True, "random", False, 3.15576e12),
    (101, 41, False, "fixed", True, 3.15576e13),    # diffusion number ~ 500 in the air layer
    (200, 101, True, "fixed", True, 3.15576e12),    # even node counts: one-cell last coarse intervals
    (257, 129, False, "sine", False, 3.15576e15),
]


@pytest.mark.parametrize("nx,ny,graded,bck,air,dt", SOLVE_CASES)
def test_heat_solve_matches_direct(gpu_ctx, nx, ny, graded, bck, air, dt):
    print(f"heat solve {nx}x{ny} graded={graded} bc={bck} air={air} dt={dt:.2e}", flush=True)
    g = graded_grid(nx, ny, 2e5, 8e4, seed=nx) if graded else uniform_grid(nx, ny, 2e5, 8e4)
    f = grid_fields(g, nx + ny)
    if air:
        f = _air_layer(f, g)
    bc = {"fixed": fixed_top_bottom_bc, "sine": dirichlet_zero_bc}.get(bck, lambda gg: random_bc(gg, 5))(g)
    RT = f["hr1"]
    if bck == "sine":
        X = np.repeat(g.gridx_b[None, :], ny, 0)
        Y = np.repeat(g.gridy_b[:, None], nx, 1)
        f["tk1"] = np.asfortranarray(800 * np.sin(np.pi * X / g.xsize) * np.sin(np.pi * Y / g.ysize))
        RT = np.zeros((ny, nx), order="F")
    exp = oracle.heat_solve(g, bc, RT, f["tk1"], f["rhocp1"], f["kt1"], dt)
    gpu_ctx.set_grid(g)
    tk2, st = heat.discretize_and_solve_conductive_heat_equation(
        gpu_ctx, to_device(RT), to_device(f["tk1"]), to_device(f["rhocp1"]), to_device(f["kt1"]), _dev(bc), dt)
    gpu_ctx.synchronize()
    got = to_host(tk2)
    err = np.abs(got - exp).max() / np.abs(exp).max()
    print(f"  iterations {st['iterations']} relres {st['relres']:.2e} max rel error {err:.2e}", flush=True)
    assert st["converged"] and st["relres"] <= 1e-12
    assert err <= 1e-10
    assert st["iterations"] <= 40
    # the residual by the oracle's own rows
    N, Li, Lj, Lv, R = oracle.heat_build(g, bc, RT, f["tk1"], f["rhocp1"], f["kt1"], dt)
    r = R - oracle.coo_spmv(N, Li, Lj, Lv, got.ravel(order="F"))
    assert np.linalg.norm(r) <= 2e-12 * np.linalg.norm(R)


def test_heat_residuals_bit_exact(gpu_ctx):
    g = graded_grid(45, 31, 2e5, 1e5, seed=11)
    f = grid_fields(g, 12)
    bc = fixed_top_bottom_bc(g)
    rng = np.random.default_rng(3)
    RT = np.asfortranarray(rng.uniform(0, 1e-6, f["tk1"].shape))
    tk2 = np.asfortranarray(f["tk1"] + rng.uniform(-5, 5, f["tk1"].shape))
    dt = 3.15576e12
    exp = oracle.heat_residuals(g, RT, f["rhocp1"], f["tk1"], tk2, f["kt1"], dt)
    gpu_ctx.set_grid(g)
    rest = heat.compute_residuals_temp(gpu_ctx, to_device(RT), to_device(f["tk1"]), to_device(f["rhocp1"]),
                                       to_device(f["kt1"]), _dev(bc), dt, to_device(tk2))
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(rest), exp)


def test_thermal_bc_bit_exact(gpu_ctx):
    g = uniform_grid(23, 17, 1e5, 5e4)
    bc = random_bc(g, 8)
    t = np.asfortranarray(np.random.default_rng(2).uniform(200, 300, (g.ynum, g.xnum)))
    exp = oracle.apply_thermal_bc(g, bc, t)
    gpu_ctx.set_grid(g)
    dt_ = to_device(t)
    heat.apply_thermal_bc(gpu_ctx, _dev(bc), dt_)
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(dt_), exp)


@pytest.mark.parametrize("limit", [60.0, 1e9])
def test_heat_loop_sub_steps(gpu_ctx, limit):
    g = uniform_grid(65, 33, 2e5, 8e4)
    f = _air_layer(grid_fields(g, 21), g)
    bc = fixed_top_bottom_bc(g)
    RT = oracle.heat_rhs(g, f, 0.0, 9.81, 0, 0)
    dt = 3.15576e14
    tk2, rest, log = oracle.heat_loop(g, bc, RT, f["tk1"], f["rhocp1"], f["kt1"], dt, limit)
    gpu_ctx.set_grid(g)
    out, st = heat.run_conductive_heat_solver_loop(gpu_ctx, to_device(RT), to_device(f["tk1"]),
                                                   to_device(f["rhocp1"]), to_device(f["kt1"]), _dev(bc), dt, limit)
    gpu_ctx.synchronize()
    print(f"heat loop limit={limit}: {st['substeps']} sub-steps, {st['solves']} solves, "
          f"{st['iterations']} iterations (oracle {len(log)} sub-steps)", flush=True)
    assert st["converged"] and st["substeps"] == len(log)
    assert st["solves"] == len(log) + sum(1 for s in log if s[2])
    np.testing.assert_allclose(st["timestep_last"], log[-1][0], rtol=1e-9)
    np.testing.assert_allclose(st["timestep_sum"], dt, rtol=1e-12)
    scale = np.abs(tk2).max()
    assert np.abs(to_host(out["tk2"]) - tk2).max() <= 1e-9 * scale
    assert np.array_equal(to_host(out["tk0"]), to_host(out["tk2"]))
    np.testing.assert_array_equal(to_host(out["dtk1"]), to_host(out["tk2"]) - f["tk1"])
    if limit < 1e9:
        assert len(log) >= 3


def _marker_dev(m, keys=("x", "y", "marker_xn", "marker_yn")):
    d = {k: to_device(m[k]) for k in keys}
    d["inside_flags"] = to_device(m["flags"])
    return d


def test_marker_temperature_updates_bit_exact(gpu_ctx):
    g = graded_grid(51, 27, 2e5, 1e5, seed=5)
    m = heat_markers(g, 4, 9)
    f = grid_fields(g, 9)
    gpu_ctx.set_grid(g)
    dm = _marker_dev(m)
    for mode in (0, 1):
        exp = oracle.grid_to_marker_temperature(g, m, f["tk1"], mode, m["TK"])
        TK = to_device(m["TK"])
        heat.grid_to_marker_temperature(gpu_ctx, dm, to_device(f["tk1"]), TK, add=bool(mode))
        gpu_ctx.synchronize()
        assert np.array_equal(to_host(TK), exp), mode
    exp = oracle.sticky_temperature_correction(m, 2, 4, 273.15)
    TK = to_device(m["TK"])
    heat.apply_sticky_temperature_correction(gpu_ctx, dict(dm, matid=to_device(m["matid"])), TK, 2, 4, 273.15)
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(TK), exp)


@pytest.mark.parametrize("subgrid", [0.0, 1.0])
def test_update_marker_temperature_matches_oracle(gpu_ctx, subgrid):
    g = uniform_grid(81, 41, 2e5, 1e5)
    m = heat_markers(g, 4, 13)
    f = _air_layer(grid_fields(g, 13), g)
    bc = fixed_top_bottom_bc(g)
    dt, limit = 3.15576e13, 50.0
    exp = oracle.update_marker_temp_for_grid_and_subgrid_changes(g, bc, m, f, dt, limit, subgrid, 0.0, 9.81, 1, 1)
    gpu_ctx.set_grid(g)
    dm = _marker_dev(m)
    dm["TK"] = to_device(m["TK"])
    r = heat.update_marker_temp_for_grid_and_subgrid_changes(gpu_ctx, dm, _dev(f), _dev(bc), dt, limit, subgrid,
                                                             1, 0.0, 9.81, 1, 1)
    gpu_ctx.synchronize()
    print(f"heat step subgrid={subgrid}: {r['stats']}", flush=True)
    assert np.array_equal(to_host(r["RT1"]), exp["RT1"])
    scale = np.abs(exp["tk2"]).max()
    assert np.abs(to_host(r["tk2"]) - exp["tk2"]).max() <= 1e-9 * scale
    assert np.abs(to_host(r["dtk1"]) - exp["dtk1"]).max() <= 1e-9 * scale
    assert np.abs(to_host(dm["TK"]) - exp["TK"]).max() <= 1e-9 * scale
    if subgrid > 0:
        dtkn = to_host(r["dtkn"])
        assert np.abs(dtkn - exp["dtkn"]).max() <= 1e-12 * max(1.0, np.abs(exp["dtkn"]).max())


def test_subgrid_heat_diffusion_close_to_oracle(gpu_ctx):
    g = graded_grid(61, 33, 2e5, 1e5, seed=3)
    m = heat_markers(g, 5, 17)
    f = grid_fields(g, 17)
    dt, coef = 3.15576e12, 1.0
    t_exp, dT = oracle.subgrid_temperature_markers(g, m, f["tk1"], f["kt1"], f["rhocp1"], dt, coef)
    (dtkn_exp,), wt_exp = oracle.scatter_basic_inclusive(g, m, [dT], [np.zeros((g.ynum, g.xnum), order="F")])
    rng = np.random.default_rng(1)
    dtk1 = np.asfortranarray(rng.uniform(-3, 3, (g.ynum, g.xnum)))
    gpu_ctx.set_grid(g)
    dm = _marker_dev(m)
    TK = to_device(m["TK"])
    d1 = to_device(dtk1)
    wt = to_device(np.zeros((g.ynum, g.xnum), order="F"))
    dtkn = heat.subgrid_heat_diffusion_update(gpu_ctx, dm, _dev(f), TK, dt, coef, d1, wtnodes=wt)
    gpu_ctx.synchronize()
    np.testing.assert_allclose(to_host(TK), t_exp, rtol=1e-14, atol=0)
    np.testing.assert_allclose(to_host(dtkn), dtkn_exp, rtol=1e-12, atol=1e-12 * np.abs(dtkn_exp).max())
    assert np.array_equal(to_host(wt), wt_exp)
    np.testing.assert_allclose(to_host(d1), dtk1 - dtkn_exp, rtol=0, atol=1e-12 * np.abs(dtkn_exp).max())


def test_subgrid_stress_close_to_oracle(gpu_ctx):
    g = graded_grid(57, 31, 2e5, 1e5, seed=6)
    m = heat_markers(g, 5, 19)
    m["mat_mu"] = MAT_MU
    ny, nx = g.ynum, g.xnum
    rng = np.random.default_rng(4)
    f = {"sxy1": np.asfortranarray(rng.uniform(-3e7, 3e7, (ny, nx))),
         "sxx1": np.asfortranarray(rng.uniform(-3e7, 3e7, (ny - 1, nx - 1))),
         "dsxy": np.asfortranarray(rng.uniform(-1e6, 1e6, (ny, nx))),
         "dsxx": np.asfortranarray(rng.uniform(-1e6, 1e6, (ny - 1, nx - 1)))}
    dt, coef = 3.15576e12, 1.0
    exp = oracle.update_subgrid_stress(g, m, f, dt, coef)
    gpu_ctx.set_grid(g)
    dm = _marker_dev(m)
    dm.update(eta=to_device(m["eta"]), matid=to_device(m["matid"]), mat_mu=to_device(MAT_MU))
    df = _dev(f)
    out = {"marker_sxy": to_device(m["sxy"]), "marker_sxx": to_device(m["sxx"])}
    update_subgrid_stress(gpu_ctx, dm, df, dt, coef, out)
    gpu_ctx.synchronize()
    for k, e in (("marker_sxy", exp["sxy"]), ("marker_sxx", exp["sxx"])):
        np.testing.assert_allclose(to_host(out[k]), e, rtol=1e-14, atol=1e-14 * np.abs(e).max())
    for k in ("dsxyn", "dsxxn"):
        np.testing.assert_allclose(to_host(out[k]), exp[k], rtol=1e-12, atol=1e-12 * np.abs(exp[k]).max())
    for k in ("dsxy", "dsxx"):
        np.testing.assert_allclose(to_host(df[k]), exp[k], rtol=1e-12, atol=1e-12 * np.abs(exp[k]).max())
    assert np.array_equal(to_host(out["wtetas"]), exp["wtetas"])
    assert np.array_equal(to_host(out["wtetan"]), exp["wtetan"])
