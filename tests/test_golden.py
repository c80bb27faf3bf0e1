"""Golden fixtures (tests/golden/, written by make_golden.py): the oracle
still reproduces them bit for bit (CPU), and the GPU path matches them
(bit-exact for RK4 and the operator; forward error <= 1e-8 for the solve)."""
import ast
import pathlib

import numpy as np
import pytest

import oracle
from earthbox_gpu import workloads
from earthbox_gpu.grid import uniform_grid
from stokes_cases import random_case

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def _load(name):
    return np.load(GOLD / name, allow_pickle=False)


def test_oracle_reproduces_rk4_fixture():
    f = _load("rk4_9x7.npz")
    g = uniform_grid(int(f["xnum"]), int(f["ynum"]), float(f["xsize"]), float(f["ysize"]))
    x, y = workloads.marker_positions(g, 3, 3, 11)
    assert np.array_equal(x, f["x"]) and np.array_equal(y, f["y"])
    ex, ey, evx, evy, esp_m, _, _ = oracle.advect_markers_and_rotate_stress_tensor(
        4, float(f["dt"]), g, f["x"], f["y"], f["xn"], f["yn"], f["flags"], f["vx1"], f["vy1"], f["esp"])
    for got, key in ((ex, "out_x"), (ey, "out_y"), (evx, "out_vx"), (evy, "out_vy"), (esp_m, "out_spin")):
        assert np.array_equal(got, f[key]), key


@pytest.mark.parametrize("name", ["stokes_9x7.npz", "stokes_ext_17x11.npz"])
def test_oracle_reproduces_stokes_fixture(name):
    f = _load(name)
    pr = random_case(**ast.literal_eval(str(f["case"])))
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    for got, key in ((Li, "Li"), (Lj, "Lj"), (Lv, "Lv"), (R, "R")):
        assert np.array_equal(got, f[key]), key
    S = oracle.direct_solve(pr.N, Li, Lj, Lv, R)
    assert np.linalg.norm(S - f["S"]) <= 1e-12 * np.linalg.norm(f["S"])


@pytest.mark.gpu
def test_gpu_matches_rk4_fixture(gpu_ctx):
    from earthbox_gpu.device import to_device, to_host
    f = _load("rk4_9x7.npz")
    g = uniform_grid(int(f["xnum"]), int(f["ynum"]), float(f["xsize"]), float(f["ysize"]))
    gpu_ctx.set_grid(g)
    import torch
    n = f["x"].size
    dx, dy = to_device(f["x"]), to_device(f["y"])
    out = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3)]
    gpu_ctx.advect_markers_and_rotate_stress_tensor(
        4, float(f["dt"]), to_device(f["flags"]), dx, dy, to_device(f["xn"]), to_device(f["yn"]),
        to_device(f["vx1"]), to_device(f["vy1"]), to_device(f["esp"]), *out)
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(dx), f["out_x"]) and np.array_equal(to_host(dy), f["out_y"])
    assert np.array_equal(to_host(out[0]), f["out_vx"]) and np.array_equal(to_host(out[1]), f["out_vy"])
    assert np.array_equal(to_host(out[2]), f["out_spin"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["stokes_9x7.npz", "stokes_ext_17x11.npz"])
def test_gpu_matches_stokes_fixture(gpu_ctx, name):
    from earthbox_gpu.device import to_device, to_host
    from earthbox_gpu.stokes import StokesSolverGPU, bc_mirror_factors
    f = _load(name)
    pr = random_case(**ast.literal_eval(str(f["case"])))
    sol = StokesSolverGPU(gpu_ctx)
    sol.set_system(pr)
    S = f["S"]
    LS = to_host(sol.apply(to_device(S)))
    assert np.array_equal(LS, oracle.coo_spmv(pr.N, f["Li"], f["Lj"], f["Lv"], S))
    assert np.array_equal(to_host(sol.rhs()), f["R"])
    Sg, st = sol.solve_system(precision=1, rtol=1e-12, max_iterations=3000, bc_d=bc_mirror_factors(pr))
    assert st["converged"], st
    assert np.linalg.norm(to_host(Sg) - S) <= 1e-8 * np.linalg.norm(S)
