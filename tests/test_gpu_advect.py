"""GPU parity of the marker kernels against the oracle (through the C ABI).

Bar: bit-exact for cell indices, velocities, spins and positions (same FP64
evaluation order, no contraction); stress rotation within 4 ulp-scale relative
error (device sin/cos vs libm).
"""
import numpy as np
import pytest
import torch

import oracle
from earthbox_gpu import workloads
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.grid import graded_grid, uniform_grid

pytestmark = pytest.mark.gpu


def _case(nx, ny, nm, seed, outside=True, graded=False):
    make = (lambda *a: graded_grid(*a, seed=seed + 100)) if graded else uniform_grid
    g = make(nx, ny, 1.5e6 * (nx - 1) / 1000, 6.0e5 * (ny - 1) / 500)
    x, y = workloads.marker_positions(g, nm, nm, seed)
    if outside:
        x[:3] = [-1.0, g.gridx_b[-1] + 1.0, np.nan]
        y[3] = -5.0
    vx1, vy1 = workloads.random_velocity_fields(g, seed + 1, vmax=1e-9)
    esp = oracle.calculate_spin(g, vx1, vy1)
    # Courant-like step: half a cell at max speed
    dt = 0.5 * g.xstpavg / 1e-9
    return g, x, y, vx1, vy1, esp, dt


@pytest.mark.parametrize("nx,ny,nm,seed,graded", [(11, 7, 3, 1, False), (101, 101, 4, 2, False),
                                                  (203, 61, 5, 3, False), (11, 7, 3, 4, True),
                                                  (97, 53, 4, 5, True)])
@pytest.mark.parametrize("mode", ["reference", "minimal"])
def test_rk4_advect_bit_exact(gpu_ctx, nx, ny, nm, seed, graded, mode):
    g, x, y, vx1, vy1, esp, dt = _case(nx, ny, nm, seed, graded=graded)
    print(f"rk4 case nx={nx} ny={ny} markers={x.size} graded={graded} mode={mode}", flush=True)
    _check_rk4(gpu_ctx, g, x, y, vx1, vy1, esp, dt, seed, mode)


@pytest.mark.parametrize("order", ["reference", "shuffled"])
@pytest.mark.parametrize("courant", [0.5, 3.0])
def test_rk4_shuffled_and_large_steps_bit_exact(gpu_ctx, order, courant):
    """Markers in random order and steps of several cells per stage (stage
    points far from the stage-1 cell, some outside the grid) match the oracle
    bit for bit."""
    g, x, y, vx1, vy1, esp, _ = _case(203, 61, 5, 7)
    dt = courant * g.xstpavg / 1e-9
    if order == "shuffled":
        p = np.random.default_rng(8).permutation(x.size)
        x, y = x[p].copy(), y[p].copy()
    print(f"rk4 window fallbacks order={order} courant={courant}", flush=True)
    _check_rk4(gpu_ctx, g, x, y, vx1, vy1, esp, dt, 7, "minimal")


def _check_rk4(gpu_ctx, g, x, y, vx1, vy1, esp, dt, seed, mode):
    flags = oracle.inside_flags(x, y, g.limits)
    xn, yn = oracle.cell_indices_basic(x, y, flags, g)
    rng = np.random.default_rng(seed)
    sxx, sxy = rng.standard_normal(x.size) * 1e7, rng.standard_normal(x.size) * 1e7
    ex, ey, evx, evy, esp_m, esxx, esxy = oracle.advect_markers_and_rotate_stress_tensor(
        4, dt, g, x, y, xn, yn, flags, vx1, vy1, esp, sxx, sxy)

    gpu_ctx.set_grid(g)
    dx, dy = to_device(x), to_device(y)
    dvx1, dvy1, desp = to_device(vx1), to_device(vy1), to_device(esp)
    n = x.size
    out = [torch.full((n,), 7.0, dtype=torch.float64, device="cuda") for _ in range(3)]
    dsxx, dsxy = to_device(sxx), to_device(sxy)
    if mode == "reference":
        gpu_ctx.advect_markers_and_rotate_stress_tensor(
            4, dt, to_device(flags), dx, dy, to_device(xn), to_device(yn), dvx1, dvy1, desp,
            *out, dsxx, dsxy)
    else:  # recompute flags and stage-1 indices on the device
        gpu_ctx.advect_markers_and_rotate_stress_tensor(
            4, dt, None, dx, dy, None, None, dvx1, dvy1, desp, *out, dsxx, dsxy)
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(out[0]), evx)
    assert np.array_equal(to_host(out[1]), evy)
    assert np.array_equal(to_host(out[2]), esp_m)
    assert np.array_equal(to_host(dx), ex, equal_nan=True)
    assert np.array_equal(to_host(dy), ey, equal_nan=True)
    np.testing.assert_allclose(to_host(dsxx), esxx, rtol=1e-14, atol=1e-7)
    np.testing.assert_allclose(to_host(dsxy), esxy, rtol=1e-14, atol=1e-7)


@pytest.mark.parametrize("graded", [False, True])
def test_rk_order1_velocity_bit_exact(gpu_ctx, graded):
    """runge_kutta_order_max != 4: the stage-1 velocity and spin only."""
    g, x, y, vx1, vy1, esp, dt = _case(61, 41, 4, 7, graded=graded)
    print(f"rk order-1 case graded={graded} markers={x.size}", flush=True)
    flags = oracle.inside_flags(x, y, g.limits)
    xn, yn = oracle.cell_indices_basic(x, y, flags, g)
    evx, evy, esp_m = oracle.rk4_velocity_and_spin(1, dt, g, x, y, xn, yn, flags, vx1, vy1, esp)
    gpu_ctx.set_grid(g)
    n = x.size
    out = [torch.full((n,), 7.0, dtype=torch.float64, device="cuda") for _ in range(3)]
    gpu_ctx.update_marker_velocity_and_spin_using_runge_kutta(
        1, dt, to_device(flags), to_device(x), to_device(y), to_device(xn), to_device(yn),
        to_device(vx1), to_device(vy1), to_device(esp), *out)
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(out[0]), evx)
    assert np.array_equal(to_host(out[1]), evy)
    assert np.array_equal(to_host(out[2]), esp_m)


@pytest.mark.parametrize("graded", [False, True])
def test_cell_indices_and_flags_bit_exact(gpu_ctx, graded):
    g = (graded_grid(41, 23, 4.0e5 / 3.0, 2.2e5, seed=8) if graded
         else uniform_grid(41, 23, 4.0e5 / 3.0, 2.2e5))
    # every node coordinate, its neighbours one ulp away, and outside probes
    xs = np.concatenate([g.gridx_b, np.nextafter(g.gridx_b, -np.inf), np.nextafter(g.gridx_b, np.inf),
                         [-0.0, -1e-300, np.inf, -np.inf, g.gridx_b[-1] * 2]])
    ys = np.concatenate([g.gridy_b, np.nextafter(g.gridy_b, -np.inf), np.nextafter(g.gridy_b, np.inf)])
    X = np.repeat(xs, ys.size)
    Y = np.tile(ys, xs.size)
    ones = np.ones(X.size, dtype=np.int8)
    exn, eyn = oracle.cell_indices_basic(X, Y, ones, g)
    eflags = oracle.inside_flags(X, Y, g.limits)
    gpu_ctx.set_grid(g)
    dX, dY = to_device(X), to_device(Y)
    xn = torch.zeros(X.size, dtype=torch.int32, device="cuda")
    yn = torch.zeros(X.size, dtype=torch.int32, device="cuda")
    gpu_ctx.calc_upper_left_basic_grid_cell_indices_for_markers(dX, dY, None, xn, yn)
    f = gpu_ctx.get_marker_inside_flags(dX, dY)
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(xn), exn)
    assert np.array_equal(to_host(yn), eyn)
    assert np.array_equal(to_host(f), eflags)


def test_spin_bit_exact(gpu_ctx):
    g = uniform_grid(57, 33, 5.6e5, 3.2e5)
    vx1, vy1 = workloads.random_velocity_fields(g, 9)
    e = oracle.calculate_spin(g, vx1, vy1)
    gpu_ctx.set_grid(g)
    out = to_device(np.zeros_like(e, order="F"))
    gpu_ctx.calculate_spin(to_device(vx1), to_device(vy1), out)
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(out), e)


def test_rk4_full_size_sampled_and_deterministic(gpu_ctx):
    """1001 x 501 nodes, 6 x 6 markers per cell (18M markers): the GPU result
    of every 997th marker equals the oracle run on just those markers (RK4 is
    per-marker independent), and two runs are bit-identical."""
    g = uniform_grid(1001, 501, 1.5e6, 6.0e5)
    x, y = workloads.marker_positions(g, 6, 6, 2)
    vx1, vy1 = workloads.random_velocity_fields(g, 12, vmax=1e-9)
    esp = oracle.calculate_spin(g, vx1, vy1)
    dt = 0.5 * g.xstpavg / 1e-9
    gpu_ctx.set_grid(g)
    res = []
    for _ in range(2):
        dx, dy = to_device(x), to_device(y)
        gpu_ctx.advect_markers_and_rotate_stress_tensor(
            4, dt, None, dx, dy, None, None, to_device(vx1), to_device(vy1), to_device(esp))
        gpu_ctx.synchronize()
        res.append((to_host(dx), to_host(dy)))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    idx = np.arange(0, x.size, 997)
    xs, ys = x[idx].copy(), y[idx].copy()
    flags = oracle.inside_flags(xs, ys, g.limits)
    xn, yn = oracle.cell_indices_basic(xs, ys, flags, g)
    ex, ey, *_ = oracle.advect_markers_and_rotate_stress_tensor(
        4, dt, g, xs, ys, xn, yn, flags, vx1, vy1, esp)
    assert np.array_equal(res[0][0][idx], ex) and np.array_equal(res[0][1][idx], ey)
