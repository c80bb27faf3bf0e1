"""GPU parity of the tolerance-mode RK4 (EBG_RK4_TOLERANCE) against the oracle.

Bar (the advection target): marker positions within 1e-10 of the domain
size of the oracle's; velocities and spins within 1e-9 of the largest grid
velocity / spin (a swapped remap, corner or stage gives O(1) errors); markers
outside the domain get zero velocity and keep their position and stress.  On
a grid that is not uniform the mode falls back to the exact kernel, which is
bit-identical to the oracle.  The kernel stages each wavefront's grid patch
in LDS when its markers sit in a few neighbouring cells and reads global
memory otherwise: a shuffled marker order (every wavefront on global reads)
and the kernel with its LDS tiles switched off give results bit-identical to
the spatially ordered markers on tiles.
"""
import numpy as np
import pytest
import torch

import oracle
from earthbox_gpu import workloads
from earthbox_gpu.context import GPUContext
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.grid import graded_grid, uniform_grid

pytestmark = pytest.mark.gpu


@pytest.fixture
def tol_ctx(gpu_ctx):
    gpu_ctx.set_rk4_arithmetic(GPUContext.RK4_TOLERANCE)
    yield gpu_ctx
    gpu_ctx.set_rk4_arithmetic(GPUContext.RK4_EXACT)


def _case(nx, ny, nm, seed, graded=False, ties=True):
    make = (lambda *a: graded_grid(*a, seed=seed + 100)) if graded else uniform_grid
    g = make(nx, ny, 1.5e6 * (nx - 1) / 1000, 6.0e5 * (ny - 1) / 500)
    x, y = workloads.marker_positions(g, nm, nm, seed)
    x[:3] = [-1.0, g.gridx_b[-1] + 1.0, np.nan]
    y[3] = -5.0
    if ties:   # markers exactly on basic nodes, cell centres and the domain edge
        rng = np.random.default_rng(seed + 5)
        k = rng.choice(np.arange(4, x.size), 300, replace=False)
        x[k[:100]] = g.gridx_b[rng.integers(0, nx, 100)]
        y[k[:100]] = g.gridy_b[rng.integers(0, ny, 100)]
        x[k[100:200]] = g.gridx_vy[1:-1][rng.integers(0, nx - 1, 100)]
        y[k[100:200]] = g.gridy_vx[1:-1][rng.integers(0, ny - 1, 100)]
        x[k[200:250]] = g.gridx_b[-1]
        y[k[250:]] = g.gridy_b[-1]
    vx1, vy1 = workloads.random_velocity_fields(g, seed + 1, vmax=1e-9)
    esp = oracle.calculate_spin(g, vx1, vy1)
    dt = 0.5 * g.xstpavg / 1e-9
    return g, x, y, vx1, vy1, esp, dt


def _run(ctx, g, x, y, vx1, vy1, esp, dt, given_cells, stress, outputs=True):
    flags = oracle.inside_flags(x, y, g.limits)
    xn, yn = oracle.cell_indices_basic(x, y, flags, g)
    rng = np.random.default_rng(3)
    sxx, sxy = rng.standard_normal(x.size) * 1e7, rng.standard_normal(x.size) * 1e7
    exp = oracle.advect_markers_and_rotate_stress_tensor(4, dt, g, x, y, xn, yn, flags, vx1, vy1, esp,
                                                         sxx if stress else None, sxy if stress else None)
    ctx.set_grid(g)
    dx, dy = to_device(x), to_device(y)
    n = x.size
    out = [torch.full((n,), 7.0, dtype=torch.float64, device="cuda") for _ in range(3)] if outputs \
        else [None, None, None]
    dsxx, dsxy = (to_device(sxx), to_device(sxy)) if stress else (None, None)
    cells = (to_device(flags), to_device(xn), to_device(yn)) if given_cells else (None, None, None)
    ctx.advect_markers_and_rotate_stress_tensor(4, dt, cells[0], dx, dy, cells[1], cells[2],
                                                to_device(vx1), to_device(vy1), to_device(esp),
                                                *out, dsxx, dsxy)
    ctx.synchronize()
    got = dict(x=to_host(dx), y=to_host(dy))
    if outputs:
        got.update(vx=to_host(out[0]), vy=to_host(out[1]), spin=to_host(out[2]))
    if stress:
        got.update(sxx=to_host(dsxx), sxy=to_host(dsxy))
    return got, dict(x=exp[0], y=exp[1], vx=exp[2], vy=exp[3], spin=exp[4], sxx=exp[5], sxy=exp[6]), flags


def _assert_close(g, got, exp, flags, vx1, vy1, esp, outputs=True, stress=False):
    L = max(g.xsize, g.ysize)
    for k in ("x", "y"):
        assert np.array_equal(np.isnan(got[k]), np.isnan(exp[k])), k
        ok = ~np.isnan(exp[k])
        err = np.abs(got[k][ok] - exp[k][ok]).max()
        assert err <= 1e-10 * L, (k, err)
        out = flags != 1
        assert np.array_equal(got[k][out], exp[k][out], equal_nan=True), k   # outside: unchanged
    if outputs:
        vmax = max(np.abs(vx1).max(), np.abs(vy1).max())
        for k, scale in (("vx", vmax), ("vy", vmax), ("spin", np.abs(esp).max())):
            ok = ~np.isnan(exp[k])
            assert np.abs(got[k][ok] - exp[k][ok]).max() <= 1e-9 * scale, k
            assert np.all(got[k][flags != 1] == 0.0), k
    if stress:
        for k in ("sxx", "sxy"):
            ok = ~np.isnan(exp[k])
            np.testing.assert_allclose(got[k][ok], exp[k][ok], rtol=1e-9, atol=1e-9 * np.abs(exp[k]).max())


@pytest.mark.parametrize("nx,ny,nm,seed", [(11, 7, 3, 1), (101, 101, 4, 2), (203, 61, 5, 3)])
@pytest.mark.parametrize("given_cells", [False, True])
@pytest.mark.parametrize("stress", [False, True])
def test_rk4_tolerance_mode(tol_ctx, nx, ny, nm, seed, given_cells, stress):
    g, x, y, vx1, vy1, esp, dt = _case(nx, ny, nm, seed)
    print(f"rk4 tolerance nx={nx} ny={ny} markers={x.size} given={given_cells} stress={stress}", flush=True)
    tol_ctx.set_grid(g)
    assert tol_ctx.rk4_tolerance_kernel_available()
    got, exp, flags = _run(tol_ctx, g, x, y, vx1, vy1, esp, dt, given_cells, stress)
    _assert_close(g, got, exp, flags, vx1, vy1, esp, stress=stress)


@pytest.mark.parametrize("courant", [0.5, 3.0])
def test_rk4_tolerance_positions_only_and_large_steps(tol_ctx, courant):
    """The headline form (no velocity/spin outputs, no stress: spin is not
    gathered) and steps of several cells, with stage points outside the grid
    (extrapolation from the edge cells)."""
    g, x, y, vx1, vy1, esp, _ = _case(203, 61, 5, 7)
    dt = courant * g.xstpavg / 1e-9
    print(f"rk4 tolerance positions only courant={courant}", flush=True)
    got, exp, flags = _run(tol_ctx, g, x, y, vx1, vy1, esp, dt, False, False, outputs=False)
    _assert_close(g, got, exp, flags, vx1, vy1, esp, outputs=False)


def test_rk4_tolerance_falls_back_to_exact_on_graded_grid(tol_ctx):
    g, x, y, vx1, vy1, esp, dt = _case(97, 53, 4, 5, graded=True)
    tol_ctx.set_grid(g)
    assert not tol_ctx.rk4_tolerance_kernel_available()
    got, exp, flags = _run(tol_ctx, g, x, y, vx1, vy1, esp, dt, False, True)
    for k in ("x", "y", "vx", "vy", "spin"):
        assert np.array_equal(got[k], exp[k], equal_nan=True), k


def test_rk4_tolerance_full_size_sampled(tol_ctx):
    """1001 x 501 nodes, 6 x 6 markers per cell (18M markers, the benchmark
    workload): every 997th marker against the oracle run on those markers."""
    g = uniform_grid(1001, 501, 1.5e6, 6.0e5)
    x, y = workloads.marker_positions(g, 6, 6, 2)
    vx1, vy1 = workloads.random_velocity_fields(g, 12, vmax=1e-9)
    esp = oracle.calculate_spin(g, vx1, vy1)
    dt = 0.5 * g.xstpavg / 1e-9
    tol_ctx.set_grid(g)
    dx, dy = to_device(x), to_device(y)
    tol_ctx.advect_markers_and_rotate_stress_tensor(4, dt, None, dx, dy, None, None, to_device(vx1),
                                                    to_device(vy1), to_device(esp))
    tol_ctx.synchronize()
    gx, gy = to_host(dx), to_host(dy)
    idx = np.arange(0, x.size, 997)
    xs, ys = x[idx].copy(), y[idx].copy()
    flags = oracle.inside_flags(xs, ys, g.limits)
    xn, yn = oracle.cell_indices_basic(xs, ys, flags, g)
    ex, ey, *_ = oracle.advect_markers_and_rotate_stress_tensor(4, dt, g, xs, ys, xn, yn, flags,
                                                                vx1, vy1, esp)
    err = max(np.abs(gx[idx] - ex).max(), np.abs(gy[idx] - ey).max())
    print(f"18M markers: max position error {err:.3e} m", flush=True)
    assert err <= 1e-10 * 1.5e6


@pytest.mark.parametrize("stress,courant", [(False, 0.5), (True, 0.5), (False, 3.0)])
def test_rk4_tolerance_marker_order_independent(tol_ctx, stress, courant):
    """Markers in the reference's order against the same markers shuffled,
    unshuffled: every output bit-identical (a marker's result depends on
    nothing but its own inputs, whichever marker shares its thread)."""
    g, x, y, vx1, vy1, esp, _ = _case(203, 61, 5, 9)
    dt = courant * g.xstpavg / 1e-9
    print(f"rk4 tolerance marker order independence, stress={stress} courant={courant}", flush=True)
    tol_ctx.set_grid(g)
    perm = np.random.default_rng(4).permutation(x.size)
    rng = np.random.default_rng(5)
    sxx, sxy = rng.standard_normal(x.size) * 1e7, rng.standard_normal(x.size) * 1e7
    outs = []
    for order in (np.arange(x.size), perm):
        dx, dy = to_device(x[order]), to_device(y[order])
        dvx, dvy, dsp = (torch.zeros(x.size, dtype=torch.float64, device="cuda") for _ in range(3))
        dsxx, dsxy = (to_device(sxx[order]), to_device(sxy[order])) if stress else (None, None)
        tol_ctx.advect_markers_and_rotate_stress_tensor(4, dt, None, dx, dy, None, None, to_device(vx1),
                                                        to_device(vy1), to_device(esp), dvx, dvy, dsp,
                                                        dsxx, dsxy)
        tol_ctx.synchronize()
        r = {"x": to_host(dx), "y": to_host(dy), "vx": to_host(dvx), "vy": to_host(dvy), "spin": to_host(dsp)}
        if stress:
            r.update(sxx=to_host(dsxx), sxy=to_host(dsxy))
        inv = np.empty_like(order)
        inv[order] = np.arange(x.size)
        outs.append({k: v[inv] for k, v in r.items()})
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k], equal_nan=True), k


@pytest.mark.parametrize("given_cells,stress,outputs,courant",
                         [(False, False, False, 0.5), (True, False, True, 0.5), (False, True, True, 0.5),
                          (True, True, True, 1.5), (False, False, True, 3.0)])
def test_rk4_tolerance_lds_tiles_bit_identical(tol_ctx, given_cells, stress, outputs, courant):
    """The LDS-tiled kernel against the same kernel reading every stage value
    from global memory: all outputs bit-identical, in the reference's marker
    order (tiles used; stage cells more than a cell away read global memory
    per lane at Courant 1.5 and 3), with markers off the grid, on nodes and
    on the domain edge, on grids whose node counts are not multiples of the
    tile shape."""
    g, x, y, vx1, vy1, esp, _ = _case(203, 61, 5, 11)
    dt = courant * g.xstpavg / 1e-9
    print(f"rk4 tiles vs global given={given_cells} stress={stress} courant={courant}", flush=True)
    res = []
    try:
        for tiles in (True, False):
            tol_ctx.set_rk4_tiling(tiles)
            got, _, _ = _run(tol_ctx, g, x, y, vx1, vy1, esp, dt, given_cells, stress, outputs)
            res.append(got)
    finally:
        tol_ctx.set_rk4_tiling(True)
    for k in res[0]:
        assert np.array_equal(res[0][k], res[1][k], equal_nan=True), k
