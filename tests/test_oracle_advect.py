"""Pins of the oracle's grid, indexing and RK4 restatement (CPU only).

The reference ships no function-level golden vectors for this path, so the
oracle is pinned by closed forms, invariants and brute force chosen so that a
plausible mistake (a swapped index, a dropped half step, a '>=' for '>')
fails at least one check.
"""
import math

import numpy as np
import pytest

import oracle
from earthbox_gpu.grid import uniform_grid
from earthbox_gpu import workloads


def test_grid_matches_package_and_closed_forms():
    for (nx, ny, xs, ys) in [(7, 5, 7.0, 3.0), (101, 101, 1.0e6, 1.0e6), (1001, 501, 1.5e6, 6.0e5)]:
        o = oracle.grid2d_uniform(nx, ny, xs, ys)
        g = uniform_grid(nx, ny, xs, ys)
        for k in ("gridx_b", "gridy_b", "xstp_b", "ystp_b", "gridx_vy", "xstp_vy",
                  "gridy_vx", "ystp_vx", "gridx_pr", "gridy_pr", "xstp_pr", "ystp_pr"):
            assert np.array_equal(o[k], getattr(g, k)), k
        # cumulative sum (Initialization.jl:35-42): x_i = x_{i-1} + h exactly
        h = xs / (nx - 1)
        acc = 0.0
        for i in range(nx):
            assert g.gridx_b[i] == acc
            acc += h
        # staggered nodes: first/last ghost half a cell outside, interior midpoints
        assert g.gridx_vy[0] == -g.xstp_b[0] / 2 and g.gridy_vx[0] == -g.ystp_b[0] / 2
        assert np.allclose(g.gridx_vy[1:-1], 0.5 * (g.gridx_b[1:] + g.gridx_b[:-1]), rtol=0, atol=1e-9 * xs)
        assert np.allclose(g.gridx_pr, g.gridx_vy[1:-1], rtol=0, atol=1e-9 * xs)


def _linear_scan(x, g):
    n = g.size
    if math.isnan(x):
        i = n
    else:
        i = 0
        for k in range(n):
            if g[k] <= x:
                i = k + 1
    return min(max(i, 1), n - 1)


def test_index_of_left_node_brute_force():
    g = uniform_grid(9, 7, 8.0e5 / 3.0, 1.0e5).gridx_b  # non-representable spacing
    probes = [-np.inf, -1e300, -0.0, 0.0, np.inf, 1e300, np.nan]
    for v in g:
        probes += [v, np.nextafter(v, -np.inf), np.nextafter(v, np.inf), v + 0.5 * g[1]]
    for x in probes:
        assert oracle.index_of_left_node(x, g) == _linear_scan(x, g), x


def test_spin_closed_form_rigid_rotation():
    g = uniform_grid(21, 13, 2.0e5, 1.2e5)
    om = 3.0e-15
    vx1, vy1 = workloads.solid_body_rotation_fields(g, om)
    esp = oracle.calculate_spin(g, vx1, vy1)
    # GridStrainRate.jl:150-160: esp = 0.5(dvy/dx - dvx/dy) = omega for rigid rotation
    assert np.allclose(esp, om, rtol=1e-9, atol=0)


def _setup(nx=31, ny=17, nm=3, seed=5):
    g = uniform_grid(nx, ny, 3.0e5, 1.6e5)
    x, y = workloads.marker_positions(g, nm, nm, seed)
    flags = oracle.inside_flags(x, y, g.limits)
    xn, yn = oracle.cell_indices_basic(x, y, flags, g)
    return g, x, y, flags, xn, yn


def test_rk4_uniform_velocity_is_exact():
    g, x, y, flags, xn, yn = _setup()
    vx1 = np.asfortranarray(np.full((g.ynum + 1, g.xnum), 2.5e-10))
    vy1 = np.asfortranarray(np.full((g.ynum, g.xnum + 1), -1.25e-10))
    esp = oracle.calculate_spin(g, vx1, vy1)
    dt = 1.0e12
    x2, y2, mvx, mvy, msp, _, _ = oracle.advect_markers_and_rotate_stress_tensor(
        4, dt, g, x, y, xn, yn, flags, vx1, vy1, esp)
    # partition of unity: bilinear weights sum to 1 up to rounding
    assert np.allclose(mvx, 2.5e-10, rtol=1e-14) and np.allclose(mvy, -1.25e-10, rtol=1e-14)
    assert np.all(msp == 0.0)
    assert np.allclose(x2, x + dt * 2.5e-10, rtol=0, atol=1e-9) and np.allclose(y2, y - dt * 1.25e-10, rtol=0, atol=1e-9)


def test_rk4_linear_shear_flow_matches_flow_map():
    """vx = a*y (linear in y, sampled on the vx grid), vy = b*x: bilinear
    interpolation is exact for these fields, so the RK4 velocity equals the
    analytic RK4 of the ODE; the ODE x' = a y, y' = b x has the closed-form
    flow map used below (error O(dt^5))."""
    g, x, y, flags, xn, yn = _setup(41, 41, 2, 11)
    a, b = 1.0e-15, -0.5e-15
    yc = 0.5 * (g.gridy_b[-1] + g.gridy_b[0])
    xc = 0.5 * (g.gridx_b[-1] + g.gridx_b[0])
    vx1 = np.asfortranarray(a * (g.gridy_vx[:, None] - yc) * np.ones((1, g.xnum)))
    vy1 = np.asfortranarray(b * (g.gridx_vy[None, :] - xc) * np.ones((g.ynum, 1)))
    esp = oracle.calculate_spin(g, vx1, vy1)
    dt = 2.0e13  # |a|*dt = 0.02
    inner = (x > g.gridx_b[1]) & (x < g.gridx_b[-2]) & (y > g.gridy_b[1]) & (y < g.gridy_b[-2])
    x2, y2, mvx, mvy, msp, _, _ = oracle.advect_markers_and_rotate_stress_tensor(
        4, dt, g, x, y, xn, yn, flags, vx1, vy1, esp)
    # exact flow of X' = M X with M = [[0, a], [b, 0]], a*b < 0: rotation-like
    w = math.sqrt(-a * b)
    X0, Y0 = x - xc, y - yc
    Xe = X0 * math.cos(w * dt) + (a / w) * Y0 * math.sin(w * dt)
    Ye = Y0 * math.cos(w * dt) + (b / w) * X0 * math.sin(w * dt)
    err = np.hypot(x2 - xc - Xe, y2 - yc - Ye)[inner]
    L = g.xsize
    assert err.max() < 1e-8 * L  # RK4 local error ~ (w dt)^5/120 * r
    # spin of this field is 0.5*(b - a) everywhere
    assert np.allclose(msp[inner], 0.5 * (b - a), rtol=1e-9)


def test_rk4_order1_uses_stage_one_only():
    g, x, y, flags, xn, yn = _setup()
    vx1, vy1 = workloads.random_velocity_fields(g, 3)
    esp = oracle.calculate_spin(g, vx1, vy1)
    v4 = oracle.rk4_velocity_and_spin(4, 1e12, g, x, y, xn, yn, flags, vx1, vy1, esp)
    v1 = oracle.rk4_velocity_and_spin(1, 1e12, g, x, y, xn, yn, flags, vx1, vy1, esp)
    v1b = oracle.rk4_velocity_and_spin(1, 0.0, g, x, y, xn, yn, flags, vx1, vy1, esp)
    assert np.array_equal(v1[0], v1b[0])          # stage 1 does not depend on dt
    assert not np.array_equal(v1[0], v4[0])       # RK4 combines later stages


def test_outside_markers_zero_and_unmoved():
    g, x, y, flags, xn, yn = _setup()
    x = x.copy()
    x[:5] = -1.0
    x[5] = np.nan
    flags = oracle.inside_flags(x, y, g.limits)
    assert np.all(flags[:6] == -1)
    vx1, vy1 = workloads.random_velocity_fields(g, 4)
    esp = oracle.calculate_spin(g, vx1, vy1)
    sxx = np.linspace(1, 2, x.size)
    sxy = np.linspace(-1, 1, x.size)
    x2, y2, mvx, mvy, msp, sxx2, sxy2 = oracle.advect_markers_and_rotate_stress_tensor(
        4, 1e12, g, x, y, xn, yn, flags, vx1, vy1, esp, sxx, sxy)
    assert np.all(mvx[:6] == 0) and np.all(msp[:6] == 0)
    assert np.array_equal(x2[:5], x[:5]) and np.isnan(x2[5])
    assert np.array_equal(sxx2[:6], sxx[:6])


def test_stress_rotation_preserves_invariant():
    """Rotation of a 2-D deviatoric tensor preserves sxx^2 + sxy^2
    (MarkersStressRotation.jl:42-58)."""
    g, x, y, flags, xn, yn = _setup()
    vx1, vy1 = workloads.solid_body_rotation_fields(g, 1e-14)
    esp = oracle.calculate_spin(g, vx1, vy1)
    rng = np.random.default_rng(0)
    sxx = rng.standard_normal(x.size) * 1e7
    sxy = rng.standard_normal(x.size) * 1e7
    *_, sxx2, sxy2 = oracle.advect_markers_and_rotate_stress_tensor(
        4, 1e12, g, x, y, xn, yn, flags, vx1, vy1, esp, sxx, sxy)
    assert np.allclose(sxx2 ** 2 + sxy2 ** 2, sxx ** 2 + sxy ** 2, rtol=1e-12)
    # rotation angle is spin*dt = 1e-2 rad: sxy changes
    assert not np.allclose(sxy2, sxy)
