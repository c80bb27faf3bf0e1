"""Known-answer pins of the oracle's scatter weight rules (CPU only).

The reference's three scatter rules (interpolation/weights/WeightFuncs.jl):
  inclusive (:67-85)   all four corners of the marker's cell, weights
                       (1-dx)(1-dy), (1-dx)dy, dx(1-dy), dx dy;
  exclusive (:122-150) the same weights, but a corner receives only markers
                       within half a cell of it in both directions
                       (ties at 0.5 included, so a cell-centre marker feeds
                       all four corners);
  central (:178-190)   the cell's pressure node, weight
                       (1-|1/2-dx|)(1-|1/2-dy|).

The checks are written node-centrically (tent functions around each node),
not as the oracle's corner loop, so a swapped corner, a wrong corner weight or
a wrong tie condition fails them:
  1. one marker per quadrant of a cell with exactly representable (dx, dy):
     which nodes receive weight, and the exact weight at each;
  2. regular markers and a quadratic field: the average at an interior node is
     the field value plus a closed-form second moment of the rule's weights
     (the first moments and the cross moment vanish by symmetry).
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from earthbox_gpu import workloads
from earthbox_gpu.grid import uniform_grid
from scatter_cases import previous_fields

H = 1024.0   # power-of-two steps: every distance and weight below is exact


def _grid():
    return uniform_grid(6, 5, 5 * H, 4 * H)


def _one_marker(g, x, y):
    m = {"x": np.array([x]), "y": np.array([y]), "eta": np.array([3.0]), "rho": np.array([2.0]),
         "mat_mu": np.array([1.0]), "matid": np.array([1], np.int16)}
    return oracle.scatter_stokes(g, m, previous_fields(g, 1))


def _tent(d):
    return 1.0 - abs(d)


QUADRANTS = [
    # (dx, dy) inside basic cell (iy, ix) = (2, 3) (1-based)
    (0.25, 0.125),   # upper-left quarter
    (0.25, 0.875),   # lower-left
    (0.75, 0.125),   # upper-right
    (0.75, 0.625),   # lower-right
    (0.5, 0.25),     # tie in x: upper-left and upper-right
    (0.25, 0.5),     # tie in y: upper-left and lower-left
    (0.5, 0.5),      # cell centre: all four corners
    (0.0, 0.375),    # on the left cell edge
]


@pytest.mark.parametrize("dx,dy", QUADRANTS)
def test_single_marker_rules(dx, dy):
    g = _grid()
    ix, iy = 3, 2
    x = g.gridx_b[ix - 1] + dx * H
    y = g.gridy_b[iy - 1] + dy * H
    out = _one_marker(g, x, y)
    excl = np.zeros((g.ynum, g.xnum))
    incl = np.zeros((g.ynum, g.xnum))
    for r in (iy, iy + 1):
        for c in (ix, ix + 1):
            ox = (x - g.gridx_b[c - 1]) / H    # node-centric offsets in cells
            oy = (y - g.gridy_b[r - 1]) / H
            w = _tent(ox) * _tent(oy)
            incl[r - 1, c - 1] = w
            if abs(ox) <= 0.5 and abs(oy) <= 0.5:
                excl[r - 1, c - 1] = w
    assert np.array_equal(out["wtnodes"], incl)
    assert np.array_equal(out["wtetas"], excl)
    # etas1 is the marker's viscosity exactly where the exclusive rule fed a node
    fed = excl != 0
    assert np.all(out["etas1"][fed] == 3.0)
    # central rule: only pressure node (iy, ix), tent around the cell centre
    cen = np.zeros((g.ynum - 1, g.xnum - 1))
    cen[iy - 1, ix - 1] = _tent(dx - 0.5) * _tent(dy - 0.5)
    assert np.array_equal(out["wtetan"], cen)
    assert out["etan1"][iy - 1, ix - 1] == 3.0


def test_single_marker_vy_grid():
    """The vy grid is offset by half a cell in x, so a marker at dx = 0.25 of
    basic cell ix sits at 0.75 of vy cell ix (the vy node left of it is the
    centre of basic cell ix-1)."""
    g = _grid()
    ix, iy = 3, 2
    x = g.gridx_b[ix - 1] + 0.25 * H
    y = g.gridy_b[iy - 1] + 0.375 * H
    out = _one_marker(g, x, y)
    exp = np.zeros((g.ynum, g.xnum + 1))
    for r in (iy, iy + 1):
        for c in range(1, g.xnum + 2):
            ox = (x - g.gridx_vy[c - 1]) / H
            oy = (y - g.gridy_b[r - 1]) / H
            if abs(ox) < 1.0:
                exp[r - 1, c - 1] = _tent(ox) * _tent(oy)
    assert np.count_nonzero(exp) == 4
    assert np.array_equal(out["wtnodes_vy"], exp)
    assert np.all(out["rho1_vy"][exp != 0] == 2.0)


def _second_moment(nm, half_cells):
    """Weighted mean of the squared offset (in cells) of regular markers
    (nm per cell, at (k+1/2)/nm) around a node, for tent weights 1-|o|,
    over offsets |o| < half_cells."""
    offs = [Fraction(2 * k + 1, 2 * nm) for k in range(int(nm * half_cells))]
    num = sum((1 - o) * o * o for o in offs)
    den = sum((1 - o) for o in offs)
    return float(num / den)


def test_second_moments_closed_form():
    # the sums written out for 4 markers per cell
    assert _second_moment(4, 1) == 176 / 1024      # offsets 1/8..7/8, weights 7/8..1/8
    assert _second_moment(4, 0.5) == 104 / 1536    # offsets 1/8, 3/8


@pytest.mark.parametrize("nm", [2, 4, 6])
def test_quadratic_field_closed_form(nm):
    nx, ny = 11, 9
    g = uniform_grid(nx, ny, (nx - 1) * H, (ny - 1) * H)
    x, y = workloads.marker_positions(g, nm, nm, 0, jitter=False)
    A, B, C, D = 3.0, -2.0, 0.5, 7.0
    xc, yc = 4.3 * H, 2.9 * H

    def f(px, py):
        return D + A * ((px - xc) / H) ** 2 + B * ((py - yc) / H) ** 2 + C * (px / H) * (py / H)

    val = f(x, y)
    m = {"x": x, "y": y, "rho": val, "eta": val, "mat_mu": np.array([1.0]),
         "matid": np.ones(x.size, np.int16)}
    out = oracle.scatter_stokes(g, m, previous_fields(g, 2))
    e_in, e_ex = _second_moment(nm, 1), _second_moment(nm, 0.5)
    X, Y = np.meshgrid(g.gridx_b, g.gridy_b)
    # inclusive (rho1) and exclusive (etas1) rules at interior basic nodes
    inner = (slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(out["rho1"][inner], (f(X, Y) + A * e_in + B * e_in)[inner], rtol=1e-12)
    np.testing.assert_allclose(out["etas1"][inner], (f(X, Y) + A * e_ex + B * e_ex)[inner], rtol=1e-12)
    # central rule (etan1) at every pressure node: the cell's own markers
    XP, YP = np.meshgrid(g.gridx_pr, g.gridy_pr)
    np.testing.assert_allclose(out["etan1"], f(XP, YP) + A * e_ex + B * e_ex, rtol=1e-12)
    # inclusive rule on the vy grid at vy nodes with a full marker neighbourhood
    XV, YV = np.meshgrid(g.gridx_vy, g.gridy_b)
    innv = (slice(1, -1), slice(2, -2))
    np.testing.assert_allclose(out["rho1_vy"][innv], (f(XV, YV) + A * e_in + B * e_in)[innv], rtol=1e-12)


def test_quadratic_moment_catches_swapped_corner_weights():
    """A corner rule with the weights of the opposite corner (dx for the left
    node) would give a different second moment: the closed form tells the
    two apart."""
    offs = [Fraction(2 * k + 1, 8) for k in range(4)]
    swapped = float(sum(o * o * o for o in offs) / sum(offs))
    assert abs(swapped - _second_moment(4, 1)) > 0.1
