"""Seeded markers and grid fields for the grid -> marker gathers after the
Stokes solve (marker strain rate, strain-rate ratio, marker pressure)."""
import numpy as np

import oracle
from earthbox_gpu.grid import graded_grid

PRESSURE = ("exx", "pr1", "sxx2", "dsxx")
BASIC = ("exy", "sxy2", "dsxy")


def node_coords(g, kind):
    """(y, x) coordinates of the basic nodes or the pressure nodes (cell
    centres), as 2-D arrays of the grid arrays' shapes."""
    if kind == "basic":
        return np.meshgrid(g.gridy_b, g.gridx_b, indexing="ij")
    return np.meshgrid(g.gridy_vx[1:-1], g.gridx_vy[1:-1], indexing="ij")


def gather_case(nx, ny, n, seed, ties=True):
    """Graded grid, n markers uniform over the box (two outside, some exactly on
    basic nodes, cell centres and cell edges when ties), their stored cells
    and inside flags, random grid fields and per-marker viscosity and material."""
    g = graded_grid(nx, ny, 2.0e5, 1.0e5, seed)
    rng = np.random.default_rng(seed + 300)
    x = rng.uniform(g.gridx_b[0], g.gridx_b[-1], n)
    y = rng.uniform(g.gridy_b[0], g.gridy_b[-1], n)
    x[:2] = [g.gridx_b[0] - 10.0, g.gridx_b[-1] + 1.0]
    if ties and n >= 64:
        k = rng.choice(np.arange(2, n), 60, replace=False)
        jx = rng.integers(0, nx, 60)
        jy = rng.integers(0, ny, 60)
        x[k[:20]], y[k[:20]] = g.gridx_b[jx[:20]], g.gridy_b[jy[:20]]
        x[k[20:40]] = g.gridx_vy[1:-1][np.minimum(jx[20:40], nx - 2)]
        y[k[20:40]] = g.gridy_vx[1:-1][np.minimum(jy[20:40], ny - 2)]
        x[k[40:]] = g.gridx_b[-1]
        y[k[40:]] = g.gridy_b[-1] if seed % 2 else g.gridy_b[0]
    inside = oracle.inside_flags(x, y, g.limits)
    xn, yn = oracle.cell_indices_basic(x, y, inside, g)
    m = {"x": x, "y": y, "xn": xn, "yn": yn, "inside": inside,
         "eta": 10 ** rng.uniform(19, 24, n), "matid": rng.integers(1, 4, n).astype(np.int16),
         "mat_mu": np.array([3e10, 6e10, 1e10])}
    f = {}
    for k in PRESSURE:
        f[k] = np.asfortranarray(rng.uniform(-1e-14, 1e-14, (ny - 1, nx - 1)) if k == "exx"
                                 else rng.uniform(-1e7, 1e7, (ny - 1, nx - 1)))
    for k in BASIC:
        f[k] = np.asfortranarray(rng.uniform(-1e-14, 1e-14, (ny, nx)) if k == "exy"
                                 else rng.uniform(-1e7, 1e7, (ny, nx)))
    return g, m, f
