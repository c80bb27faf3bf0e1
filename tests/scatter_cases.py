"""Seeded marker sets with every Stokes property for scatter tests."""
import numpy as np

from earthbox_gpu import workloads
from earthbox_gpu.grid import uniform_grid

FIELDS = ("rho1", "rho1_vy", "etas1", "etan1", "eta_flow", "mus1", "mun1", "sxy1", "sxx1",
          "plastics", "plasticn", "cohesion_grid", "fric_degrees_grid", "dilatation_grid",
          "extractable_meltfrac_grid")


def marker_set(nx, ny, nm, seed, jitter=True, outside=True):
    g = uniform_grid(nx, ny, 2.0e5 * (nx - 1) / 20, 1.0e5 * (ny - 1) / 10)
    x, y = workloads.marker_positions(g, nm, nm, seed, jitter=jitter)
    rng = np.random.default_rng(seed + 100)
    n = x.size
    if outside:
        x[:2] = [-5.0, g.gridx_b[-1] + 3.0]
    mat_mu = np.array([3e10, 6e10, 1e10])
    m = {
        "x": x, "y": y,
        "rho": 3000 + 300 * rng.random(n),
        "cohesion": 1e7 * rng.random(n),
        "fric": 0.6 * rng.random(n),
        "eta_flow": 10 ** rng.uniform(19, 23, n),
        "eta": 10 ** rng.uniform(19, 23, n),
        "sxy": 1e6 * rng.standard_normal(n),
        "sxx": 1e6 * rng.standard_normal(n),
        "meltfrac": rng.random(n) * 0.1,
        "pfailure": (rng.random(n) < 0.3).astype(np.float32),
        "dilatation": (5 * rng.random(n)).astype(np.float32),
        "matid": rng.integers(1, 4, n).astype(np.int16),
        "mat_mu": mat_mu,
    }
    return g, m


def previous_fields(g, seed):
    """Previous grid values (the backup where no marker weight lands)."""
    from earthbox_gpu.scatter import grid_field_shape
    rng = np.random.default_rng(seed + 7)
    return {k: np.asfortranarray(rng.random(grid_field_shape(k, g.xnum, g.ynum)) + 1.0)
            for k in FIELDS}
