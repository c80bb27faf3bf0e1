"""Shared inputs of the heat-path tests: thermal boundary conditions, grid
fields and marker states on a 2-D staggered grid (column-major arrays)."""
import numpy as np

import oracle
from earthbox_gpu import workloads

SEC_PER_MYR = 3.15576e13


def fixed_top_bottom_bc(g, t_top=273.0, t_bottom=1573.0):
    """Fixed temperature at the top and bottom, zero heat flux on the sides:
    the usual lithosphere setting (TemperatureBC.jl, ZeroHeatFluxBC.jl)."""
    nx, ny = g.xnum, g.ynum
    btopt = np.zeros((nx, 2), order="F")
    bbottomt = np.zeros((nx, 2), order="F")
    btopt[:, 0] = t_top
    bbottomt[:, 0] = t_bottom
    bleftt = np.zeros((ny, 2), order="F")
    brightt = np.zeros((ny, 2), order="F")
    bleftt[:, 1] = 1.0
    brightt[:, 1] = 1.0
    return {"btopt": btopt, "bbottomt": bbottomt, "bleftt": bleftt, "brightt": brightt}


def random_bc(g, seed):
    """Every coefficient random: value and neighbour factor on all four sides."""
    rng = np.random.default_rng(seed)
    nx, ny = g.xnum, g.ynum
    mk = lambda n: np.asfortranarray(np.column_stack([rng.uniform(200, 1600, n), rng.uniform(-1, 1, n)]))
    return {"btopt": mk(nx), "bbottomt": mk(nx), "bleftt": mk(ny), "brightt": mk(ny)}


def dirichlet_zero_bc(g):
    nx, ny = g.xnum, g.ynum
    return {"btopt": np.zeros((nx, 2), order="F"), "bbottomt": np.zeros((nx, 2), order="F"),
            "bleftt": np.zeros((ny, 2), order="F"), "brightt": np.zeros((ny, 2), order="F")}


def grid_fields(g, seed):
    """Random heat-transport fields of crustal magnitude."""
    rng = np.random.default_rng(seed)
    ny, nx = g.ynum, g.xnum
    F = lambda lo, hi, s: np.asfortranarray(rng.uniform(lo, hi, s))
    yy = np.asfortranarray(np.repeat(g.gridy_b[:, None], nx, axis=1))
    return {
        "tk1": np.asfortranarray(273.0 + 1300.0 * yy / g.gridy_b[-1] + rng.uniform(-50, 50, (ny, nx))),
        "kt1": F(1.5, 4.0, (ny, nx)), "rhocp1": F(2.4e6, 3.6e6, (ny, nx)), "hr1": F(0.0, 2e-6, (ny, nx)),
        "ha1": F(0.0, 0.05, (ny, nx)), "rho1": F(2700.0, 3300.0, (ny, nx)),
        "vx1": F(-3e-10, 3e-10, (ny + 1, nx)), "vy1": F(-3e-10, 3e-10, (ny, nx + 1)),
        "sxy2": F(-3e7, 3e7, (ny, nx)), "etas1": np.asfortranarray(10 ** rng.uniform(19, 23, (ny, nx))),
        "sxx2": F(-3e7, 3e7, (ny - 1, nx - 1)),
        "etan1": np.asfortranarray(10 ** rng.uniform(19, 23, (ny - 1, nx - 1))),
    }


def heat_markers(g, per_cell, seed, nmat=4, outside=2):
    """Jittered markers with thermal properties, material ids 1..nmat and a
    few markers outside the domain."""
    x, y = workloads.marker_positions(g, per_cell, per_cell, seed)
    rng = np.random.default_rng(seed + 100)
    if outside:
        x[:outside] = g.gridx_b[-1] + 1.0 + np.arange(outside)
    n = x.size
    flags = oracle.inside_flags(x, y, g.limits)
    xn, yn = oracle.cell_indices_basic(x, y, flags, g)
    return {
        "x": x, "y": y, "flags": flags, "marker_xn": xn, "marker_yn": yn,
        "TK": 273.0 + 1300.0 * y / g.gridy_b[-1] + rng.uniform(-20, 20, n),
        "kt": rng.uniform(1.5, 4.0, n), "rhocp": rng.uniform(2.4e6, 3.6e6, n), "ha": rng.uniform(0, 0.05, n),
        "matid": rng.integers(1, nmat + 1, n).astype(np.int16),
        "serp": rng.uniform(0, 1e-7, n),
        "eta": 10 ** rng.uniform(19, 23, n), "sxy": rng.uniform(-3e7, 3e7, n), "sxx": rng.uniform(-3e7, 3e7, n),
    }


MAT_HR = np.array([1e-6, 0.5e-6, 0.0, 2e-6])
MAT_MU = np.array([3e10, 2.5e10, 6.7e10, 1e10])
