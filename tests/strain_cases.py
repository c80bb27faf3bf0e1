"""Seeded velocity and rheology fields for the grid strain-rate and stress tests."""
import numpy as np

from earthbox_gpu.grid import graded_grid


def strain_fields(nx, ny, seed, linear=None):
    """Random (or, with linear=(a, b, c, d), vx = a x + b y and vy = c x + d y
    sampled at the staggered nodes, ghosts included) velocities plus random
    viscosity, modulus and old stress."""
    g = graded_grid(nx, ny, 2.0e5, 1.0e5, seed)
    rng = np.random.default_rng(seed + 90)
    if linear is None:
        vx1 = rng.uniform(-1e-9, 1e-9, (ny + 1, nx))
        vy1 = rng.uniform(-1e-9, 1e-9, (ny, nx + 1))
    else:
        a, b, c, d = linear
        vx1 = a * g.gridx_b[None, :] + b * g.gridy_vx[:, None]
        vy1 = c * g.gridx_vy[None, :] + d * g.gridy_b[:, None]
    b_, p_ = (ny, nx), (ny - 1, nx - 1)
    f = {"etas1": 10 ** rng.uniform(19, 24, b_), "mus1": 10 ** rng.uniform(10, 10.8, b_),
         "sxy1": rng.uniform(-1e7, 1e7, b_), "etan1": 10 ** rng.uniform(19, 24, p_),
         "mun1": 10 ** rng.uniform(10, 10.8, p_), "sxx1": rng.uniform(-1e7, 1e7, p_)}
    f = {k: np.asfortranarray(v) for k, v in f.items()}
    return g, np.asfortranarray(vx1), np.asfortranarray(vy1), f
