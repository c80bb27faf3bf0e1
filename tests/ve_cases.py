"""Seeded grid fields for the viscoelastic-terms and right-hand-side tests."""
import numpy as np

from earthbox_gpu.grid import graded_grid


def ve_fields(nx, ny, seed, dilatation=True):
    g = graded_grid(nx, ny, 2.0e5, 1.0e5, seed)
    rng = np.random.default_rng(seed + 50)

    def F(shape, lo, hi, log=False):
        v = rng.uniform(lo, hi, shape)
        return np.asfortranarray(10 ** v if log else v)
    b, p, vy = (ny, nx), (ny - 1, nx - 1), (ny, nx + 1)
    f = {"etas1": F(b, 18, 24, True), "mus1": F(b, 9.5, 11, True), "sxy1": F(b, -1e7, 1e7),
         "rho1": F(b, 2700, 3300), "etan1": F(p, 18, 24, True), "mun1": F(p, 9.5, 11, True),
         "sxx1": F(p, -1e7, 1e7), "rho1_vy": F(vy, 2700, 3300)}
    if dilatation:
        f["dilatation_grid"] = F(p, 0, 30)
        f["eii_plastic_pressure"] = F(p, 0, 1e-14)
    return g, f
