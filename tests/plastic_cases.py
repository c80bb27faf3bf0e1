"""Seeded grid fields for the nodal plastic-yield tests."""
import numpy as np

from earthbox_gpu.grid import graded_grid, uniform_grid

PARAMS = dict(timestep=3.15e11, timesum=3.0e13, ntimestep=5, viscosity_min=1e18, viscosity_max=1e25,
              iuse_fluid_pressure_for_yield=0, y_sealevel=1.0e4, no_yielding_in_mobile_wall=0,
              no_yielding_in_plate_extension=0, velocity_internal_x=1e-10, x_left_mobile_wall=5.0e4,
              x_right_mobile_wall=8.0e4, y_top_mobile_wall=2.0e4, y_bottom_mobile_wall=6.0e4,
              plate_extension_thickness=1.0e4, plate_extension_width=3.0e4)


def plastic_fields(nx, ny, seed, graded=True):
    g = graded_grid(nx, ny, 2.0e5, 1.0e5, seed) if graded else uniform_grid(nx, ny, 2.0e5, 1.0e5)
    rng = np.random.default_rng(seed + 70)
    b, p = (ny, nx), (ny - 1, nx - 1)

    def F(shape, lo, hi, log=False):
        v = rng.uniform(lo, hi, shape)
        return np.asfortranarray(10 ** v if log else v)
    f = {"plastic_yield": np.asfortranarray((rng.random(b) < 0.3).astype(np.float64)),
         "etas1": F(b, 19, 24, True), "eta_flow": F(b, 20, 25, True), "mus1": F(b, 10, 10.8, True),
         "cohesion_grid": F(b, 0, 2e7), "fric_degrees_grid": F(b, 0, 35), "sxy2": F(b, -8e7, 8e7),
         "sxx2": F(p, -8e7, 8e7), "pr1": F(p, 0, 3e8)}
    off = rng.random(b) < 0.1     # some nodes without cohesion and friction
    f["cohesion_grid"][off] = 0.0
    f["fric_degrees_grid"][off] = 0.0
    yield_error = np.asfortranarray(rng.uniform(-1, 1, b))
    return g, f, yield_error
