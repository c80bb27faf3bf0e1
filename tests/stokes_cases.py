"""Seeded Stokes-continuity test systems covering the reference's boundary
variants (free slip, no slip, prescribed normal velocity, internal velocity
zones, pressure modes 0/1/2, interface stabilization)."""
import numpy as np

from earthbox_gpu import workloads
from earthbox_gpu.grid import uniform_grid
from earthbox_gpu.stokes_problem import StokesProblem, free_slip_bcs, no_slip_bcs


def random_case(nx, ny, seed, bc="free", mode=0, stab=0, zone=False):
    rng = np.random.default_rng(seed)
    g = uniform_grid(nx, ny, 1.0e5 * (nx - 1) / 10, 1.0e5 * (ny - 1) / 10)
    F = lambda s, lo, hi: np.asfortranarray(10 ** rng.uniform(lo, hi, s))
    etas = F((ny, nx), 19, 23)
    etan = F((ny - 1, nx - 1), 19, 23)
    rho = np.asfortranarray(3000 + 300 * rng.random((ny, nx + 1)))
    RX = np.asfortranarray(rng.standard_normal((ny + 1, nx)) * 1e3)
    RY = np.asfortranarray(rng.standard_normal((ny, nx + 1)) * 1e4)
    RC = np.asfortranarray(rng.standard_normal((ny - 1, nx - 1)) * 1e-15)
    if bc == "free":
        b = free_slip_bcs(g)
    elif bc == "noslip":
        b = no_slip_bcs(g)
    else:  # extension: prescribed normal velocities and a mixed top
        b = free_slip_bcs(g, vx_left=-1.6e-10, vx_right=1.6e-10, vy_top=0.0, vy_bottom=-2e-11)
        b.btopx[:, 1] = -1.0
    if zone:
        b.bintern_zone = np.array([nx // 2, 2, ny - 3, -1, nx // 3, 3, ny - 4, -1], dtype=np.int64)
        b.bintern_velocity = np.array([0, 0, 0, 3e-10, 0, 0, 0, -2e-10], dtype=np.float64)
    return StokesProblem(grid=g, etas0=etas, etan0=etan, rho1_vy=rho, RX1=RX, RY1=RY, RC1=RC,
                         bcs=b, pressure_bc_mode=mode, pressure_bc=1e5 if mode else 3e4,
                         timestep=3e12 if stab else 0.0, gravity_y=9.81,
                         iuse_interface_stabilization=stab)


CASES = [
    dict(nx=9, ny=7, seed=1),
    dict(nx=12, ny=9, seed=2, bc="noslip"),
    dict(nx=17, ny=11, seed=3, bc="ext", stab=1),
    dict(nx=15, ny=13, seed=4, zone=True),
    dict(nx=11, ny=10, seed=5, mode=1),
    dict(nx=10, ny=11, seed=6, mode=2, bc="noslip", stab=1),
    dict(nx=64, ny=33, seed=7, bc="ext", stab=1, zone=True),
]
