"""Seeded synthetic inputs for the hot path (markers, velocity fields, models).

All generators use numpy.random.default_rng(seed) so the GPU path, the CPU
oracle and the benchmark see the same bits.
"""
from __future__ import annotations

import numpy as np

from .grid import StaggeredGrid2D


def marker_positions(g: StaggeredGrid2D, nmx: int, nmy: int, seed: int,
                     jitter: bool = True) -> tuple[np.ndarray, np.ndarray]:
    """Markers nmx x nmy per cell, ordered x-major then y as the reference
    (markers/coordinates/init_manager/types/MarkersRandomized.jl:5-38 with
    mxnum=(xnum-1)*nmx, mxstep=xsize/mxnum, ModelDataContainer.jl:355-392).

    Coordinates are i*step - step/2 + (r - 0.5)*step.  Unlike the reference,
    which reuses one random draw for both axes, x and y use independent draws.
    """
    mxnum = (g.xnum - 1) * nmx
    mynum = (g.ynum - 1) * nmy
    mxstep = g.xsize / mxnum
    mystep = g.ysize / mynum
    rng = np.random.default_rng(seed)
    ix = np.repeat(np.arange(1, mxnum + 1, dtype=np.float64), mynum)
    iy = np.tile(np.arange(1, mynum + 1, dtype=np.float64), mxnum)
    if jitter:
        rx = rng.random(mxnum * mynum)
        ry = rng.random(mxnum * mynum)
    else:
        rx = np.full(mxnum * mynum, 0.5)
        ry = np.full(mxnum * mynum, 0.5)
    x = ix * mxstep - mxstep / 2.0 + (rx - 0.5) * mxstep
    y = iy * mystep - mystep / 2.0 + (ry - 0.5) * mystep
    return x, y


def random_velocity_fields(g: StaggeredGrid2D, seed: int, vmax: float = 1e-9):
    """Smooth-ish random vx1 ((ynum+1) x xnum) and vy1 (ynum x (xnum+1)),
    Fortran-ordered, used for advection parity tests."""
    rng = np.random.default_rng(seed)
    ny, nx = g.ynum, g.xnum
    def field(shape):
        base = rng.standard_normal(shape)
        # cheap smoothing so neighbouring nodes correlate
        for _ in range(2):
            base = (base + np.roll(base, 1, 0) + np.roll(base, 1, 1)) / 3.0
        return np.asfortranarray(vmax * base / max(1e-300, np.abs(base).max()))
    return field((ny + 1, nx)), field((ny, nx + 1))


def solid_body_rotation_fields(g: StaggeredGrid2D, omega: float):
    """Rigid rotation about the box centre sampled on the staggered nodes
    (the closed-form field of kinematics/Kinematics.jl:14-43):
    vx = -omega*(y - yc), vy = omega*(x - xc)."""
    xc = 0.5 * (g.gridx_b[0] + g.gridx_b[-1])
    yc = 0.5 * (g.gridy_b[0] + g.gridy_b[-1])
    vx1 = np.asfortranarray(-omega * (g.gridy_vx[:, None] - yc) * np.ones((1, g.xnum)))
    vy1 = np.asfortranarray(omega * (g.gridx_vy[None, :] - xc) * np.ones((g.ynum, 1)))
    return vx1, vy1
