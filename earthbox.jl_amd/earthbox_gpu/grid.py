"""Host-side staggered-grid construction (uniform and graded basic grids).

Restates the reference's coordinate and spacing loops so the arrays handed to
the GPU are the same bits the reference would hold in ModelData:
  basic:    staggered_grid/types_manager/uniform/core/Initialization.jl:27-43
            (cumulative sum, so gridx_b[end] need not equal xsize)
  spacing:  staggered_grid/utils/Spacing.jl:6-18, :124-133
  vx / vy:  staggered_grid/utils/VxGridTools.jl:40-99, VyGridTools.jl:41-100
  pressure: staggered_grid/utils/PressureGridTools.jl:34-86
Arrays are 0-based numpy vectors holding the reference's 1-based entries.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def _basic_axis(n: int, step) -> tuple[np.ndarray, np.ndarray]:
    """Cumulative-sum coordinates; step is one value or n-1 per-cell steps."""
    steps = np.broadcast_to(np.asarray(step, dtype=np.float64), (n - 1,))
    g = np.empty(n)
    acc = 0.0
    g[0] = acc
    for i in range(1, n):  # sequential FP64 accumulation, as the reference
        acc = acc + float(steps[i - 1])
        g[i] = acc
    stp = g[1:] - g[:-1]
    return g, stp


def _staggered_axis(n: int, g: np.ndarray, stp: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    gs = np.empty(n + 1)
    gs[0] = g[0] - stp[0] / 2.0
    gs[n] = g[n - 1] + stp[n - 2] / 2.0
    gs[1:n] = (g[1:n] + g[0:n - 1]) / 2.0
    stps = np.empty(n)
    stps[0] = stp[0]
    stps[n - 1] = stp[n - 2]
    stps[1:n - 1] = (g[2:n] - g[0:n - 2]) / 2.0
    return gs, stps


def _pressure_axis(n: int, g: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    stpp = (g[2:n] - g[0:n - 2]) / 2.0
    gp = np.empty(n - 1)
    acc = (g[0] + g[1]) / 2.0
    gp[0] = acc
    for i in range(1, n - 1):
        acc = acc + stpp[i - 1]
        gp[i] = acc
    return gp, stpp


@dataclass
class StaggeredGrid2D:
    xnum: int
    ynum: int
    xsize: float
    ysize: float
    xstpavg: float
    ystpavg: float
    gridx_b: np.ndarray
    gridy_b: np.ndarray
    xstp_b: np.ndarray
    ystp_b: np.ndarray
    gridx_vy: np.ndarray
    xstp_vy: np.ndarray
    gridy_vx: np.ndarray
    ystp_vx: np.ndarray
    gridx_pr: np.ndarray
    gridy_pr: np.ndarray
    xstp_pr: np.ndarray
    ystp_pr: np.ndarray

    @property
    def limits(self) -> tuple[float, float, float, float]:
        """(xmin, xmax, ymin, ymax) as staggered_grid/utils/Dimensions.jl:6-19."""
        return (float(self.gridx_b[0]), float(self.gridx_b[-1]),
                float(self.gridy_b[0]), float(self.gridy_b[-1]))


def _assemble(xnum, ynum, xsize, ysize, xstep, ystep) -> StaggeredGrid2D:
    xstpavg = xsize / (xnum - 1)
    ystpavg = ysize / (ynum - 1)
    gx, sx = _basic_axis(xnum, xstep)
    gy, sy = _basic_axis(ynum, ystep)
    gxv, sxv = _staggered_axis(xnum, gx, sx)
    gyv, syv = _staggered_axis(ynum, gy, sy)
    gxp, sxp = _pressure_axis(xnum, gx)
    gyp, syp = _pressure_axis(ynum, gy)
    return StaggeredGrid2D(xnum, ynum, xsize, ysize, xstpavg, ystpavg,
                           gx, gy, sx, sy, gxv, sxv, gyv, syv, gxp, gyp, sxp, syp)


def graded_grid(xnum: int, ynum: int, xsize: float, ysize: float, seed: int,
                spread: float = 0.6) -> StaggeredGrid2D:
    """Non-uniform staggered grid: per-cell steps drawn in
    [(1-spread), (1+spread)] x the mean step and rescaled to the box size, then
    accumulated as the reference accumulates a refined grid's steps."""
    if xnum < 3 or ynum < 3:
        raise ValueError("xnum and ynum must be >= 3")
    rng = np.random.default_rng(seed)

    def steps(n, size):
        w = 1.0 + spread * (2.0 * rng.random(n - 1) - 1.0)
        return w * (size / w.sum())
    return _assemble(xnum, ynum, xsize, ysize, steps(xnum, xsize), steps(ynum, ysize))


def uniform_grid(xnum: int, ynum: int, xsize: float, ysize: float) -> StaggeredGrid2D:
    """Uniform staggered grid as the reference builds it."""
    if xnum < 3 or ynum < 3:
        raise ValueError("xnum and ynum must be >= 3")
    return _assemble(xnum, ynum, xsize, ysize, xsize / (xnum - 1), ysize / (ynum - 1))
