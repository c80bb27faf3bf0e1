"""Host-side container of one Stokes-continuity system on the assembled-operator
conventions (the inputs StokesBuildData gathers from ModelData,
solver/build_system/utils/BuildStructs.jl:37-177), plus BC-array helpers.

Matrices are numpy Fortran-ordered with the reference's shapes:
  etas0 (ny, nx), etan0 (ny-1, nx-1), rho1_vy (ny, nx+1),
  RX1 (ny+1, nx), RY1 (ny, nx+1), RC1 (ny-1, nx-1),
  btop*/bbottom* (nx+1, 2), bleft*/bright* (ny+1, 2).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .grid import StaggeredGrid2D


def F(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


@dataclass
class VelocityBCs:
    """Ghost-node coefficient arrays b*[:,1] (C) and b*[:,2] (D): v_ghost = C + D*v."""
    btopx: np.ndarray
    btopy: np.ndarray
    bbottomx: np.ndarray
    bbottomy: np.ndarray
    bleftx: np.ndarray
    blefty: np.ndarray
    brightx: np.ndarray
    brighty: np.ndarray
    bintern_zone: np.ndarray = field(default_factory=lambda: np.full(8, -1, dtype=np.int64))
    bintern_velocity: np.ndarray = field(default_factory=lambda: np.zeros(8))


def free_slip_bcs(g: StaggeredGrid2D, vx_left=0.0, vx_right=0.0, vy_top=0.0,
                  vy_bottom=0.0) -> VelocityBCs:
    """Free slip on all walls with prescribed normal velocities, as
    boundary_conditions/bc_reset/types/ViscousBlockBC.jl:9-32 sets them
    (StokesGradientBC zero tangential gradient: C=0, D=1; VelocityBC normal
    velocity: C=v, D=0)."""
    nx1, ny1 = g.xnum + 1, g.ynum + 1
    def arr(n, c, d):
        a = np.empty((n, 2), order="F")
        a[:, 0] = c
        a[:, 1] = d
        return a
    return VelocityBCs(
        btopx=arr(nx1, 0.0, 1.0), btopy=arr(nx1, vy_top, 0.0),
        bbottomx=arr(nx1, 0.0, 1.0), bbottomy=arr(nx1, vy_bottom, 0.0),
        bleftx=arr(ny1, vx_left, 0.0), blefty=arr(ny1, 0.0, 1.0),
        brightx=arr(ny1, vx_right, 0.0), brighty=arr(ny1, 0.0, 1.0))


def no_slip_bcs(g: StaggeredGrid2D) -> VelocityBCs:
    """No slip on all walls (StokesBuildManager.jl:523-548): tangential ghost
    velocity mirrors the interior (D=-1), normal velocity zero."""
    nx1, ny1 = g.xnum + 1, g.ynum + 1
    def arr(n, c, d):
        a = np.empty((n, 2), order="F")
        a[:, 0] = c
        a[:, 1] = d
        return a
    return VelocityBCs(
        btopx=arr(nx1, 0.0, -1.0), btopy=arr(nx1, 0.0, 0.0),
        bbottomx=arr(nx1, 0.0, -1.0), bbottomy=arr(nx1, 0.0, 0.0),
        bleftx=arr(ny1, 0.0, 0.0), blefty=arr(ny1, 0.0, -1.0),
        brightx=arr(ny1, 0.0, 0.0), brighty=arr(ny1, 0.0, -1.0))


@dataclass
class StokesProblem:
    grid: StaggeredGrid2D
    etas0: np.ndarray
    etan0: np.ndarray
    rho1_vy: np.ndarray
    RX1: np.ndarray
    RY1: np.ndarray
    RC1: np.ndarray
    bcs: VelocityBCs
    pressure_bc_mode: int = 0
    pressure_bc: float = 0.0
    timestep: float = 0.0
    gravity_y: float = 10.0
    iuse_interface_stabilization: int = 0
    pscale: float = 0.0

    @property
    def N(self) -> int:
        return 3 * (self.grid.xnum - 1) * (self.grid.ynum - 1)

    def __post_init__(self):
        g = self.grid
        ny, nx = g.ynum, g.xnum
        shapes = {"etas0": (ny, nx), "etan0": (ny - 1, nx - 1), "rho1_vy": (ny, nx + 1),
                  "RX1": (ny + 1, nx), "RY1": (ny, nx + 1), "RC1": (ny - 1, nx - 1)}
        for k, s in shapes.items():
            a = F(getattr(self, k))
            if a.shape != s:
                raise ValueError(f"{k} has shape {a.shape}, expected {s}")
            setattr(self, k, a)
        if self.pscale == 0.0:
            # SystemSolver.jl:108-115
            self.pscale = 2.0 * float(self.etan0[0, 0]) / (g.xstpavg + g.ystpavg)
