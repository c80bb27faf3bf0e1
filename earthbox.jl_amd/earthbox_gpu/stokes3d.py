"""Host-side mirror of the reference's experimental 3-D multigrid and the 3-D
Stokes solve on its system.

Multigrid3DGPU drives the reference's 3-D multigrid
(multigrid/multigrid_solver/MultigridSolver.jl:113-135 for MultigridData3d:
LevelManager, Smoother, Restriction, Prolongation, MultigridVCycle) with the
device-resident FP64 hierarchy.  Stokes3DSolverGPU solves the same residual
equations (Residuals.jl:10-560, free-slip ghost passes of
BoundaryConditions.jl:92-174, pressure gauge of PressureUpdate.jl:5-21) with
FP64 FGMRES preconditioned by one V-cycle in FP32 or FP64.

Array layout: column-major (y fastest, then x, then z) with the reference's
LevelData shapes; vectors of the solver are [vx | vy | vz | pr].
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import MGRunParams, SolveStats, check, lib
from .context import GPUContext
from .device import empty, ptr
from .stokes import default_opts

MG3_FIELDS = {"etan": 0, "etaxy": 1, "etaxz": 2, "etayz": 3, "vx": 4, "vy": 5, "vz": 6, "pr": 7,
              "RX": 8, "RY": 9, "RZ": 10, "RC": 11, "rx": 12, "ry": 13, "rz": 14, "rc": 15}


def shapes3(xnum: int, ynum: int, znum: int) -> dict:
    """LevelData array shapes (y, x, z) of one level (LevelManager.jl:78-134)."""
    ny, nx, nz = ynum, xnum, znum
    vx, vy, vz = (ny + 1, nx, nz + 1), (ny, nx + 1, nz + 1), (ny + 1, nx + 1, nz)
    pr = (ny - 1, nx - 1, nz - 1)
    return {"vx": vx, "vy": vy, "vz": vz, "pr": pr, "etan": pr, "etaxy": (ny, nx, nz - 1),
            "etaxz": (ny - 1, nx, nz), "etayz": (ny, nx - 1, nz), "RX": vx, "RY": vy, "RZ": vz,
            "RC": pr, "rx": vx, "ry": vy, "rz": vz, "rc": pr}


class Multigrid3DGPU:
    """The reference's experimental 3-D multigrid hierarchy on the device (FP64)."""

    def __init__(self, ctx: GPUContext, xnum, ynum, znum, xsize, ysize, zsize, nlevels=0):
        self.ctx = ctx
        n = C.c_int()
        check(lib.ebg_mg3d_create(ctx.handle, xnum, ynum, znum, xsize, ysize, zsize, nlevels,
                                  C.byref(n)), "ebg_mg3d_create")
        self.nlevels = n.value
        self.shapes = []
        for k in range(self.nlevels):
            x, y, z = C.c_int64(), C.c_int64(), C.c_int64()
            check(lib.ebg_mg3d_level_shape(ctx.handle, k, C.byref(x), C.byref(y), C.byref(z)),
                  "ebg_mg3d_level_shape")
            self.shapes.append((x.value, y.value, z.value))

    def field_shape(self, level, name):
        return shapes3(*self.shapes[level])[name]

    def get(self, level, name) -> torch.Tensor:
        t = empty(self.field_shape(level, name), np.float64)
        check(lib.ebg_mg3d_copy_field(self.ctx.handle, level, MG3_FIELDS[name], ptr(t), None),
              "ebg_mg3d_copy_field")
        return t

    def set(self, level, name, t: torch.Tensor) -> None:
        if tuple(t.shape) != self.field_shape(level, name):
            raise ValueError(f"{name}: expected shape {self.field_shape(level, name)}")
        check(lib.ebg_mg3d_copy_field(self.ctx.handle, level, MG3_FIELDS[name], None,
                                      ptr(t, np.float64)), "ebg_mg3d_copy_field")

    def set_fused_limit(self, nodes: int):
        """Levels of at most `nodes` nodes smooth in one single-workgroup launch."""
        check(lib.ebg_mg3d_set_fused_limit(self.ctx.handle, int(nodes)), "ebg_mg3d_set_fused_limit")

    def restrict_viscosity(self):
        check(lib.ebg_mg3d_restrict_viscosity(self.ctx.handle), "ebg_mg3d_restrict_viscosity")

    def set_bc(self, level):
        check(lib.ebg_mg3d_set_bc(self.ctx.handle, level), "ebg_mg3d_set_bc")

    def smoother(self, level, nsweeps, theta_s=0.9, theta_c=0.3, prnorm=0.0):
        check(lib.ebg_mg3d_smoother(self.ctx.handle, level, nsweeps, theta_s, theta_c, prnorm),
              "ebg_mg3d_smoother")

    def restrict(self, level):
        check(lib.ebg_mg3d_restrict(self.ctx.handle, level), "ebg_mg3d_restrict")

    def prolong(self, level, relax_v=1.0, relax_p=1.0):
        check(lib.ebg_mg3d_prolong(self.ctx.handle, level, relax_v, relax_p), "ebg_mg3d_prolong")

    def vcycles(self, ncycles, finest_sweeps=2, multiplier=2, cap=0, theta_s=0.9, theta_c=0.3,
                relax_v=1.0, relax_p=1.0, prnorm=0.0):
        check(lib.ebg_mg3d_vcycles(self.ctx.handle, ncycles, finest_sweeps, multiplier, cap,
                                   theta_s, theta_c, relax_v, relax_p, prnorm), "ebg_mg3d_vcycles")

    def run(self, nvcycles, convergence_criterion, nviscosity_jumps=12.0, nvcycles_viscosity_jump=15,
            resmax=1e-6, stokesscale=1.0, continscale=1.0, finest_sweeps=2, multiplier=2, cap=0,
            theta_s=0.9, theta_c=0.3, relax_v=1.0, relax_p=1.0, pressure_bc=0.0):
        """MultigridSolver.run_multigrid_solver (MultigridSolver.jl:113-135,
        use_multimulti = false): viscosity continuation, V-cycles, mean
        principal residuals and the reference's stopping test.  Level 0 must
        hold the viscosity, right-hand sides and initial solution.  Returns
        (cycles run, converged, history[cycles, 4] of log10 RMS residuals
        x, y, z, continuity)."""
        prm = MGRunParams(nvcycles=nvcycles, nvcycles_viscosity_jump=nvcycles_viscosity_jump,
                          finest_sweeps=finest_sweeps, sweep_multiplier=multiplier, sweep_cap=cap,
                          nviscosity_jumps=nviscosity_jumps,
                          convergence_criterion=convergence_criterion, resmax=resmax,
                          stokesscale=stokesscale, continscale=continscale, theta_s=theta_s,
                          theta_c=theta_c, relax_v=relax_v, relax_p=relax_p, pressure_bc=pressure_bc)
        hist = np.zeros((max(nvcycles, 1), 4))
        done, conv = C.c_int(), C.c_int()
        check(lib.ebg_mg3d_run(self.ctx.handle, C.byref(prm), C.byref(done), C.byref(conv),
                               hist.ctypes.data), "ebg_mg3d_run")
        return done.value, bool(conv.value), hist[:done.value].copy()

    def run_multimulti(self, nvcycles, convergence_criterion, nviscosity_jumps=5.9999, nvcycles_viscosity_jump=3,
                       stokesscale=1.0, continscale=1.0, finest_sweeps=2, multiplier=2, cap=0, theta_s=0.9,
                       theta_c=0.3, relax_v=1.0, relax_p=1.0, pressure_bc=0.0):
        """MultigridSolver.run_multigrid_solver with use_multimulti = true:
        level 0 holds the global solution, level 1 solves for corrections.
        Fill level 0 as for run(); on return it holds the global solution.
        Returns (cycles run, converged, principal history[cycles, 4], global
        history[updates, 4])."""
        prm = MGRunParams(nvcycles=nvcycles, nvcycles_viscosity_jump=nvcycles_viscosity_jump,
                          finest_sweeps=finest_sweeps, sweep_multiplier=multiplier, sweep_cap=cap,
                          nviscosity_jumps=nviscosity_jumps,
                          convergence_criterion=convergence_criterion, resmax=0.0,
                          stokesscale=stokesscale, continscale=continscale, theta_s=theta_s,
                          theta_c=theta_c, relax_v=relax_v, relax_p=relax_p, pressure_bc=pressure_bc)
        hist = np.zeros((max(nvcycles, 1), 4))
        cap_g = max(nvcycles, 1)
        ghist = np.zeros((cap_g, 4))
        done, conv, ng = C.c_int(), C.c_int(), C.c_int()
        check(lib.ebg_mg3d_run_multimulti(self.ctx.handle, C.byref(prm), C.byref(done), C.byref(conv),
                                          hist.ctypes.data, C.byref(ng), ghist.ctypes.data, cap_g),
              "ebg_mg3d_run_multimulti")
        return done.value, bool(conv.value), hist[:done.value].copy(), ghist[:ng.value].copy()


class Stokes3DSolverGPU:
    """FGMRES(FP64) + V-cycle (FP32 or FP64) solve of the 3-D multigrid's system."""

    def __init__(self, ctx: GPUContext):
        self.ctx = ctx
        self.dims = None

    def set_system(self, xnum, ynum, znum, xsize, ysize, zsize, etan, etaxy, etaxz, etayz):
        sh = shapes3(xnum, ynum, znum)
        for name, t in (("etan", etan), ("etaxy", etaxy), ("etaxz", etaxz), ("etayz", etayz)):
            if tuple(t.shape) != sh[name]:
                raise ValueError(f"{name}: expected shape {sh[name]}")
        check(lib.ebg_stokes3d_set_system(self.ctx.handle, xnum, ynum, znum, xsize, ysize, zsize,
                                          ptr(etan, np.float64), ptr(etaxy, np.float64),
                                          ptr(etaxz, np.float64), ptr(etayz, np.float64)),
              "ebg_stokes3d_set_system")
        self.dims = (xnum, ynum, znum)
        self.sh = sh
        n = C.c_int64()
        check(lib.ebg_stokes3d_vector_length(self.ctx.handle, C.byref(n)), "ebg_stokes3d_vector_length")
        self.n = n.value
        self.offsets = {}
        o = 0
        for k in ("vx", "vy", "vz", "pr"):
            self.offsets[k] = o
            o += int(np.prod(sh[k]))
        assert o == self.n

    def vector(self) -> torch.Tensor:
        return torch.zeros(self.n, dtype=torch.float64, device=self.ctx.device)

    def split(self, v: torch.Tensor) -> dict:
        """3-D column-major views of the four parts of a solver vector."""
        out = {}
        for k in ("vx", "vy", "vz", "pr"):
            ny, nx, nz = self.sh[k]
            o = self.offsets[k]
            out[k] = v[o:o + ny * nx * nz].view(nz, nx, ny).permute(2, 1, 0)
        return out

    def pack(self, parts: dict) -> torch.Tensor:
        v = self.vector()
        views = self.split(v)
        for k in ("vx", "vy", "vz", "pr"):
            views[k].copy_(parts[k])
        return v

    def apply(self, x: torch.Tensor, y: torch.Tensor) -> None:
        check(lib.ebg_stokes3d_apply(self.ctx.handle, ptr(x, np.float64, self.n),
                                     ptr(y, np.float64, self.n)), "ebg_stokes3d_apply")

    def precondition(self, r: torch.Tensor, z: torch.Tensor, **opts) -> None:
        o = default_opts(**{**self.DEFAULTS, **opts})
        check(lib.ebg_stokes3d_precondition(self.ctx.handle, ptr(r, np.float64, self.n),
                                            ptr(z, np.float64, self.n), C.byref(o)),
              "ebg_stokes3d_precondition")

    # preconditioner settings measured best on the 3-D inclusion problems
    # (six sweeps on the finest level, pressure relaxation 1, coarse cap 64)
    DEFAULTS = {"finest_sweeps": 6, "theta_c": 1.0, "sweep_cap": 64, "restart": 30}

    def solve(self, rhs: torch.Tensor, prnorm: float = 0.0, x: torch.Tensor | None = None, **opts):
        """Solve for rhs = [RX | RY | RZ | RC]; returns (x, stats dict)."""
        o = default_opts(**{**self.DEFAULTS, **opts})
        if x is None:
            x = self.vector()
        st = SolveStats()
        rc = lib.ebg_stokes3d_solve(self.ctx.handle, ptr(rhs, np.float64, self.n), float(prnorm),
                                    ptr(x, np.float64, self.n), C.byref(o), C.byref(st))
        if rc < 0:
            check(rc, "ebg_stokes3d_solve")
        stats = {"iterations": st.iterations, "restarts": st.restarts, "relres": st.relres,
                 "relres_initial": st.relres_initial, "seconds": st.seconds,
                 "converged": rc == 0}
        return x, stats
