"""Host-side mirror of the reference's Stokes solve seam and multigrid driver.

StokesSolverGPU replaces, per Picard iteration, the reference's
  SystemSolver.solve_system_optimized (solver/system_solver/SystemSolver.jl:67-91),
  SystemSolver.process_stokes_solution! (:141-151) and
  StokesResiduals.stokes_calc_nonlinear_system_residual_optimized!
  (solver/residuals/StokesResiduals.jl:63-82),
on device-resident fields.  MultigridGPU drives the reference's experimental
2-D multigrid (multigrid/multigrid_solver/MultigridSolver.jl:113-135).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import MGRunParams, SolveOpts, SolveStats, check, lib
from .context import GPUContext
from .device import empty, ptr, to_device
from .stokes_problem import StokesProblem

MG_FIELDS = {"etas": 0, "etan": 1, "vx": 2, "vy": 3, "pr": 4, "RX": 5, "RY": 6, "RC": 7,
             "rx": 8, "ry": 9, "rc": 10}


RESTART_2D = 100   # FGMRES restart of the 2-D solves (the C default, 30, suits the 3-D basis memory)


def default_opts(**kw) -> SolveOpts:
    o = SolveOpts()
    check(lib.ebg_stokes_solve_opts_default(C.byref(o)), "ebg_stokes_solve_opts_default")
    for k, v in kw.items():
        if k == "bc_d":
            for q in range(4):
                o.bc_d[q] = float(v[q])
        else:
            setattr(o, k, v)
    return o


def bc_mirror_factors(prob: StokesProblem) -> tuple[float, float, float, float]:
    """Tangential ghost factors D of the walls (top/bottom for vx, left/right
    for vy) for the preconditioner's homogeneous boundary conditions."""
    b = prob.bcs
    return (float(b.btopx[1, 1]), float(b.bbottomx[1, 1]), float(b.blefty[1, 1]),
            float(b.brighty[1, 1]))


class StokesSolverGPU:
    """One assembled Stokes-continuity system resident on the GPU."""

    def __init__(self, ctx: GPUContext):
        self.ctx = ctx
        self.prob: StokesProblem | None = None

    def set_system(self, prob: StokesProblem, fields: dict | None = None) -> None:
        """Upload the fields (numpy, or device tensors in `fields`)."""
        if self.ctx.grid is not prob.grid:
            self.ctx.set_grid(prob.grid)
        f = fields or {}
        self._dev = {k: f.get(k) if f.get(k) is not None else to_device(getattr(prob, k))
                     for k in ("etas0", "etan0", "rho1_vy", "RX1", "RY1", "RC1")}
        b = prob.bcs
        self._host = [np.asfortranarray(a, dtype=np.float64) for a in
                      (b.btopx, b.btopy, b.bbottomx, b.bbottomy, b.bleftx, b.blefty,
                       b.brightx, b.brighty)]
        self._zone = np.ascontiguousarray(b.bintern_zone, dtype=np.int64)
        self._zvel = np.ascontiguousarray(b.bintern_velocity, dtype=np.float64)
        d = self._dev
        check(lib.ebg_stokes_set_system(
            self.ctx.handle, ptr(d["etas0"]), ptr(d["etan0"]), ptr(d["rho1_vy"]),
            ptr(d["RX1"]), ptr(d["RY1"]), ptr(d["RC1"]),
            *[a.ctypes.data for a in self._host], self._zone.ctypes.data, self._zvel.ctypes.data,
            int(prob.pressure_bc_mode), float(prob.pressure_bc), float(prob.pscale),
            float(prob.timestep), float(prob.gravity_y), int(prob.iuse_interface_stabilization)),
            "ebg_stokes_set_system")
        self.prob = prob

    def new_vector(self) -> torch.Tensor:
        return torch.zeros(self.prob.N, dtype=torch.float64, device=self.ctx.device)

    def apply(self, S: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        out = self.new_vector() if out is None else out
        check(lib.ebg_stokes_apply(self.ctx.handle, ptr(S, np.float64, self.prob.N),
                                   ptr(out, np.float64, self.prob.N)), "ebg_stokes_apply")
        return out

    def rhs(self) -> torch.Tensor:
        out = self.new_vector()
        check(lib.ebg_stokes_rhs(self.ctx.handle, ptr(out, np.float64, self.prob.N)), "ebg_stokes_rhs")
        return out

    def stokes_calc_nonlinear_system_residual(self, S: torch.Tensor) -> torch.Tensor:
        out = self.new_vector()
        check(lib.ebg_stokes_residual(self.ctx.handle, ptr(S, np.float64, self.prob.N),
                                      ptr(out, np.float64, self.prob.N)), "ebg_stokes_residual")
        return out

    def solve_system(self, S: torch.Tensor | None = None, **opts) -> tuple[torch.Tensor, dict]:
        """FGMRES + multigrid; returns (S, stats dict).  Raises on hard errors;
        stats["converged"] is False when the iteration budget ran out."""
        S = self.new_vector() if S is None else S
        o = default_opts(**{"restart": RESTART_2D, **opts})
        st = SolveStats()
        rc = check(lib.ebg_stokes_solve(self.ctx.handle, ptr(S, np.float64, self.prob.N),
                                        C.byref(o), C.byref(st)), "ebg_stokes_solve")
        stats = {k: getattr(st, k) for k, _ in SolveStats._fields_}
        stats["converged"] = (rc == 0)
        return S, stats

    def solve_rhs(self, R: torch.Tensor, S: torch.Tensor | None = None, **opts):
        """solve_system with a caller right-hand side R (reference order)."""
        S = self.new_vector() if S is None else S
        o = default_opts(**{"restart": RESTART_2D, **opts})
        st = SolveStats()
        rc = check(lib.ebg_stokes_solve_rhs(self.ctx.handle, ptr(R, np.float64, self.prob.N),
                                            ptr(S, np.float64, self.prob.N), C.byref(o), C.byref(st)),
                   "ebg_stokes_solve_rhs")
        stats = {k: getattr(st, k) for k, _ in SolveStats._fields_}
        stats["converged"] = (rc == 0)
        return S, stats

    def precondition(self, r: torch.Tensor, out: torch.Tensor | None = None, **opts) -> torch.Tensor:
        """One application z = M r of the solve's multigrid preconditioner."""
        out = self.new_vector() if out is None else out
        o = default_opts(**opts)
        check(lib.ebg_stokes_precondition(self.ctx.handle, ptr(r, np.float64, self.prob.N),
                                          ptr(out, np.float64, self.prob.N), C.byref(o)),
              "ebg_stokes_precondition")
        return out

    def compute_stokes_residuals(self, vx1, vy1, pr1):
        """StokesResiduals.compute_stokes_residuals! (StokesResiduals.jl:94-137):
        (resx1, resy1, resc1) on the vx, vy and pressure grids."""
        g = self.prob.grid
        rx = empty((g.ynum + 1, g.xnum), np.float64)
        ry = empty((g.ynum, g.xnum + 1), np.float64)
        rc = empty((g.ynum - 1, g.xnum - 1), np.float64)
        check(lib.ebg_compute_stokes_residuals(self.ctx.handle, ptr(vx1), ptr(vy1), ptr(pr1), ptr(rx),
                                               ptr(ry), ptr(rc)), "ebg_compute_stokes_residuals")
        return rx, ry, rc

    def process_stokes_solution(self, S: torch.Tensor):
        g = self.prob.grid
        vx1 = empty((g.ynum + 1, g.xnum), np.float64)
        vy1 = empty((g.ynum, g.xnum + 1), np.float64)
        pr1 = empty((g.ynum - 1, g.xnum - 1), np.float64)
        check(lib.ebg_stokes_unpack(self.ctx.handle, ptr(S, np.float64, self.prob.N),
                                    ptr(vx1), ptr(vy1), ptr(pr1)), "ebg_stokes_unpack")
        return vx1, vy1, pr1


class MultigridGPU:
    """The reference's experimental 2-D multigrid hierarchy on the device (FP64)."""

    def __init__(self, ctx: GPUContext, xnum, ynum, xsize, ysize, nlevels=0):
        self.ctx = ctx
        n = C.c_int()
        check(lib.ebg_mg2d_create(ctx.handle, xnum, ynum, xsize, ysize, nlevels, C.byref(n)),
              "ebg_mg2d_create")
        self.nlevels = n.value
        self.shapes = []
        for k in range(self.nlevels):
            x, y = C.c_int64(), C.c_int64()
            check(lib.ebg_mg2d_level_shape(ctx.handle, k, C.byref(x), C.byref(y)), "level_shape")
            self.shapes.append((x.value, y.value))

    def field_shape(self, level, name):
        x, y = self.shapes[level]
        return {"etas": (y, x), "etan": (y - 1, x - 1), "vx": (y + 1, x), "vy": (y, x + 1),
                "pr": (y - 1, x - 1), "RX": (y + 1, x), "RY": (y, x + 1), "RC": (y - 1, x - 1),
                "rx": (y + 1, x), "ry": (y, x + 1), "rc": (y - 1, x - 1)}[name]

    def get(self, level, name) -> torch.Tensor:
        t = empty(self.field_shape(level, name), np.float64)
        check(lib.ebg_mg2d_copy_field(self.ctx.handle, level, MG_FIELDS[name], ptr(t), None),
              "ebg_mg2d_copy_field")
        return t

    def set(self, level, name, t: torch.Tensor) -> None:
        check(lib.ebg_mg2d_copy_field(self.ctx.handle, level, MG_FIELDS[name], None, ptr(t)),
              "ebg_mg2d_copy_field")

    def restrict_viscosity(self):
        check(lib.ebg_mg2d_restrict_viscosity(self.ctx.handle), "ebg_mg2d_restrict_viscosity")

    def smoother(self, level, nsweeps, theta_s=0.9, theta_c=0.3, prnorm=0.0, bc_d=(1, 1, 1, 1)):
        d = (C.c_double * 4)(*bc_d)
        check(lib.ebg_mg2d_smoother(self.ctx.handle, level, nsweeps, theta_s, theta_c, prnorm, d),
              "ebg_mg2d_smoother")

    def smoother_tiled(self, level, nsweeps, local_sweeps=1, tile=0, theta_s=0.9, theta_c=0.3,
                       prnorm=0.0, bc_d=(1, 1, 1, 1)):
        """The smoother on LDS tiles; tile=0 keeps the whole level in one tile."""
        d = (C.c_double * 4)(*bc_d)
        check(lib.ebg_mg2d_smoother_tiled(self.ctx.handle, level, nsweeps, local_sweeps, tile,
                                          theta_s, theta_c, prnorm, d), "ebg_mg2d_smoother_tiled")

    def vcycles(self, ncycles, finest_sweeps=2, multiplier=2, cap=48, theta_s=0.9, theta_c=0.3,
                relax_v=1.0, relax_p=1.0, prnorm=0.0, bc_d=(1, 1, 1, 1)):
        d = (C.c_double * 4)(*bc_d)
        check(lib.ebg_mg2d_vcycles(self.ctx.handle, ncycles, finest_sweeps, multiplier, cap,
                                   theta_s, theta_c, relax_v, relax_p, prnorm, d),
              "ebg_mg2d_vcycles")

    def direct_solve(self, level, bc_d=(1, 1, 1, 1)):
        """Exact solve of the level's equations (gauge p(1,1) = 0) into vx, vy, pr."""
        d = (C.c_double * 4)(*bc_d)
        check(lib.ebg_mg2d_direct_solve(self.ctx.handle, level, d), "ebg_mg2d_direct_solve")

    def run(self, nvcycles, convergence_criterion, nviscosity_jumps=12.0, nvcycles_viscosity_jump=15,
            stokesscale=1.0, continscale=1.0, finest_sweeps=2, multiplier=2, cap=48, theta_s=0.9,
            theta_c=0.3, relax_v=1.0, relax_p=1.0, pressure_bc=0.0, bc_d=(1, 1, 1, 1)):
        """MultigridSolver.run_multigrid_solver (MultigridSolver.jl:113-135,
        use_multimulti = false): viscosity continuation, V-cycles, mean
        principal residuals and the reference's stopping test.  Level 0 must
        hold the viscosity, right-hand sides and initial solution.  Returns
        (cycles run, converged, history[cycles, 3] of log10 RMS residuals
        x, y, continuity)."""
        prm = MGRunParams(nvcycles=nvcycles, nvcycles_viscosity_jump=nvcycles_viscosity_jump,
                          finest_sweeps=finest_sweeps, sweep_multiplier=multiplier, sweep_cap=cap,
                          nviscosity_jumps=nviscosity_jumps,
                          convergence_criterion=convergence_criterion, resmax=0.0,
                          stokesscale=stokesscale, continscale=continscale, theta_s=theta_s,
                          theta_c=theta_c, relax_v=relax_v, relax_p=relax_p, pressure_bc=pressure_bc)
        d = (C.c_double * 4)(*bc_d)
        hist = np.zeros((max(nvcycles, 1), 3))
        done, conv = C.c_int(), C.c_int()
        check(lib.ebg_mg2d_run(self.ctx.handle, C.byref(prm), d, C.byref(done), C.byref(conv),
                               hist.ctypes.data), "ebg_mg2d_run")
        return done.value, bool(conv.value), hist[:done.value].copy()

    def run_multimulti(self, nvcycles, convergence_criterion, nviscosity_jumps=5.9999, nvcycles_viscosity_jump=3,
                       stokesscale=1.0, continscale=1.0, finest_sweeps=2, multiplier=2, cap=48, theta_s=0.9,
                       theta_c=0.3, relax_v=1.0, relax_p=1.0, pressure_bc=0.0, bc_d=(1, 1, 1, 1)):
        """MultigridSolver.run_multigrid_solver with use_multimulti = true
        (the reference's run_stokes_sinker_2d_multimulti.jl driver): level 0
        holds the global solution, level 1 solves for corrections.  Level 0
        (the device level 0 here) is filled as for run(); on return it holds
        the global solution.  Returns (cycles run, converged, principal
        history[cycles, 3], global history[updates, 3])."""
        prm = MGRunParams(nvcycles=nvcycles, nvcycles_viscosity_jump=nvcycles_viscosity_jump,
                          finest_sweeps=finest_sweeps, sweep_multiplier=multiplier, sweep_cap=cap,
                          nviscosity_jumps=nviscosity_jumps,
                          convergence_criterion=convergence_criterion, resmax=0.0,
                          stokesscale=stokesscale, continscale=continscale, theta_s=theta_s,
                          theta_c=theta_c, relax_v=relax_v, relax_p=relax_p, pressure_bc=pressure_bc)
        d = (C.c_double * 4)(*bc_d)
        hist = np.zeros((max(nvcycles, 1), 3))
        cap_g = max(nvcycles, 1)
        ghist = np.zeros((cap_g, 3))
        done, conv, ng = C.c_int(), C.c_int(), C.c_int()
        check(lib.ebg_mg2d_run_multimulti(self.ctx.handle, C.byref(prm), d, C.byref(done), C.byref(conv),
                                          hist.ctypes.data, C.byref(ng), ghist.ctypes.data, cap_g),
              "ebg_mg2d_run_multimulti")
        return done.value, bool(conv.value), hist[:done.value].copy(), ghist[:ng.value].copy()


def change_norms(ctx, v, v_old=None):
    """MathTools.get_L2_norms / get_inf_norms (src/utils/MathTools.jl:64-84) over
    the concatenation of the device arrays in v (and v_old; None = zeros):
    (abs_L2, rel_L2, abs_inf, rel_inf)."""
    v = list(v) if isinstance(v, (list, tuple)) else [v]
    vo = list(v_old) if isinstance(v_old, (list, tuple)) else ([v_old] * len(v))
    if not 1 <= len(v) <= 4 or len(vo) != len(v):
        raise ValueError("1..4 segments, v and v_old of equal count")
    k = len(v)
    n = (C.c_int64 * k)(*[t.numel() for t in v])
    pv = (C.c_void_p * k)(*[t.data_ptr() for t in v])
    po = (C.c_void_p * k)(*[(t.data_ptr() if t is not None else None) for t in vo])
    for a, b in zip(v, vo):
        if b is not None and b.numel() != a.numel():
            raise ValueError("v and v_old segments differ in size")
    out = (C.c_double * 4)()
    check(lib.ebg_change_norms(ctx.handle, k, n, pv, po, out), "ebg_change_norms")
    return tuple(out)


def calculate_new_timestep(ctx, vx1, vy1, timestep_old, marker_cell_displ_max):
    """Update.calculate_new_time_step_using_average_grid_spacing
    (time_step/update/Update.jl:43-61): (timestep_new, updated)."""
    dt = C.c_double()
    up = C.c_int()
    check(lib.ebg_calculate_new_timestep(ctx.handle, ptr(vx1), ptr(vy1), float(timestep_old),
                                         float(marker_cell_displ_max), C.byref(dt), C.byref(up)),
          "ebg_calculate_new_timestep")
    return dt.value, bool(up.value)
