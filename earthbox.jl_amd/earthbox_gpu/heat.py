"""The heat-equation step on device arrays (host side).

Mirrors the reference's heat path of one time step:
MarkerToHeatArrays.interpolate_markers_to_transport_arrays!
(interpolation/utils/MarkerToHeatArrays.jl:29-35) and
GridTemperatureToMarkers.interpolate! before the Stokes solve, then
HeatSolver.update_marker_temp_for_grid_and_subgrid_changes!
(heat_solver/HeatSolver.jl:77-101): right-hand-side terms, the adaptive heat
loop around the solve (HeatLoop.jl:12-78), the grid temperature change, the
sub-grid diffusion (SubgridDiffusion.jl) and the marker temperature update
(MarkerTemperature.jl), and after advection the sticky temperature
correction.  Every step runs in the HIP library through the C ABI; device
arrays are column-major (ny, nx) torch views as the reference's Julia
matrices.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import (HeatGridFields, HeatLoopIn, HeatLoopOut, HeatLoopStats, HeatRhsIn, HeatSolveOpts, HeatSystem,
                   MarkerCells, MarkerHeatProps, SolveStats, SubgridHeatIn, ThermalBC, check, lib)
from .context import GPUContext
from .device import empty, ptr

HEAT_FIELDS = ("tk1", "kt1", "rhocp1", "hr1", "ha1")


def _grid(ctx: GPUContext):
    if ctx.grid is None:
        raise RuntimeError("set_grid() must be called first")
    return ctx.grid


def _bc(ctx: GPUContext, bc: dict) -> ThermalBC:
    g = _grid(ctx)
    b = ThermalBC()
    b.btopt = ptr(bc["btopt"], np.float64, 2 * g.xnum)
    b.bbottomt = ptr(bc["bbottomt"], np.float64, 2 * g.xnum)
    b.bleftt = ptr(bc["bleftt"], np.float64, 2 * g.ynum)
    b.brightt = ptr(bc["brightt"], np.float64, 2 * g.ynum)
    return b


def _cells(ctx: GPUContext, markers: dict) -> MarkerCells:
    n = int(markers["x"].numel())
    m = MarkerCells()
    m.marknum = n
    m.x = ptr(markers["x"], np.float64, n)
    m.y = ptr(markers["y"], np.float64, n)
    m.inside_flags = ptr(markers.get("inside_flags"), np.int8, n)
    m.marker_xn = ptr(markers.get("marker_xn"), np.int32, n)
    m.marker_yn = ptr(markers.get("marker_yn"), np.int32, n)
    return m


def interpolate_marker_heat_parameters_to_grid(ctx: GPUContext, markers: dict, mats: dict, fields: dict,
                                               bc: dict | None = None) -> dict:
    """Interpolation.interpolate_marker_heat_parameters_to_grid!
    (interpolation/Interpolation.jl:77-85).  markers: device x, y, optional
    inside_flags, TK, kt, rhocp, ha, matid (int16) and optional
    serpentinization_heat_production; mats: mat_hr.  fields: device
    tk1, kt1, rhocp1, hr1, ha1 updated in place; tk1 must hold the previous
    heat solution tk2 on entry (the reference's backup of tk1 is tk2,
    BackupTransportArrays.jl:154-163).  Optional wtnodes, marker_xn,
    marker_yn receive the weights and cells.  With bc the thermal boundary
    conditions are applied to tk1 afterwards (MarkerToHeatArrays.jl:33)."""
    g = _grid(ctx)
    n = int(markers["x"].numel())
    nb = g.xnum * g.ynum
    p = MarkerHeatProps()
    p.marknum = n
    p.x = ptr(markers["x"], np.float64, n)
    p.y = ptr(markers["y"], np.float64, n)
    p.inside_flags = ptr(markers.get("inside_flags"), np.int8, n)
    for k in ("TK", "kt", "rhocp", "ha"):
        setattr(p, k, ptr(markers.get(k), np.float64, n))
    p.matid = ptr(markers.get("matid"), np.int16, n)
    mat_hr = mats.get("mat_hr")
    p.mat_hr = ptr(mat_hr, np.float64)
    p.nmat = 0 if mat_hr is None else int(mat_hr.numel())
    p.serpentinization_heat_production = ptr(markers.get("serpentinization_heat_production"), np.float64, n)
    f = HeatGridFields()
    for k in HEAT_FIELDS + ("wtnodes",):
        setattr(f, k, ptr(fields.get(k), np.float64, nb))
    f.marker_xn = ptr(fields.get("marker_xn"), np.int32, n)
    f.marker_yn = ptr(fields.get("marker_yn"), np.int32, n)
    check(lib.ebg_interpolate_marker_heat_parameters_to_grid(ctx.handle, C.byref(p), C.byref(f)),
          "ebg_interpolate_marker_heat_parameters_to_grid")
    if bc is not None and fields.get("tk1") is not None:
        apply_thermal_bc(ctx, bc, fields["tk1"])
    return fields


def apply_thermal_bc(ctx: GPUContext, bc: dict, tk1: torch.Tensor) -> torch.Tensor:
    """ApplyThermalBC.apply_thermal_bc_along_boundaries! (ApplyThermalBC.jl:20-44)."""
    g = _grid(ctx)
    b = _bc(ctx, bc)
    check(lib.ebg_apply_thermal_bc(ctx.handle, C.byref(b), ptr(tk1, np.float64, g.xnum * g.ynum)),
          "ebg_apply_thermal_bc")
    return tk1


def calculate_heat_rhs_terms_grid(ctx: GPUContext, fields: dict, gravity_x: float = 0.0, gravity_y: float = 0.0,
                                  iuse_adiabatic_heating: int = 0, iuse_shear_heating: int = 0,
                                  RT1: torch.Tensor | None = None) -> torch.Tensor:
    """HeatSolver.calculate_heat_rhs_terms_grid! (HeatSolver.jl:107-118)."""
    g = _grid(ctx)
    ny, nx = g.ynum, g.xnum
    a = HeatRhsIn()
    sizes = {"hr1": nx * ny, "ha1": nx * ny, "rho1": nx * ny, "vx1": (ny + 1) * nx, "vy1": ny * (nx + 1),
             "sxy2": nx * ny, "etas1": nx * ny, "sxx2": (nx - 1) * (ny - 1), "etan1": (nx - 1) * (ny - 1)}
    for k, sz in sizes.items():
        setattr(a, k, ptr(fields.get(k), np.float64, sz))
    a.gravity_x = float(gravity_x)
    a.gravity_y = float(gravity_y)
    a.iuse_adiabatic_heating = int(iuse_adiabatic_heating)
    a.iuse_shear_heating = int(iuse_shear_heating)
    if RT1 is None:
        RT1 = empty((ny, nx), np.float64)
    check(lib.ebg_heat_rhs(ctx.handle, C.byref(a), ptr(RT1, np.float64, nx * ny)), "ebg_heat_rhs")
    return RT1


def _system(ctx: GPUContext, RT1, tk0, rhocp1, kt1, bc: dict, timestep: float) -> HeatSystem:
    g = _grid(ctx)
    nb = g.xnum * g.ynum
    s = HeatSystem()
    s.RT1 = ptr(RT1, np.float64, nb)
    s.tk0 = ptr(tk0, np.float64, nb)
    s.rhocp1 = ptr(rhocp1, np.float64, nb)
    s.kt1 = ptr(kt1, np.float64, nb)
    s.bc = _bc(ctx, bc)
    s.timestep = float(timestep)
    return s


def _opts(restart=0, max_iterations=0, rtol=0.0, pre_sweeps=0, post_sweeps=0) -> HeatSolveOpts:
    o = HeatSolveOpts()
    o.restart, o.max_iterations, o.rtol = int(restart), int(max_iterations), float(rtol)
    o.pre_sweeps, o.post_sweeps = int(pre_sweeps), int(post_sweeps)
    return o


def heat_system_apply(ctx: GPUContext, RT1, tk0, rhocp1, kt1, bc: dict, timestep: float,
                      T: torch.Tensor | None = None):
    """The rows of HeatBuildManager.build_system_of_equations applied to T
    (L*T, or None without T) and the right-hand side R, both (ny, nx)."""
    g = _grid(ctx)
    s = _system(ctx, RT1, tk0, rhocp1, kt1, bc, timestep)
    R = empty((g.ynum, g.xnum), np.float64)
    LT = empty((g.ynum, g.xnum), np.float64) if T is not None else None
    nb = g.xnum * g.ynum
    check(lib.ebg_heat_system_apply(ctx.handle, C.byref(s), ptr(T, np.float64, nb), ptr(LT, np.float64, nb),
                                    ptr(R, np.float64, nb)), "ebg_heat_system_apply")
    return LT, R


def discretize_and_solve_conductive_heat_equation(ctx: GPUContext, RT1, tk0, rhocp1, kt1, bc: dict,
                                                  timestep: float, tk2: torch.Tensor | None = None, **opts):
    """SystemSolver.discretize_and_solve_conductive_heat_equation
    (SystemSolver.jl:32-66): tk2 and the solve statistics."""
    g = _grid(ctx)
    s = _system(ctx, RT1, tk0, rhocp1, kt1, bc, timestep)
    if tk2 is None:
        tk2 = empty((g.ynum, g.xnum), np.float64)
    st = SolveStats()
    o = _opts(**opts)
    rc = lib.ebg_heat_solve(ctx.handle, C.byref(s), ptr(tk2, np.float64, g.xnum * g.ynum), C.byref(o),
                            C.byref(st))
    check(rc, "ebg_heat_solve")
    return tk2, {"iterations": st.iterations, "restarts": st.restarts, "relres": st.relres,
                 "converged": rc == 0}


def compute_residuals_temp(ctx: GPUContext, RT1, tk0, rhocp1, kt1, bc: dict, timestep: float, tk2,
                           rest: torch.Tensor | None = None) -> torch.Tensor:
    """HeatResiduals.compute_residuals_temp! (HeatResiduals.jl:12-162)."""
    g = _grid(ctx)
    s = _system(ctx, RT1, tk0, rhocp1, kt1, bc, timestep)
    if rest is None:
        rest = empty((g.ynum, g.xnum), np.float64)
    nb = g.xnum * g.ynum
    check(lib.ebg_heat_residuals(ctx.handle, C.byref(s), ptr(tk2, np.float64, nb), ptr(rest, np.float64, nb)),
          "ebg_heat_residuals")
    return rest


def run_conductive_heat_solver_loop(ctx: GPUContext, RT1, tk1, rhocp1, kt1, bc: dict, timestep: float,
                                    max_temp_change: float, out: dict | None = None, max_substeps: int = 0,
                                    **opts):
    """HeatLoop.run_conductive_heat_solver_loop! (HeatLoop.jl:12-78) from
    tk0 = tk1 (HeatSolver.initialize_old_temperature_array!).  Returns the
    dict of device arrays tk0 (= tk2 at the end), tk2, rest and dtk1 = tk2 - tk1
    (HeatSolver.calculate_grid_temp_change!), and the loop statistics."""
    g = _grid(ctx)
    nb = g.xnum * g.ynum
    out = out if out is not None else {}
    for k in ("tk0", "tk2", "rest", "dtk1"):
        if k not in out:
            out[k] = empty((g.ynum, g.xnum), np.float64)
    out["tk0"].copy_(tk1)
    a = HeatLoopIn()
    a.RT1 = ptr(RT1, np.float64, nb)
    a.rhocp1 = ptr(rhocp1, np.float64, nb)
    a.kt1 = ptr(kt1, np.float64, nb)
    a.bc = _bc(ctx, bc)
    a.timestep = float(timestep)
    a.max_temp_change = float(max_temp_change)
    a.max_substeps = int(max_substeps)
    a.tk1 = ptr(tk1, np.float64, nb)
    o = HeatLoopOut()
    o.tk0 = ptr(out["tk0"], np.float64, nb)
    o.tk2 = ptr(out["tk2"], np.float64, nb)
    o.rest = ptr(out["rest"], np.float64, nb)
    o.dtk1 = ptr(out["dtk1"], np.float64, nb)
    st = HeatLoopStats()
    op = _opts(**opts)
    rc = lib.ebg_run_conductive_heat_solver_loop(ctx.handle, C.byref(a), C.byref(o), C.byref(op), C.byref(st))
    check(rc, "ebg_run_conductive_heat_solver_loop")
    stats = {k: getattr(st, k) for k, _ in HeatLoopStats._fields_}
    stats["converged"] = rc == 0
    return out, stats


def grid_to_marker_temperature(ctx: GPUContext, markers: dict, grid: torch.Tensor, marker_TK: torch.Tensor,
                               add: bool = False) -> torch.Tensor:
    """add False: GridTemperatureToMarkers.interpolate! (the startup
    interpolation); add True: MarkerTemperature.add_remaining_grid_change_to_marker_temperature!
    with grid = dtk1.  marker_TK is updated in place."""
    g = _grid(ctx)
    m = _cells(ctx, markers)
    check(lib.ebg_grid_to_marker_temperature(ctx.handle, C.byref(m), ptr(grid, np.float64, g.xnum * g.ynum),
                                             1 if add else 0, ptr(marker_TK, np.float64, m.marknum)),
          "ebg_grid_to_marker_temperature")
    return marker_TK


def interpolate_temperature_back_to_markers_at_startup(ctx: GPUContext, markers: dict, tk1: torch.Tensor,
                                                       marker_TK: torch.Tensor, timesum: float,
                                                       iuse_initial_temp_interp: int) -> torch.Tensor:
    """Interpolation.interpolate_temperature_back_to_markers_at_startup!
    (interpolation/Interpolation.jl:116-131)."""
    if iuse_initial_temp_interp == 1 and timesum == 0:
        grid_to_marker_temperature(ctx, markers, tk1, marker_TK)
    return marker_TK


def apply_sticky_temperature_correction(ctx: GPUContext, markers: dict, marker_TK: torch.Tensor,
                                        matid_sticky_air: int, matid_sticky_water: int,
                                        temperature_top: float) -> torch.Tensor:
    """MarkerTemperature.apply_sticky_temperature_correction! (MarkerTemperature.jl:52-72)."""
    m = _cells(ctx, markers)
    check(lib.ebg_apply_sticky_temperature_correction(ctx.handle, C.byref(m),
                                                      ptr(markers["matid"], np.int16, m.marknum),
                                                      int(matid_sticky_air), int(matid_sticky_water),
                                                      float(temperature_top), ptr(marker_TK, np.float64, m.marknum)),
          "ebg_apply_sticky_temperature_correction")
    return marker_TK


def subgrid_heat_diffusion_update(ctx: GPUContext, markers: dict, fields: dict, marker_TK: torch.Tensor,
                                  timestep: float, subgrid_diff_coef_temp: float, dtk1: torch.Tensor,
                                  dtkn: torch.Tensor | None = None, wtnodes: torch.Tensor | None = None):
    """SubgridDiffusion.subgrid_heat_diffusion_update! followed by
    remove_subgrid_effects_from_grid_changes! (SubgridDiffusion.jl:52-185):
    marker_TK and dtk1 updated in place; returns dtkn."""
    g = _grid(ctx)
    nb = g.xnum * g.ynum
    a = SubgridHeatIn()
    a.markers = _cells(ctx, markers)
    a.tk1 = ptr(fields["tk1"], np.float64, nb)
    a.kt1 = ptr(fields["kt1"], np.float64, nb)
    a.rhocp1 = ptr(fields["rhocp1"], np.float64, nb)
    a.timestep = float(timestep)
    a.subgrid_diff_coef_temp = float(subgrid_diff_coef_temp)
    if dtkn is None:
        dtkn = empty((g.ynum, g.xnum), np.float64)
    check(lib.ebg_subgrid_heat_diffusion_update(ctx.handle, C.byref(a),
                                                ptr(marker_TK, np.float64, a.markers.marknum),
                                                ptr(dtkn, np.float64, nb), ptr(dtk1, np.float64, nb),
                                                ptr(wtnodes, np.float64, nb)),
          "ebg_subgrid_heat_diffusion_update")
    return dtkn


def use_heat_solver(timestep: float, max_temp_change: float, iuse_heat: int) -> bool:
    """HeatSolver.use_heat_solver (HeatSolver.jl:120-128)."""
    return timestep > 0 and max_temp_change > 0 and iuse_heat == 1


def update_marker_temp_for_grid_and_subgrid_changes(ctx: GPUContext, markers: dict, fields: dict, bc: dict,
                                                    timestep: float, max_temp_change: float,
                                                    subgrid_diff_coef_temp: float = 0.0, iuse_heat: int = 1,
                                                    gravity_x: float = 0.0, gravity_y: float = 0.0,
                                                    iuse_adiabatic_heating: int = 0, iuse_shear_heating: int = 0,
                                                    **opts) -> dict:
    """HeatSolver.update_marker_temp_for_grid_and_subgrid_changes!
    (HeatSolver.jl:77-101).  markers: device x, y, marker_xn, marker_yn,
    optional inside_flags and TK (updated in place); fields: device tk1,
    kt1, rhocp1, hr1 (and ha1, rho1, vx1, vy1, sxy2, etas1, sxx2, etan1 for
    the optional terms).  Returns RT1, tk2, tk0, rest, dtk1, dtkn (None
    without sub-grid diffusion) and the heat-loop statistics; an empty dict
    when the heat solver is off."""
    if not use_heat_solver(timestep, max_temp_change, iuse_heat):
        return {}
    RT1 = calculate_heat_rhs_terms_grid(ctx, fields, gravity_x, gravity_y, iuse_adiabatic_heating,
                                        iuse_shear_heating)
    loop, stats = run_conductive_heat_solver_loop(ctx, RT1, fields["tk1"], fields["rhocp1"], fields["kt1"], bc,
                                                  timestep, max_temp_change, **opts)
    dtk1 = loop["dtk1"]
    dtkn = None
    if subgrid_diff_coef_temp > 0:
        dtkn = subgrid_heat_diffusion_update(ctx, markers, fields, markers["TK"], timestep, subgrid_diff_coef_temp,
                                             dtk1)
    grid_to_marker_temperature(ctx, markers, dtk1, markers["TK"], add=True)
    return {"RT1": RT1, "tk2": loop["tk2"], "tk0": loop["tk0"], "rest": loop["rest"], "dtk1": dtk1,
            "dtkn": dtkn, "stats": stats}
