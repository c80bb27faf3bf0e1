"""Grid -> marker gathers after the Stokes solve: the marker steps of
StokesContinuitySolver.execute_post_stokes_solver_steps!
(stokes_continuity_solver/StokesContinuitySolver.jl:272-279) on the GPU.

``update_markers_after_stokes`` computes, in one pass over the markers, what
MarkerStrainRate.calculate_marker_strain_rate! (marker_exx, marker_exy),
MarkerStrainRateRatio.calculate_marker_strain_rate_ratio! (marker_sr_ratio)
and MarkerPressure.calculate_marker_pressure! (marker_pr) write; the three
named functions below run the same kernel for their own outputs.  Results are
bit-identical to the reference loops; markers outside the grid keep their
values.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import MarkerGatherIn, MarkerGatherOut, SubgridStressIn, SubgridStressOut, check, lib
from .context import GPUContext
from .device import empty, ptr

GRID_PRESSURE = ("exx", "pr1", "sxx2", "dsxx")
GRID_BASIC = ("exy", "sxy2", "dsxy")
OUTPUTS = ("marker_exx", "marker_exy", "marker_pr", "marker_sr_ratio")
STRESS = ("marker_sxx", "marker_sxy")  # updated in place, so they must be given in ``out``


def update_markers_after_stokes(ctx: GPUContext, markers: dict, fields: dict, timestep: float = 0.0,
                                out: dict | None = None, outputs=OUTPUTS,
                                subgrid_diff_coef_stress: float = 0.0) -> dict:
    """markers: device vectors x, y, marker_xn, marker_yn (int32, 1-based, as
    the scatter stores them), optional inside_flags (int8), and for the ratio
    eta, matid (int16) and mat_mu.  fields: device grid arrays exx, pr1, sxx2,
    dsxx (ynum-1, xnum-1) and exy, sxy2, dsxy (ynum, xnum), as far as the
    requested outputs need them.  out: device vectors to update (created,
    zero-filled, for the requested outputs when missing).  Asking for
    marker_sxx and marker_sxy (given in out) adds the remaining grid stress
    change dsxx, dsxy at each marker, as StressCorrection.
    apply_marker_stress_correction_for_remaining_grid_stress! does.  That is
    the whole marker stress update only when subgrid_diff_coef_stress is 0:
    with a positive coefficient the reference first runs update_subgrid_stress!
    (StokesContinuitySolver.jl:365-381), so the correction is refused."""
    if subgrid_diff_coef_stress > 0 and any(k in STRESS for k in outputs):
        raise ValueError("the marker stress correction alone is the reference's stress update only when "
                         f"subgrid_diff_coef_stress is 0 (got {subgrid_diff_coef_stress})")
    g = ctx.grid
    if g is None:
        raise RuntimeError("set_grid() must be called first")
    nx, ny = g.xnum, g.ynum
    n = int(markers["x"].numel())
    a = MarkerGatherIn()
    a.marknum = n
    for k in ("x", "y", "eta"):
        setattr(a, k, ptr(markers.get(k), np.float64, n))
    a.marker_xn = ptr(markers["marker_xn"], np.int32, n)
    a.marker_yn = ptr(markers["marker_yn"], np.int32, n)
    a.inside_flags = ptr(markers.get("inside_flags"), np.int8, n)
    a.matid = ptr(markers.get("matid"), np.int16, n)
    mat_mu = markers.get("mat_mu")
    a.mat_mu = ptr(mat_mu, np.float64)
    a.nmat = 0 if mat_mu is None else int(mat_mu.numel())
    for k in GRID_PRESSURE:
        setattr(a, k, ptr(fields.get(k), np.float64, (nx - 1) * (ny - 1)))
    for k in GRID_BASIC:
        setattr(a, k, ptr(fields.get(k), np.float64, nx * ny))
    a.timestep = float(timestep)
    if out is None:
        out = {}
    dev = markers["x"].device
    o = MarkerGatherOut()
    for k in outputs:
        if k not in OUTPUTS + STRESS:
            raise KeyError(k)
        if k in STRESS and k not in out:
            raise KeyError(f"{k} is updated in place and must be given in out")
        if k not in out:
            out[k] = torch.zeros(n, dtype=torch.float64, device=dev)
        setattr(o, k, ptr(out[k], np.float64, n))
    check(lib.ebg_marker_strain_rate_and_pressure(ctx.handle, C.byref(a), C.byref(o)),
          "ebg_marker_strain_rate_and_pressure")
    return out


def calculate_marker_strain_rate(ctx: GPUContext, markers: dict, fields: dict, out: dict | None = None) -> dict:
    """MarkerStrainRate.calculate_marker_strain_rate! (MarkerStrainRate.jl:23-69)."""
    return update_markers_after_stokes(ctx, markers, fields, 0.0, out, ("marker_exx", "marker_exy"))


def calculate_marker_strain_rate_ratio(ctx: GPUContext, markers: dict, fields: dict, timestep: float,
                                       out: dict | None = None) -> dict:
    """MarkerStrainRateRatio.calculate_marker_strain_rate_ratio!
    (MarkerStrainRateRatio.jl:71-188); the marker strain rates it reads are
    recomputed from exx and exy in-register, as the reference's preceding step
    would have stored them."""
    return update_markers_after_stokes(ctx, markers, fields, timestep, out, ("marker_sr_ratio",))


def apply_marker_stress_correction_for_remaining_grid_stress(ctx: GPUContext, markers: dict, fields: dict,
                                                             out: dict) -> dict:
    """StressCorrection.apply_marker_stress_correction_for_remaining_grid_stress!
    (subgrid_stress/stress_correction/StressCorrection.jl:27-83): out holds the
    device vectors marker_sxx and marker_sxy, updated in place."""
    return update_markers_after_stokes(ctx, markers, fields, 0.0, out, STRESS)


def calculate_marker_pressure(ctx: GPUContext, markers: dict, fields: dict, out: dict | None = None) -> dict:
    """MarkerPressure.calculate_marker_pressure! (MarkerPressure.jl:17-49)."""
    return update_markers_after_stokes(ctx, markers, fields, 0.0, out, ("marker_pr",))


def update_subgrid_stress(ctx: GPUContext, markers: dict, fields: dict, timestep: float,
                          subgrid_diff_coef_stress: float, out: dict) -> dict:
    """SubGridStress.Update.update_subgrid_stress! (subgrid_stress/update/Update.jl:55-71).
    markers: device x, y, marker_xn, marker_yn, optional inside_flags, eta,
    matid (int16), mat_mu; fields: sxy1 (ny, nx) and sxx1 (ny-1, nx-1), the
    pre-solve grid stresses, and dsxy, dsxx, updated in place to the remaining
    changes.  out: device marker_sxy, marker_sxx (updated in place); dsxyn,
    dsxxn, wtetas, wtetan are created when missing and returned in out."""
    g = ctx.grid
    if g is None:
        raise RuntimeError("set_grid() must be called first")
    nx, ny = g.xnum, g.ynum
    n = int(markers["x"].numel())
    a = SubgridStressIn()
    m = a.markers
    m.marknum = n
    m.x = ptr(markers["x"], np.float64, n)
    m.y = ptr(markers["y"], np.float64, n)
    m.inside_flags = ptr(markers.get("inside_flags"), np.int8, n)
    m.marker_xn = ptr(markers["marker_xn"], np.int32, n)
    m.marker_yn = ptr(markers["marker_yn"], np.int32, n)
    a.matid = ptr(markers["matid"], np.int16, n)
    a.mat_mu = ptr(markers["mat_mu"], np.float64)
    a.nmat = int(markers["mat_mu"].numel())
    a.eta = ptr(markers["eta"], np.float64, n)
    a.sxy1 = ptr(fields["sxy1"], np.float64, nx * ny)
    a.sxx1 = ptr(fields["sxx1"], np.float64, (nx - 1) * (ny - 1))
    a.timestep = float(timestep)
    a.subgrid_diff_coef_stress = float(subgrid_diff_coef_stress)
    dev = markers["x"].device
    for k, shape in (("dsxyn", (ny, nx)), ("dsxxn", (ny - 1, nx - 1)), ("wtetas", (ny, nx)),
                     ("wtetan", (ny - 1, nx - 1))):
        if k not in out:
            out[k] = empty(shape, np.float64, dev)
    o = SubgridStressOut()
    o.marker_sxy = ptr(out["marker_sxy"], np.float64, n)
    o.marker_sxx = ptr(out["marker_sxx"], np.float64, n)
    o.dsxy = ptr(fields["dsxy"], np.float64, nx * ny)
    o.dsxx = ptr(fields["dsxx"], np.float64, (nx - 1) * (ny - 1))
    o.dsxyn = ptr(out["dsxyn"], np.float64, nx * ny)
    o.dsxxn = ptr(out["dsxxn"], np.float64, (nx - 1) * (ny - 1))
    o.wtetas = ptr(out["wtetas"], np.float64, nx * ny)
    o.wtetan = ptr(out["wtetan"], np.float64, (nx - 1) * (ny - 1))
    check(lib.ebg_update_subgrid_stress(ctx.handle, C.byref(a), C.byref(o)), "ebg_update_subgrid_stress")
    return out


def update_marker_stress_for_grid_and_subgrid_changes(ctx: GPUContext, markers: dict, fields: dict,
                                                      timestep: float, subgrid_diff_coef_stress: float,
                                                      out: dict) -> dict:
    """StokesContinuitySolver.update_marker_stress_for_grid_and_subgrid_changes!
    (StokesContinuitySolver.jl:365-381): the sub-grid stress update when its
    coefficient is positive, then the marker stress correction for the
    remaining grid stress change."""
    if subgrid_diff_coef_stress > 0:
        update_subgrid_stress(ctx, markers, fields, timestep, subgrid_diff_coef_stress, out)
    return update_markers_after_stokes(ctx, markers, fields, 0.0, out, STRESS)
