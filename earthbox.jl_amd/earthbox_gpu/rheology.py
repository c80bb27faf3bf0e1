"""Marker rheology on device arrays (host side).

update_marker_rheology mirrors Rheology.update_marker_rheology! (failure
properties with strain weakening, FailureProperties.jl; composite-creep flow
viscosity, CompositeViscosity.jl) and Rheology.update_marker_viscoplastic_viscosity!
(ViscoplasticViscosity.jl:21-61), rheology/Rheology.jl:36-80.
update_marker_plastic_yielding mirrors MarkerPlasticity.update_plastic_yielding!
(rheology/plastic_failure/markers/MarkerPlasticity.jl:34-43), the marker step
of the Picard loop over markers.  calc_marker_viscoplastic_viscosity mirrors
GridViscoplasticityToMarkers.calc_marker_viscoplastic_viscosity!, and
update_marker_strain mirrors Advection.update_marker_strain!
(advection/Advection.jl:169-178, MarkersStrain.jl).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import MarkerPlasticIn, MarkerPlasticOut, MarkerRheologyIn, MarkerRheologyOut, check, lib
from .context import GPUContext
from .device import ptr

FAILURE, FLOW, VISCOPLASTIC = 1, 2, 4
MARKER_INPUTS = ("strain_plastic", "fric_ini", "melt_damage", "sxx", "sxy", "exx", "exy", "TK", "pr", "GII",
                 "sr_ratio")
SCALARS = ("timestep", "temperature_top", "viscosity_min", "viscosity_max", "powerlaw_stress_min",
           "gas_constant", "iuse_viscous_strain_soft", "vsoftfac", "itype_global", "ntimestep")


def update_marker_rheology(ctx: GPUContext, markers: dict, mats: dict, params: dict, out: dict,
                           stages: int = FAILURE | FLOW | VISCOPLASTIC) -> dict:
    """markers: device vectors keyed like MARKER_INPUTS (as the stages need
    them; melt_damage optional), matid (int16), optional inside_flags (int8);
    mats: device mat_flow_type (int64, nmat), mat_flow (nmat x 12, column
    major), mat_plastic (nmat x 6), mat_mu (nmat); params: SCALARS; out:
    device vectors cohesion, fric, eta_flow, preexp, eta (updated in place)."""
    n = int(markers["matid"].numel())
    nmat = int(mats["mat_mu"].numel())
    a = MarkerRheologyIn()
    a.marknum = n
    a.inside_flags = ptr(markers.get("inside_flags"), np.int8, n)
    a.matid = ptr(markers["matid"], np.int16, n)
    for k in MARKER_INPUTS:
        setattr(a, k, ptr(markers.get(k), np.float64, n))
    a.nmat = nmat
    a.mat_flow_type = ptr(mats.get("mat_flow_type"), np.int64, nmat)
    a.mat_flow = ptr(mats.get("mat_flow"), np.float64, nmat * 12)
    a.mat_plastic = ptr(mats.get("mat_plastic"), np.float64, nmat * 6)
    a.mat_mu = ptr(mats.get("mat_mu"), np.float64, nmat)
    for k in SCALARS:
        setattr(a, k, params.get(k, 0))
    a.stages = int(stages)
    o = MarkerRheologyOut()
    for k in ("cohesion", "fric", "eta_flow", "preexp", "eta"):
        setattr(o, k, ptr(out.get(k), np.float64, n))
    check(lib.ebg_update_marker_rheology(ctx.handle, C.byref(a), C.byref(o)), "ebg_update_marker_rheology")
    return out


PLASTIC_INPUTS = ("y", "cohesion", "fric", "exx", "exy", "sr_ratio", "pr")
PLASTIC_SCALARS = ("ntimestep", "timestep", "y_sealevel", "viscosity_min", "viscosity_max", "yield_stress_min",
                   "yield_stress_max", "iuse_fluid_pressure_for_yield", "no_yielding_in_mobile_wall")
# itype_plasticity option ids (plastic_failure/markers/options/Options.jl)
VISCOELASTIC_FORECAST, ELASTIC_FORECAST = 0, 1


def update_marker_plastic_yielding(ctx: GPUContext, markers: dict, mat_mu, params: dict,
                                   forecast: int = VISCOELASTIC_FORECAST, matid_sticky_air: int = -1,
                                   matid_mobile_wall: int = -1, matid_plate_extension: int = -1) -> None:
    """markers: device inside_flags (int8), matid (int16) and the PLASTIC_INPUTS
    vectors; sxx, sxy, eta and pfailure (float32) updated in place.  mat_mu:
    device shear moduli (nmat); params: PLASTIC_SCALARS.  Matching the
    reference, only the viscoelastic forecast (0) corrects the stresses."""
    n = int(markers["matid"].numel())
    nmat = int(mat_mu.numel())
    a = MarkerPlasticIn()
    a.marknum = n
    a.inside_flags = ptr(markers["inside_flags"], np.int8, n)
    a.matid = ptr(markers["matid"], np.int16, n)
    a.mat_mu = ptr(mat_mu, np.float64, nmat)
    a.nmat = nmat
    for k in PLASTIC_INPUTS:
        setattr(a, k, ptr(markers[k], np.float64, n))
    for k in PLASTIC_SCALARS:
        setattr(a, k, params[k])
    a.forecast = int(forecast)
    a.matid_sticky_air, a.matid_mobile_wall, a.matid_plate_extension = (
        int(matid_sticky_air), int(matid_mobile_wall), int(matid_plate_extension))
    o = MarkerPlasticOut(ptr(markers["sxx"], np.float64, n), ptr(markers["sxy"], np.float64, n),
                         ptr(markers["eta"], np.float64, n), ptr(markers["pfailure"], np.float32, n))
    check(lib.ebg_update_marker_plastic_yielding(ctx.handle, C.byref(a), C.byref(o)),
          "ebg_update_marker_plastic_yielding")


def calc_marker_viscoplastic_viscosity(ctx: GPUContext, markers: dict, plastic_yield, etas1) -> None:
    """markers: x, y, marker_xn, marker_yn, optional inside_flags, eta_flow;
    eta and pfailure (float32) updated in place."""
    n = int(markers["x"].numel())
    g = ctx.grid
    nb = g.xnum * g.ynum
    check(lib.ebg_calc_marker_viscoplastic_viscosity(
        ctx.handle, n, ptr(markers["x"], np.float64, n), ptr(markers["y"], np.float64, n),
        ptr(markers["marker_xn"], np.int32, n), ptr(markers["marker_yn"], np.int32, n),
        ptr(markers.get("inside_flags"), np.int8, n), ptr(plastic_yield, np.float64, nb),
        ptr(etas1, np.float64, nb), ptr(markers["eta_flow"], np.float64, n), ptr(markers["eta"], np.float64, n),
        ptr(markers["pfailure"], np.float32, n)), "ebg_calc_marker_viscoplastic_viscosity")


def update_marker_strain(ctx: GPUContext, markers: dict, timestep: float, plastic_healing_rate: float = 0.0) -> None:
    """markers: exx, exy, sxx, sxy, eta, eta_flow, pfailure (float32), optional
    inside_flags; GII, strain_plastic, strain_rate_plastic updated in place."""
    n = int(markers["exx"].numel())
    f = lambda k, dt=np.float64: ptr(markers[k], dt, n)
    check(lib.ebg_update_marker_strain(ctx.handle, n, ptr(markers.get("inside_flags"), np.int8, n),
                                       float(timestep), float(plastic_healing_rate), f("exx"), f("exy"),
                                       f("sxx"), f("sxy"), f("eta"), f("eta_flow"), f("pfailure", np.float32),
                                       f("GII"), f("strain_plastic"), f("strain_rate_plastic")),
          "ebg_update_marker_strain")
