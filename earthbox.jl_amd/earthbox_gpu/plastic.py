"""Nodal plastic yielding (host side): one step of the reference's Picard loop
over nodes.

update_nodes_for_plastic_yielding mirrors
PlasticFailureNodes.update_nodes_for_plastic_yielding
(rheology/plastic_failure/nodes/PlasticFailureNodes.jl:24-319) on device arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import PlasticNodesIn, PlasticNodesOut, check, lib
from .context import GPUContext
from .device import empty, ptr

FIELDS = {"plastic_yield": "b", "etas1": "b", "eta_flow": "b", "mus1": "b", "cohesion_grid": "b",
          "fric_degrees_grid": "b", "sxy2": "b", "sxx2": "p", "pr1": "p"}
SCALARS = tuple(k for k, _ in PlasticNodesIn._fields_[9:])


def update_nodes_for_plastic_yielding(ctx: GPUContext, fields: dict, yield_error, params: dict,
                                      no_yielding_in_mobile_wall: bool = False,
                                      no_yielding_in_plate_extension: bool = False, out: dict | None = None,
                                      noalloc: bool = False):
    """fields: device arrays keyed like FIELDS; yield_error: device (ynum, xnum),
    updated in place; params: the remaining scalars of PlasticNodesIn (missing
    ones 0).  Returns (out dict with etavp_new, plastic_strain_rate and
    plastic_yield_new, nnode_yield, global_yield_error), as the reference's
    tuple.  noalloc: use the allocation-free entry point (its reduction
    scratch is a static device array; pass `out` to allocate nothing here
    either)."""
    g = ctx.grid
    if g is None:
        raise RuntimeError("set_grid() must be called first")
    nx, ny = g.xnum, g.ynum
    nb, npr = nx * ny, (nx - 1) * (ny - 1)
    a = PlasticNodesIn()
    for k, kind in FIELDS.items():
        setattr(a, k, ptr(fields[k], np.float64, nb if kind == "b" else npr))
    for k in SCALARS:
        setattr(a, k, params.get(k, 0))
    a.no_yielding_in_mobile_wall = int(bool(no_yielding_in_mobile_wall))
    a.no_yielding_in_plate_extension = int(bool(no_yielding_in_plate_extension))
    if out is None:
        out = {k: empty((ny, nx), np.float64) for k in ("etavp_new", "plastic_strain_rate", "plastic_yield_new")}
    o = PlasticNodesOut()
    for k in ("etavp_new", "plastic_strain_rate", "plastic_yield_new"):
        setattr(o, k, ptr(out[k], np.float64, nb))
    o.yield_error = ptr(yield_error, np.float64, nb)
    fn = "ebg_update_nodes_for_plastic_yielding" + ("_noalloc" if noalloc else "")
    check(getattr(lib, fn)(ctx.handle, C.byref(a), C.byref(o)), fn)
    return out, int(o.nnode_yield), float(o.global_yield_error)


def interpolate_basic_viscoplastic_viscosity_to_pressure_grid(ctx: GPUContext, etas1, plastics, plasticn,
                                                              plastic_yield, etan1,
                                                              iuse_harmonic_avg_normal_viscosity: int = 1):
    """Interpolation.interpolate_basic_viscoplastic_viscosity_to_pressure_grid!
    (Interpolation.jl:97-109): etan1 updated in place when the harmonic option is on."""
    if iuse_harmonic_avg_normal_viscosity != 1:
        return etan1
    g = ctx.grid
    nb, npr = g.xnum * g.ynum, (g.xnum - 1) * (g.ynum - 1)
    check(lib.ebg_harmonic_normal_viscosity(ctx.handle, ptr(etas1, np.float64, nb), ptr(plastics, np.float64, nb),
                                            ptr(plasticn, np.float64, npr), ptr(plastic_yield, np.float64, nb),
                                            ptr(etan1, np.float64, npr)), "ebg_harmonic_normal_viscosity")
    return etan1


def plastic_strain_rate_at_pressure_nodes(ctx: GPUContext, eii_plastic_basic, eii_plastic_pressure):
    """PlasticStrainRateInterpolation.interpolate_at_pressure_nodes!."""
    g = ctx.grid
    nb, npr = g.xnum * g.ynum, (g.xnum - 1) * (g.ynum - 1)
    check(lib.ebg_plastic_strain_rate_at_pressure_nodes(ctx.handle, ptr(eii_plastic_basic, np.float64, nb),
                                                        ptr(eii_plastic_pressure, np.float64, npr)),
          "ebg_plastic_strain_rate_at_pressure_nodes")
    return eii_plastic_pressure
