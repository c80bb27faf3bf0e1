"""Marker -> grid interpolation of the Stokes transport properties (host side).

interpolate_marker_stokes_parameters_to_grid mirrors the reference's
Interpolation.initialize_bilinear_interpolation! + interpolate_marker_stokes_
parameters_to_grid! (interpolation/Interpolation.jl:61-95) on device arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import GRID_FIELD_NAMES, MarkerStokesProps, StokesGridFields, check, lib
from .context import GPUContext
from .device import empty, ptr

PROP_DTYPES = {"x": np.float64, "y": np.float64, "inside_flags": np.int8, "rho": np.float64,
               "cohesion": np.float64, "fric": np.float64, "eta_flow": np.float64,
               "eta": np.float64, "sxy": np.float64, "sxx": np.float64, "meltfrac": np.float64,
               "pfailure": np.float32, "dilatation": np.float32, "matid": np.int16,
               "mat_mu": np.float64}

BASIC = ("rho1", "etas1", "eta_flow", "mus1", "sxy1", "plastics", "cohesion_grid",
         "fric_degrees_grid", "wtnodes", "wtetas")
PRESSURE = ("etan1", "mun1", "sxx1", "plasticn", "dilatation_grid", "extractable_meltfrac_grid",
            "wtetan")
VY = ("rho1_vy", "wtnodes_vy")


def grid_field_shape(name: str, xnum: int, ynum: int):
    if name in BASIC:
        return (ynum, xnum)
    if name in PRESSURE:
        return (ynum - 1, xnum - 1)
    if name in VY:
        return (ynum, xnum + 1)
    raise KeyError(name)


def new_grid_fields(xnum, ynum, fill=0.0) -> dict:
    out = {}
    for k in GRID_FIELD_NAMES:
        if k.startswith("marker_"):
            continue
        out[k] = empty(grid_field_shape(k, xnum, ynum), np.float64)
        out[k].fill_(fill)
    return out


def interpolate_marker_stokes_parameters_to_grid(ctx: GPUContext, markers: dict, fields: dict) -> None:
    """markers: device arrays keyed by PROP_DTYPES names (missing = NULL);
    fields: device grid arrays keyed by GRID_FIELD_NAMES (in/out)."""
    g = ctx.grid
    n = markers["x"].numel()
    pr = MarkerStokesProps()
    pr.marknum = n
    for k, dt in PROP_DTYPES.items():
        t = markers.get(k)
        if t is None:
            continue
        setattr(pr, k, ptr(t, dt, None if k == "mat_mu" else n))
    if markers.get("mat_mu") is not None:
        pr.nmat = markers["mat_mu"].numel()
    gf = StokesGridFields()
    for k in GRID_FIELD_NAMES:
        t = fields.get(k)
        if t is None:
            continue
        if k.startswith("marker_"):
            setattr(gf, k, ptr(t, np.int32, n))
        else:
            s = grid_field_shape(k, g.xnum, g.ynum)
            setattr(gf, k, ptr(t, np.float64, s[0] * s[1]))
    check(lib.ebg_interpolate_marker_stokes_parameters_to_grid(ctx.handle, C.byref(pr), C.byref(gf)),
          "ebg_interpolate_marker_stokes_parameters_to_grid")
