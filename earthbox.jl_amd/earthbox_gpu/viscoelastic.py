"""Viscoelastic terms before the Stokes solve and grid strain rates and
stresses after it (host side).

calc_viscoelastic_terms_and_rhs mirrors the reference's back-to-back calls
StokesViscoelasticTerms.calc_viscoelastic_terms!(model) and
StokesContinuityRhs.viscoelastic_rhs!(model) (StokesContinuitySolver.jl:189-190);
update_grid_strain_rate_and_stress mirrors the grid part of
stokes_solution_processing! (GridStrainRate.jl:33-159, GridStress.jl), on
device arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import StrainStressIn, StrainStressOut, ViscoelasticIn, ViscoelasticOut, check, lib
from .context import GPUContext
from .device import empty, ptr

IN_SHAPES = {"etas1": "b", "mus1": "b", "sxy1": "b", "rho1": "b", "etan1": "p", "mun1": "p",
             "sxx1": "p", "rho1_vy": "vy", "dilatation_grid": "p", "eii_plastic_pressure": "p"}
OUT_SHAPES = {"etas0": "b", "sxy0": "b", "etan0": "p", "sxx0": "p", "RX1": "vx", "RY1": "vy", "RC1": "p"}


def _shape(kind: str, xnum: int, ynum: int):
    return {"b": (ynum, xnum), "p": (ynum - 1, xnum - 1), "vx": (ynum + 1, xnum),
            "vy": (ynum, xnum + 1)}[kind]


def calc_viscoelastic_terms_and_rhs(ctx: GPUContext, fields: dict, timestep: float, gravity_x: float,
                                    gravity_y: float, iuse_interface_stabilization: int = 0,
                                    out: dict | None = None) -> dict:
    """fields: device arrays keyed like IN_SHAPES (rho1_vy, dilatation_grid and
    eii_plastic_pressure may be missing); returns (or fills) the outputs keyed
    like OUT_SHAPES."""
    g = ctx.grid
    if g is None:
        raise RuntimeError("set_grid() must be called first")
    nx, ny = g.xnum, g.ynum
    a = ViscoelasticIn()
    for k, kind in IN_SHAPES.items():
        t = fields.get(k)
        if t is not None:
            s = _shape(kind, nx, ny)
            setattr(a, k, ptr(t, np.float64, s[0] * s[1]))
    a.timestep = float(timestep)
    a.gravity_x = float(gravity_x)
    a.gravity_y = float(gravity_y)
    a.iuse_interface_stabilization = int(iuse_interface_stabilization)
    if out is None:
        out = {k: empty(_shape(kind, nx, ny), np.float64) for k, kind in OUT_SHAPES.items()}
    o = ViscoelasticOut()
    for k, kind in OUT_SHAPES.items():
        s = _shape(kind, nx, ny)
        setattr(o, k, ptr(out[k], np.float64, s[0] * s[1]))
    check(lib.ebg_stokes_viscoelastic_rhs(ctx.handle, C.byref(a), C.byref(o)), "ebg_stokes_viscoelastic_rhs")
    return out


STRAIN_IN = {"vx1": "vx", "vy1": "vy", "etas1": "b", "mus1": "b", "sxy1": "b", "etan1": "p", "mun1": "p",
             "sxx1": "p"}
STRAIN_OUT = {"exx": "p", "exy": "b", "eii": "p", "sxy2": "b", "sxx2": "p", "dsxy": "b", "dsxx": "p"}


def update_grid_strain_rate_and_stress(ctx: GPUContext, fields: dict, timestep: float,
                                       out: dict | None = None) -> dict:
    """fields: device arrays keyed like STRAIN_IN; returns (or fills) the
    outputs keyed like STRAIN_OUT."""
    g = ctx.grid
    if g is None:
        raise RuntimeError("set_grid() must be called first")
    nx, ny = g.xnum, g.ynum
    a = StrainStressIn()
    for k, kind in STRAIN_IN.items():
        s = _shape(kind, nx, ny)
        setattr(a, k, ptr(fields[k], np.float64, s[0] * s[1]))
    a.timestep = float(timestep)
    if out is None:
        out = {k: empty(_shape(kind, nx, ny), np.float64) for k, kind in STRAIN_OUT.items()}
    o = StrainStressOut()
    for k, kind in STRAIN_OUT.items():
        s = _shape(kind, nx, ny)
        setattr(o, k, ptr(out[k], np.float64, s[0] * s[1]))
    check(lib.ebg_grid_strain_rate_and_stress(ctx.handle, C.byref(a), C.byref(o)),
          "ebg_grid_strain_rate_and_stress")
    return out
