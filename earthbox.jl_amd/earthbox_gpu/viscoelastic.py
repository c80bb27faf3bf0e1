"""Viscoelastic terms and right-hand sides of the Stokes system (host side).

calc_viscoelastic_terms_and_rhs mirrors the reference's back-to-back calls
StokesViscoelasticTerms.calc_viscoelastic_terms!(model) and
StokesContinuityRhs.viscoelastic_rhs!(model) (StokesContinuitySolver.jl:189-190)
on device arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import ViscoelasticIn, ViscoelasticOut, check, lib
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
