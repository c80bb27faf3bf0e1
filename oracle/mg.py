"""Oracle (test infrastructure only): driver of the restated reference 2-D
multigrid (liboracle.so, mg2d.c): level hierarchy, grid maps and V-cycle.

Follows multigrid/level_manager/LevelManager.jl:411-519 (levels),
multigrid_data/MultigridDataManager.jl:223-241 (smoothing counts),
vcycle/core/SmoothAndRestrict.jl:44-71 and SmoothAndProlongate.jl:60-115.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import grid2d_uniform, lib, p

_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double


class _MGLevel(C.Structure):
    _fields_ = [("xnum", _I64), ("ynum", _I64),
                ("xstp_b", _P), ("ystp_b", _P), ("xstp_vy", _P), ("ystp_vx", _P),
                ("etas", _P), ("etan", _P), ("vx", _P), ("vy", _P), ("pr", _P),
                ("RX", _P), ("RY", _P), ("RC", _P), ("rx", _P), ("ry", _P), ("rc", _P)]


class _MGMap(C.Structure):
    _fields_ = [(k, _P) for k in ("vx_iy", "vx_jx", "vy_iy", "vy_jx", "pr_iy", "pr_jx",
                                  "b_iy", "b_jx", "vx_dy", "vx_dx", "vy_dy", "vy_dx",
                                  "pr_dy", "pr_dx", "b_dy", "b_dx")]


def _sigs():
    L = lib()
    PL, PM = C.POINTER(_MGLevel), C.POINTER(_MGMap)
    table = {
        "ebo_mg_set_bc": (None, [PL, _D, _D, _D, _D]),
        "ebo_mg_sweep": (None, [PL, _D, _D]),
        "ebo_mg_normalize_pressure": (None, [PL, _D]),
        "ebo_mg_residuals": (None, [PL]),
        "ebo_mg_smoother": (None, [PL, _I64, _D, _D, _D, _P]),
        "ebo_mg_axis_map": (None, [_I64, _P, _I64, _P, _P, _P]),
        "ebo_mg_restrict": (None, [PL, PL, PM, _P, _P, _P]),
        "ebo_mg_prolong": (None, [PL, PL, PM, _D, _D]),
        "ebo_mg_viscosity_restrict": (None, [PL, PL, PM, _P, _P]),
        "ebo_mg_scale_viscosity": (None, [_I64, _P, _P, _D, _D, _D, _D]),
    }
    for k, (r, a) in table.items():
        getattr(L, k).restype = r
        getattr(L, k).argtypes = a
    return L


def F(shape):
    return np.zeros(shape, order="F")


class Level:
    def __init__(self, xnum, ynum, xsize, ysize):
        self.xnum, self.ynum = xnum, ynum
        self.g = grid2d_uniform(xnum, ynum, xsize, ysize)
        ny, nx = ynum, xnum
        self.etas, self.etan = F((ny, nx)), F((ny - 1, nx - 1))
        self.vx, self.vy, self.pr = F((ny + 1, nx)), F((ny, nx + 1)), F((ny - 1, nx - 1))
        self.RX, self.RY, self.RC = F((ny + 1, nx)), F((ny, nx + 1)), F((ny - 1, nx - 1))
        self.rx, self.ry, self.rc = F((ny + 1, nx)), F((ny, nx + 1)), F((ny - 1, nx - 1))

    def cs(self) -> _MGLevel:
        g = self.g
        return _MGLevel(self.xnum, self.ynum, p(g["xstp_b"]), p(g["ystp_b"]), p(g["xstp_vy"]),
                        p(g["ystp_vx"]), p(self.etas), p(self.etan), p(self.vx), p(self.vy),
                        p(self.pr), p(self.RX), p(self.RY), p(self.RC), p(self.rx), p(self.ry),
                        p(self.rc))

    def zero_solution(self):
        self.vx[...] = 0.0
        self.vy[...] = 0.0
        self.pr[...] = 0.0


def level_sizes(xnum, ynum, max_levels=10):
    """LevelManager.calculate_multigrid_levels_2d (:441-457)."""
    ys, xs = [ynum], [xnum]
    while min(ys[-1], xs[-1]) > 4:
        yn, xn = -(-ys[-1] // 2), -(-xs[-1] // 2)
        if min(yn, xn) < 4 or len(ys) >= max_levels:
            break
        ys.append(yn)
        xs.append(xn)
    return list(zip(xs, ys))


def axis_map(ga, gb):
    idx = np.empty(ga.size, dtype=np.int32)
    d = np.empty(ga.size)
    _sigs().ebo_mg_axis_map(ga.size, p(ga), gb.size, p(gb), p(idx), p(d))
    return idx, d


class Map:
    """fine -> coarse maps (GridMappingManager.calculate_mapping_group!)."""

    def __init__(self, f: Level, c: Level):
        gf, gc = f.g, c.g
        self.a = {}
        self.a["vx_iy"], self.a["vx_dy"] = axis_map(gf["gridy_vx"], gc["gridy_vx"])
        self.a["vx_jx"], self.a["vx_dx"] = axis_map(gf["gridx_b"], gc["gridx_b"])
        self.a["vy_iy"], self.a["vy_dy"] = axis_map(gf["gridy_b"], gc["gridy_b"])
        self.a["vy_jx"], self.a["vy_dx"] = axis_map(gf["gridx_vy"], gc["gridx_vy"])
        self.a["pr_iy"], self.a["pr_dy"] = axis_map(gf["gridy_pr"], gc["gridy_pr"])
        self.a["pr_jx"], self.a["pr_dx"] = axis_map(gf["gridx_pr"], gc["gridx_pr"])
        self.a["b_iy"], self.a["b_dy"] = axis_map(gf["gridy_b"], gc["gridy_b"])
        self.a["b_jx"], self.a["b_dx"] = axis_map(gf["gridx_b"], gc["gridx_b"])

    def cs(self) -> _MGMap:
        return _MGMap(*[p(self.a[k]) for k, _ in _MGMap._fields_])


def smoothing_counts(nlev, finest, multiplier=2, cap=None):
    s = [finest]
    for _ in range(1, nlev):
        s.append(s[-1] * multiplier)
    if cap is not None:
        s = [s[0]] + [min(v, cap) for v in s[1:]]
    return s


class Hierarchy:
    def __init__(self, xnum, ynum, xsize, ysize, nlevels=None, bc_d=(1.0, 1.0, 1.0, 1.0)):
        sizes = level_sizes(xnum, ynum)
        if nlevels is not None:
            sizes = sizes[:nlevels]
        self.levels = [Level(x, y, xsize, ysize) for x, y in sizes]
        self.maps = [Map(self.levels[k], self.levels[k + 1]) for k in range(len(sizes) - 1)]
        self.bc_d = np.asarray(bc_d, dtype=np.float64)

    def restrict_viscosity(self):
        L = _sigs()
        for k, m in enumerate(self.maps):
            f, c = self.levels[k], self.levels[k + 1]
            wts, wtn = F(c.etas.shape), F(c.etan.shape)
            L.ebo_mg_viscosity_restrict(C.byref(f.cs()), C.byref(c.cs()), C.byref(m.cs()),
                                        p(wts), p(wtn))

    def smoother(self, k, nsweeps, theta_s, theta_c, prnorm):
        _sigs().ebo_mg_smoother(C.byref(self.levels[k].cs()), nsweeps, theta_s, theta_c, prnorm,
                                p(self.bc_d))

    def vcycle(self, smoothing, theta_s=0.9, theta_c=0.3, relax_v=1.0, relax_p=1.0, prnorm=0.0):
        """One V-cycle in the reference's order; returns the finest-level
        residual arrays after the final smoothing."""
        L = _sigs()
        nl = len(self.levels)
        for n in range(nl):  # SmoothAndRestrict.jl:50-71
            lev = self.levels[n]
            if n > 0:
                lev.zero_solution()
            self.smoother(n, smoothing[n], theta_s, theta_c, prnorm)
            if n < nl - 1:
                c = self.levels[n + 1]
                wtx, wty, wtc = F(c.RX.shape), F(c.RY.shape), F(c.RC.shape)
                L.ebo_mg_restrict(C.byref(lev.cs()), C.byref(c.cs()), C.byref(self.maps[n].cs()),
                                  p(wtx), p(wty), p(wtc))
        for n in range(nl - 1, -1, -1):  # SmoothAndProlongate.jl:72-115
            if n == 0:
                self.smoother(0, smoothing[0], theta_s, theta_c, prnorm)
            else:
                self.smoother(n, smoothing[n], theta_s, theta_c, 0.0)
                L.ebo_mg_prolong(C.byref(self.levels[n].cs()), C.byref(self.levels[n - 1].cs()),
                                 C.byref(self.maps[n - 1].cs()), relax_v, relax_p)
        f = self.levels[0]
        return f.rx, f.ry, f.rc
