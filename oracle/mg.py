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
        "ebo_mg_mean_residuals": (None, [PL, _D, _D, _P]),
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

    def mean_residuals(self, stokesscale, continscale):
        out = np.zeros(3)
        _sigs().ebo_mg_mean_residuals(C.byref(self.levels[0].cs()), stokesscale, continscale, p(out))
        return out

    def level0_residual(self, lev0, pressure_bc):
        """Smoother.calculate_residuals_without_smoothing on the global level 0
        (Smoother.jl:142-150): ghost values, pressure normalization, residuals;
        written into the level-1 right-hand sides
        (update_rhs_parts_on_level1_using_global_level0_residuals, :117-131)."""
        L = _sigs()
        c = lev0.cs()
        L.ebo_mg_set_bc(C.byref(c), *self.bc_d)
        L.ebo_mg_normalize_pressure(C.byref(c), pressure_bc)
        L.ebo_mg_residuals(C.byref(c))
        f = self.levels[0]
        f.RX[...] = lev0.rx
        f.RY[...] = lev0.ry
        f.RC[...] = lev0.rc

    def run_multimulti(self, nvcycles, convergence_criterion, nviscosity_jumps=5.9999,
                       nvcycles_viscosity_jump=3, stokesscale=1.0, continscale=1.0, finest_sweeps=2,
                       multiplier=2, cap=None, theta_s=0.9, theta_c=0.3, relax_v=1.0, relax_p=1.0,
                       pressure_bc=0.0):
        """MultigridSolver.run_multigrid_solver for MultigridData2d with
        use_multimulti = true (MultigridSolver.jl:113-135): level 0 holds the
        global solution with the original viscosity and right-hand sides
        (MultigridDataManager.jl:376-421 puts the initial guess and right-hand
        sides there); the hierarchy from level 1 down solves for corrections
        whose right-hand side is level 0's residual.  After each V-cycle
        (MultigridVCycle.jl:197-267) the multimulti continuation
        (ViscosityScaling.jl:170-205, 308-337): the maximum viscosity grows by
        its factor every nvcycles_viscosity_jump cycles; on its second
        overstep of the full maximum, level 0 takes level 1's correction
        (:207-215), level 1's right-hand side becomes level 0's new residual,
        whose mean scaled residuals are recorded as the global residuals
        (:381-404), level 1's solution is zeroed and the ramp restarts from
        constant viscosity.  The stopping test uses the last global residuals
        once two global updates exist (MultigridVCycle.jl:424-431, 488-504).
        The caller fills level 1 (self.levels[0]) as for run(); the global
        solution is left in self.level0.  Returns (cycles run, converged,
        principal history[cycles, 3], global history[updates, 3])."""
        L = _sigs()
        names = ("etan", "etas")
        f0 = self.levels[0]
        lev0 = Level(f0.xnum, f0.ynum, 1.0, 1.0)
        lev0.g = f0.g
        for k in ("etas", "etan", "vx", "vy", "pr", "RX", "RY", "RC"):
            setattr(lev0, k, getattr(f0, k).copy(order="F"))
        self.level0 = lev0
        f0.zero_solution()
        self.level0_residual(lev0, pressure_bc)
        vmin = min(float(getattr(f0, n).min()) for n in names)
        vmax = max(float(getattr(f0, n).max()) for n in names)
        min_o, max_o = vmin, vmax
        min_cur = max_cur = max_start = vmin
        max_factor = 10.0 ** (np.log10(vmax / vmin) / nviscosity_jumps)
        self.restrict_viscosity()
        orig = [{n: getattr(lv, n).copy(order="F") for n in names} for lv in self.levels]
        omin = min(float(a.min()) for o in orig for a in o.values())

        def scale_all():
            coef = np.log(max_cur / min_cur) / np.log(max_o / min_o)
            if omin < 0 and coef != np.floor(coef):
                raise ValueError("restricted viscosity is negative on a coarse level")
            for lv, o in zip(self.levels, orig):
                for n in names:
                    L.ebo_mg_scale_viscosity(o[n].size, p(o[n]), p(getattr(lv, n)), min_o, max_o,
                                             min_cur, max_cur)
        if max_cur < max_o or min_cur > min_o:
            scale_all()
        sm = smoothing_counts(len(self.levels), finest_sweeps, multiplier, cap)
        jump, overstep, hist, ghist, converged = 1, 0, [], [], False
        for ivcycle in range(1, nvcycles + 1):
            self.vcycle(sm, theta_s, theta_c, relax_v, relax_p, pressure_bc)
            hist.append(self.mean_residuals(stokesscale, continscale))
            jump += 1
            if jump > nvcycles_viscosity_jump:
                jump = 1
                max_cur = max_cur * max_factor
                if max_cur >= max_o:
                    overstep += 1
                if max_cur > max_o:
                    max_cur = max_o
                if overstep > 1:
                    max_cur = max_start
                    overstep = 0
                    for k in ("vx", "vy", "pr"):
                        getattr(lev0, k)[...] += getattr(f0, k)
                    lev0.etas[...] = orig[0]["etas"]
                    lev0.etan[...] = orig[0]["etan"]
                    self.level0_residual(lev0, pressure_bc)
                    out = np.zeros(3)
                    L.ebo_mg_mean_residuals(C.byref(lev0.cs()), stokesscale, continscale, p(out))
                    ghist.append(out)
                    f0.zero_solution()
                scale_all()
            if len(ghist) > 1:
                rmax = float(np.max(np.abs(10.0 ** ghist[-1])))
            else:
                rmax = 1e38
            if rmax < convergence_criterion:
                converged = True
                break
        return len(hist), converged, np.array(hist).reshape(-1, 3), np.array(ghist).reshape(-1, 3)

    def run(self, nvcycles, convergence_criterion, nviscosity_jumps=12.0, nvcycles_viscosity_jump=15,
            stokesscale=1.0, continscale=1.0, finest_sweeps=2, multiplier=2, cap=None,
            theta_s=0.9, theta_c=0.3, relax_v=1.0, relax_p=1.0, pressure_bc=0.0):
        """MultigridSolver.run_multigrid_solver for MultigridData2d with
        use_multimulti = false (MultigridSolver.jl:113-135): finest-level
        viscosity range (MultigridDataManager.jl:321-373), viscosity
        restriction, saved originals, continuation (ViscosityScaling.jl:26-39,
        62-79, 112-131, 267-309: a jump every nvcycles_viscosity_jump cycles,
        no residual trigger in 2-D), V-cycles (MultigridVCycle.jl:197-267), mean
        principal residuals (CalculateMeanResiduals.jl:130-172) and the test of
        MultigridVCycle.jl:414-434, 454-468.  Returns (cycles run, converged,
        history[cycles, 3])."""
        L = _sigs()
        names = ("etan", "etas")
        f0 = self.levels[0]
        vmin = min(float(getattr(f0, n).min()) for n in names)
        vmax = max(float(getattr(f0, n).max()) for n in names)
        min_o, max_o = vmin, vmax
        min_cur = max_cur = vmin
        min_factor = 1.0
        max_factor = 10.0 ** (np.log10(vmax / vmin) / nviscosity_jumps)
        self.restrict_viscosity()
        orig = [{n: getattr(lv, n).copy(order="F") for n in names} for lv in self.levels]
        omin = min(float(a.min()) for o in orig for a in o.values())

        def scale_all():
            coef = np.log(max_cur / min_cur) / np.log(max_o / min_o)
            if omin < 0 and coef != np.floor(coef):
                # Julia's ^ raises DomainError for a negative base and a
                # fractional exponent
                raise ValueError("restricted viscosity is negative on a coarse level")
            for lv, o in zip(self.levels, orig):
                for n in names:
                    L.ebo_mg_scale_viscosity(o[n].size, p(o[n]), p(getattr(lv, n)), min_o, max_o,
                                             min_cur, max_cur)
        if max_cur < max_o or min_cur > min_o:
            scale_all()
        sm = smoothing_counts(len(self.levels), finest_sweeps, multiplier, cap)
        jump, hist, converged = 1, [], False
        for ivcycle in range(1, nvcycles + 1):
            self.vcycle(sm, theta_s, theta_c, relax_v, relax_p, pressure_bc)
            res = self.mean_residuals(stokesscale, continscale)
            hist.append(res)
            jump += 1
            if jump > nvcycles_viscosity_jump and (max_cur < max_o or min_cur > min_o):
                jump = 1
                min_cur = max(min_cur * min_factor, min_o)
                max_cur = min(max_cur * max_factor, max_o)
                scale_all()
            rmax = float(np.max(np.abs(10.0 ** res))) if ivcycle > 2 else 1e38   # NaN propagates
            if rmax < convergence_criterion and ivcycle > nvcycles_viscosity_jump * nviscosity_jumps:
                converged = True
                break
        return len(hist), converged, np.array(hist).reshape(-1, 3)
