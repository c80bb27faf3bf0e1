"""Oracle (test infrastructure only): driver of the restated reference 3-D
multigrid (liboracle.so, mg3d.c): level hierarchy, grid maps, V-cycle and the
assembled matrix of the 3-D operator for direct solves.

Follows multigrid/level_manager/LevelManager.jl:411-435 (levels),
grid_mapping/GridMappingManager.jl:101-249 (maps),
vcycle/core/SmoothAndRestrict.jl:15-46 and SmoothAndProlongate.jl:14-60.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib, p

_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double

KB, KV, KP, KS = 0, 1, 2, 3   # axis grid kinds: basic, staggered, pressure, shear centre


class _Level(C.Structure):
    _fields_ = [("xnum", _I64), ("ynum", _I64), ("znum", _I64)] + [
        (k, _P) for k in ("xstp_b", "ystp_b", "zstp_b", "ystp_vx", "zstp_vx", "xstp_vy",
                          "etan", "etaxy", "etaxz", "etayz", "vx", "vy", "vz", "pr",
                          "RX", "RY", "RZ", "RC", "rx", "ry", "rz", "rc")]


class _Map(C.Structure):
    _fields_ = [("iy", _P * 4), ("jx", _P * 4), ("kz", _P * 4),
                ("dy", _P * 4), ("dx", _P * 4), ("dz", _P * 4)]


_done = False


def _sigs():
    global _done
    L = lib()
    if not _done:
        PL, PM = C.POINTER(_Level), C.POINTER(_Map)
        table = {
            "ebo_axis_uniform": (None, [_I64, _D] + [_P] * 7),
            "ebo_axis_from_basic": (None, [_I64] + [_P] * 7),
            "ebo_mg_axis_map": (None, [_I64, _P, _I64, _P, _P, _P]),
            "ebo_mg3_set_bc": (None, [PL]),
            "ebo_mg3_sweep": (None, [PL, _D, _D]),
            "ebo_mg3_sweep_serial": (None, [PL, _D, _D]),
            "ebo_mg3_normalize_pressure": (None, [PL, _D]),
            "ebo_mg3_residuals": (None, [PL]),
            "ebo_mg3_smoother": (None, [PL, _I64, _D, _D, _D, C.c_int]),
            "ebo_mg3_smoother_velocity": (None, [PL, _I64, _D]),
            "ebo_mg3_restrict": (None, [PL, PL, PM, _P, _P, _P, _P]),
            "ebo_mg3_prolong": (None, [PL, PL, PM, _D, _D]),
            "ebo_mg3_viscosity_restrict": (None, [PL, PL, PM, _P, _P, _P, _P]),
            "ebo_mg3_apply": (None, [PL]),
            "ebo_mg3_mean_residuals": (None, [PL, _D, _D, _P]),
            "ebo_mg_scale_viscosity": (None, [_I64, _P, _P, _D, _D, _D, _D]),
        }
        for k, (r, a) in table.items():
            getattr(L, k).restype = r
            getattr(L, k).argtypes = a
        _done = True
    return L


def F(shape):
    return np.zeros(shape, order="F")


def axis_uniform(n: int, size: float) -> dict:
    """Coordinates and steps of one axis of a uniform level grid."""
    a = {"b": np.empty(n), "stp_b": np.empty(n - 1), "v": np.empty(n + 1), "stp_v": np.empty(n),
         "p": np.empty(n - 1), "stp_p": np.empty(max(n - 2, 1)), "s": np.empty(n - 1)}
    _sigs().ebo_axis_uniform(n, size, p(a["b"]), p(a["stp_b"]), p(a["v"]), p(a["stp_v"]),
                             p(a["p"]), p(a["stp_p"]), p(a["s"]))
    return a


def axis_from_basic(g_b: np.ndarray) -> dict:
    """The same for given basic node coordinates (non-uniform axes)."""
    n = g_b.size
    a = {"b": np.ascontiguousarray(g_b, dtype=np.float64), "stp_b": np.empty(n - 1),
         "v": np.empty(n + 1), "stp_v": np.empty(n), "p": np.empty(n - 1),
         "stp_p": np.empty(max(n - 2, 1)), "s": np.empty(n - 1)}
    _sigs().ebo_axis_from_basic(n, p(a["b"]), p(a["stp_b"]), p(a["v"]), p(a["stp_v"]),
                                p(a["p"]), p(a["stp_p"]), p(a["s"]))
    return a


def axis_kinds(a: dict) -> list:
    """Node coordinates of one axis by kind (basic, staggered, pressure, shear centre)."""
    return [a["b"], a["v"], a["p"], a["s"]]


def shapes(xnum, ynum, znum) -> dict:
    ny, nx, nz = ynum, xnum, znum
    return {"vx": (ny + 1, nx, nz + 1), "vy": (ny, nx + 1, nz + 1), "vz": (ny + 1, nx + 1, nz),
            "pr": (ny - 1, nx - 1, nz - 1), "etaxy": (ny, nx, nz - 1),
            "etaxz": (ny - 1, nx, nz), "etayz": (ny, nx - 1, nz)}


class Level:
    def __init__(self, xnum, ynum, znum, xsize, ysize, zsize, axes=None):
        self.xnum, self.ynum, self.znum = xnum, ynum, znum
        if axes is None:
            axes = (axis_uniform(xnum, xsize), axis_uniform(ynum, ysize), axis_uniform(znum, zsize))
        self.ax, self.ay, self.az = axes
        sh = shapes(xnum, ynum, znum)
        self.sh = sh
        self.etan, self.etaxy = F(sh["pr"]), F(sh["etaxy"])
        self.etaxz, self.etayz = F(sh["etaxz"]), F(sh["etayz"])
        self.vx, self.vy, self.vz, self.pr = F(sh["vx"]), F(sh["vy"]), F(sh["vz"]), F(sh["pr"])
        self.RX, self.RY, self.RZ, self.RC = F(sh["vx"]), F(sh["vy"]), F(sh["vz"]), F(sh["pr"])
        self.rx, self.ry, self.rz, self.rc = F(sh["vx"]), F(sh["vy"]), F(sh["vz"]), F(sh["pr"])

    def cs(self) -> _Level:
        return _Level(self.xnum, self.ynum, self.znum, p(self.ax["stp_b"]), p(self.ay["stp_b"]),
                      p(self.az["stp_b"]), p(self.ay["stp_v"]), p(self.az["stp_v"]),
                      p(self.ax["stp_v"]), *[p(getattr(self, k)) for k in (
                          "etan", "etaxy", "etaxz", "etayz", "vx", "vy", "vz", "pr", "RX", "RY",
                          "RZ", "RC", "rx", "ry", "rz", "rc")])

    def zero_solution(self):
        for a in (self.vx, self.vy, self.vz, self.pr):
            a[...] = 0.0


def level_sizes(xnum, ynum, znum, max_levels=10):
    """LevelManager.calculate_multigrid_levels_3d (:411-435)."""
    ys, xs, zs = [ynum], [xnum], [znum]
    while min(ys[-1], xs[-1], zs[-1]) > 4:
        yn, xn, zn = -(-ys[-1] // 2), -(-xs[-1] // 2), -(-zs[-1] // 2)
        if min(yn, xn, zn) < 4 or len(ys) >= max_levels:
            break
        ys.append(yn)
        xs.append(xn)
        zs.append(zn)
    return list(zip(xs, ys, zs))


def axis_map(ga, gb):
    idx = np.empty(ga.size, dtype=np.int32)
    d = np.empty(ga.size)
    _sigs().ebo_mg_axis_map(ga.size, p(ga), gb.size, p(gb), p(idx), p(d))
    return idx, d


class Map:
    """fine -> coarse maps of all axis kinds (GridMappingManager.calculate_mapping_group!)."""

    def __init__(self, f: Level, c: Level):
        self.a = {}
        for name, af, ac in (("y", f.ay, c.ay), ("x", f.ax, c.ax), ("z", f.az, c.az)):
            for kind, (gf, gc) in enumerate(zip(axis_kinds(af), axis_kinds(ac))):
                self.a[name, kind] = axis_map(gf, gc)

    def cs(self) -> _Map:
        m = _Map()
        for kind in range(4):
            m.iy[kind] = p(self.a["y", kind][0])
            m.jx[kind] = p(self.a["x", kind][0])
            m.kz[kind] = p(self.a["z", kind][0])
            m.dy[kind] = p(self.a["y", kind][1])
            m.dx[kind] = p(self.a["x", kind][1])
            m.dz[kind] = p(self.a["z", kind][1])
        return m


def smoothing_counts(nlev, finest, multiplier=2, cap=None):
    s = [finest]
    for _ in range(1, nlev):
        s.append(s[-1] * multiplier)
    if cap is not None:
        s = [s[0]] + [min(v, cap) for v in s[1:]]
    return s


class Hierarchy:
    def __init__(self, xnum, ynum, znum, xsize, ysize, zsize, nlevels=None):
        sizes = level_sizes(xnum, ynum, znum)
        if nlevels is not None:
            sizes = sizes[:nlevels]
        self.levels = [Level(x, y, z, xsize, ysize, zsize) for x, y, z in sizes]
        self.maps = [Map(self.levels[k], self.levels[k + 1]) for k in range(len(sizes) - 1)]

    def restrict_viscosity(self):
        L = _sigs()
        for k, m in enumerate(self.maps):
            f, c = self.levels[k], self.levels[k + 1]
            w = [F(c.sh[n]) for n in ("etaxy", "etaxz", "etayz", "pr")]
            L.ebo_mg3_viscosity_restrict(C.byref(f.cs()), C.byref(c.cs()), C.byref(m.cs()),
                                         *[p(a) for a in w])

    def smoother(self, k, nsweeps, theta_s, theta_c, prnorm, serial=False):
        _sigs().ebo_mg3_smoother(C.byref(self.levels[k].cs()), nsweeps, theta_s, theta_c, prnorm,
                                 1 if serial else 0)

    def restrict(self, k):
        f, c = self.levels[k], self.levels[k + 1]
        w = [F(c.sh[n]) for n in ("vx", "vy", "vz", "pr")]
        _sigs().ebo_mg3_restrict(C.byref(f.cs()), C.byref(c.cs()), C.byref(self.maps[k].cs()),
                                 *[p(a) for a in w])

    def prolong(self, k, relax_v=1.0, relax_p=1.0):
        """Add the level-k correction to level k-1."""
        _sigs().ebo_mg3_prolong(C.byref(self.levels[k].cs()), C.byref(self.levels[k - 1].cs()),
                                C.byref(self.maps[k - 1].cs()), relax_v, relax_p)

    def vcycle(self, smoothing, theta_s=0.9, theta_c=0.3, relax_v=1.0, relax_p=1.0, prnorm=0.0,
               serial=False, velocity_only=False):
        """One V-cycle in the reference's order; returns the finest-level
        residual arrays after the final smoothing.  velocity_only: the
        momentum block alone (pressure held at its level-0 values, zero on
        coarse levels)."""
        nl = len(self.levels)

        def smooth(n, pn):
            if velocity_only:
                _sigs().ebo_mg3_smoother_velocity(C.byref(self.levels[n].cs()), smoothing[n], theta_s)
            else:
                self.smoother(n, smoothing[n], theta_s, theta_c, pn, serial)
        for n in range(nl):  # SmoothAndRestrict.jl:15-46
            if n > 0:
                self.levels[n].zero_solution()
            smooth(n, prnorm)
            if n < nl - 1:
                self.restrict(n)
        for n in range(nl - 1, -1, -1):  # SmoothAndProlongate.jl:14-60
            if n == 0:
                smooth(0, prnorm)
            else:
                smooth(n, 0.0)
                self.prolong(n, relax_v, relax_p)
        f = self.levels[0]
        return f.rx, f.ry, f.rz, f.rc

    def mean_residuals(self, stokesscale, continscale):
        out = np.zeros(4)
        _sigs().ebo_mg3_mean_residuals(C.byref(self.levels[0].cs()), stokesscale, continscale, p(out))
        return out

    def level0_residual(self, lev0, pressure_bc):
        """Smoother.calculate_residuals_without_smoothing on the global level 0
        (Smoother.jl:133-141) written into the level-1 right-hand sides
        (update_rhs_parts_on_level1_using_global_level0_residuals, :101-115)."""
        L = _sigs()
        c = lev0.cs()
        L.ebo_mg3_set_bc(C.byref(c))
        L.ebo_mg3_normalize_pressure(C.byref(c), pressure_bc)
        L.ebo_mg3_residuals(C.byref(c))
        f = self.levels[0]
        for a, b in (("RX", "rx"), ("RY", "ry"), ("RZ", "rz"), ("RC", "rc")):
            getattr(f, a)[...] = getattr(lev0, b)

    def run_multimulti(self, nvcycles, convergence_criterion, nviscosity_jumps=5.9999, nvcycles_viscosity_jump=3,
                       stokesscale=1.0, continscale=1.0, finest_sweeps=2, multiplier=2, cap=None, theta_s=0.9,
                       theta_c=0.3, relax_v=1.0, relax_p=1.0, pressure_bc=0.0):
        """MultigridSolver.run_multigrid_solver for MultigridData3d with
        use_multimulti = true: as the 2-D restatement (oracle/mg.py
        Hierarchy.run_multimulti) with the 3-D level 0 (ViscosityScaling.jl:133-168,
        217-265, 353-381) and the global-residual test (MultigridVCycle.jl:400-411,
        470-486).  Returns (cycles run, converged, principal history[cycles, 4],
        global history[updates, 4]); the global solution is left in self.level0."""
        L = _sigs()
        names = ("etan", "etaxy", "etaxz", "etayz")
        f0 = self.levels[0]
        lev0 = Level(f0.xnum, f0.ynum, f0.znum, 1.0, 1.0, 1.0, axes=(f0.ax, f0.ay, f0.az))
        for k in names + ("vx", "vy", "vz", "pr", "RX", "RY", "RZ", "RC"):
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
                    for k in ("vx", "vy", "vz", "pr"):
                        getattr(lev0, k)[...] += getattr(f0, k)
                    for n in names:
                        getattr(lev0, n)[...] = orig[0][n]
                    self.level0_residual(lev0, pressure_bc)
                    out = np.zeros(4)
                    L.ebo_mg3_mean_residuals(C.byref(lev0.cs()), stokesscale, continscale, p(out))
                    ghist.append(out)
                    f0.zero_solution()
                scale_all()
            rmax = float(np.max(np.abs(10.0 ** ghist[-1]))) if len(ghist) > 1 else 1e38
            if rmax < convergence_criterion:
                converged = True
                break
        return len(hist), converged, np.array(hist).reshape(-1, 4), np.array(ghist).reshape(-1, 4)

    def run(self, nvcycles, convergence_criterion, nviscosity_jumps=12.0, nvcycles_viscosity_jump=15,
            resmax=1e-6, stokesscale=1.0, continscale=1.0, finest_sweeps=2, multiplier=2, cap=None,
            theta_s=0.9, theta_c=0.3, relax_v=1.0, relax_p=1.0, pressure_bc=0.0):
        """MultigridSolver.run_multigrid_solver with use_multimulti = false
        (MultigridSolver.jl:113-135).  Finest-level viscosity range
        (MultigridDataManager.jl:305-373), viscosity restriction, saved
        originals, continuation (ViscosityScaling.jl:11-24, 41-60, 81-110 with
        the counters of ViscosityScaling.jl:267-309), V-cycles (MultigridVCycle.jl:79-135)
        and the convergence test (MultigridVCycle.jl:390-412, 436-452).  Returns
        (cycles run, converged, history[cycles, 4])."""
        L = _sigs()
        names = ("etan", "etaxy", "etaxz", "etayz")
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
            t = 10.0 ** res[0]
            restot = t + t + t + t   # the reference sums the x residual four times
            if (jump > nvcycles_viscosity_jump or restot < resmax) and (max_cur < max_o or min_cur > min_o):
                jump = 1
                min_cur = max(min_cur * min_factor, min_o)
                max_cur = min(max_cur * max_factor, max_o)
                scale_all()
            rmax = float(np.max(np.abs(10.0 ** res))) if ivcycle > 2 else 1e38   # NaN propagates
            if rmax < convergence_criterion and ivcycle > nvcycles_viscosity_jump * nviscosity_jumps:
                converged = True
                break
        return len(hist), converged, np.array(hist).reshape(-1, 4)


def interior_masks(lv: Level) -> dict:
    """Unknown (non-ghost) nodes of each solution array."""
    sh = lv.sh
    mx = np.zeros(sh["vx"], dtype=bool, order="F")
    mx[1:-1, 1:-1, 1:-1] = True
    my = np.zeros(sh["vy"], dtype=bool, order="F")
    my[1:-1, 1:-1, 1:-1] = True
    mz = np.zeros(sh["vz"], dtype=bool, order="F")
    mz[1:-1, 1:-1, 1:-1] = True
    mp = np.ones(sh["pr"], dtype=bool, order="F")
    return {"vx": mx, "vy": my, "vz": mz, "pr": mp}


def apply(lv: Level, x: dict) -> dict:
    """A x of the 3-D problem (mg3d.c ebo_mg3_apply); x holds vx, vy, vz, pr."""
    t = Level(lv.xnum, lv.ynum, lv.znum, 0, 0, 0, axes=(lv.ax, lv.ay, lv.az))
    for k in ("etan", "etaxy", "etaxz", "etayz"):
        setattr(t, k, getattr(lv, k))
    for k in ("vx", "vy", "vz", "pr"):
        getattr(t, k)[...] = x[k]
    _sigs().ebo_mg3_apply(C.byref(t.cs()))
    return {"vx": t.rx, "vy": t.ry, "vz": t.rz, "pr": t.rc}


def assemble(lv: Level):
    """Sparse matrix of the 3-D operator on the unknowns (interior velocities
    and all pressures), one column per unit-vector probe of ebo_mg3_apply
    (small grids only).  Returns (A, masks)."""
    import scipy.sparse as sp

    masks = interior_masks(lv)
    names = ("vx", "vy", "vz", "pr")
    rows, cols, vals = [], [], []
    n = sum(int(masks[k].sum()) for k in names)
    x = {q: F(lv.sh[q]) for q in names}
    col = 0
    for k in names:
        for idx in np.argwhere(masks[k]):
            x[k][tuple(idx)] = 1.0
            y = pack(apply(lv, x), masks)
            x[k][tuple(idx)] = 0.0
            nzr = np.nonzero(y)[0]
            rows.append(nzr)
            cols.append(np.full(nzr.size, col))
            vals.append(y[nzr])
            col += 1
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    return A, masks


def pack(x: dict, masks: dict) -> np.ndarray:
    return np.concatenate([x[k][masks[k]] for k in ("vx", "vy", "vz", "pr")])


def unpack(v: np.ndarray, lv: Level, masks: dict) -> dict:
    out, o = {}, 0
    for k in ("vx", "vy", "vz", "pr"):
        a = F(lv.sh[k])
        n = int(masks[k].sum())
        a[masks[k]] = v[o:o + n]
        o += n
        out[k] = a
    return out
