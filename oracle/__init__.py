"""CPU oracle for the EarthBox hot path -- TEST INFRASTRUCTURE ONLY.

A plain-C restatement (liboracle.so, built by oracle/Makefile) of the reference
Julia loops plus a scipy direct solve of the restated assembled system.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and
only as the checker or the timed CPU baseline -- never as a product path.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double

_SIGS = {
    "ebo_grid2d_uniform": (None, [_I64, _I64, _D, _D] + [_P] * 14),
    "ebo_index_of_left_node": (C.c_int32, [_D, _P, _I64]),
    "ebo_inside_flags": (None, [_I64, _P, _P, _D, _D, _D, _D, _P]),
    "ebo_cell_indices_basic": (None, [_I64, _P, _P, _P, _P, _I64, _P, _I64, _P, _P]),
    "ebo_rk4_velocity_and_spin": (None, [_I64, C.c_int, _D, _I64, _I64] + [_P] * 19),
    "ebo_update_marker_location": (None, [_I64, _D, _P, _P, _P, _P, _P]),
    "ebo_rotate_marker_stress": (None, [_I64, _D, _P, _P, _P, _P]),
    "ebo_calculate_spin": (None, [_I64, _I64, _P, _P, _P, _P, _P]),
    "ebo_set_threads": (None, [C.c_int]),
}

_lib = None


def build() -> None:
    """Compile liboracle.so from the C sources in this directory."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def p(a: np.ndarray | None):
    """Pointer of a numpy array (Fortran order for matrices), or NULL."""
    if a is None:
        return None
    if a.ndim >= 2 and not a.flags.f_contiguous:
        raise ValueError("matrices must be Fortran-ordered (Julia layout)")
    if a.ndim == 1 and not a.flags.c_contiguous:
        raise ValueError("vectors must be contiguous")
    return a.ctypes.data


def grid2d_uniform(xnum, ynum, xsize, ysize) -> dict:
    out = {
        "gridx_b": np.empty(xnum), "gridy_b": np.empty(ynum),
        "xstp_b": np.empty(xnum - 1), "ystp_b": np.empty(ynum - 1),
        "gridx_vy": np.empty(xnum + 1), "xstp_vy": np.empty(xnum),
        "gridy_vx": np.empty(ynum + 1), "ystp_vx": np.empty(ynum),
        "gridx_pr": np.empty(xnum - 1), "gridy_pr": np.empty(ynum - 1),
        "xstp_pr": np.empty(xnum - 2), "ystp_pr": np.empty(ynum - 2),
    }
    xa, ya = C.c_double(), C.c_double()
    lib().ebo_grid2d_uniform(xnum, ynum, xsize, ysize, C.addressof(xa), C.addressof(ya),
                             *[p(out[k]) for k in ("gridx_b", "gridy_b", "xstp_b", "ystp_b",
                                                   "gridx_vy", "xstp_vy", "gridy_vx", "ystp_vx",
                                                   "gridx_pr", "gridy_pr", "xstp_pr", "ystp_pr")])
    out["xstpavg"], out["ystpavg"] = xa.value, ya.value
    return out


def index_of_left_node(x: float, grid: np.ndarray) -> int:
    return int(lib().ebo_index_of_left_node(float(x), p(grid), grid.size))


def inside_flags(x, y, limits) -> np.ndarray:
    f = np.empty(x.size, dtype=np.int8)
    lib().ebo_inside_flags(x.size, p(x), p(y), *limits, p(f))
    return f


def cell_indices_basic(x, y, flags, g) -> tuple[np.ndarray, np.ndarray]:
    xn = np.zeros(x.size, dtype=np.int32)
    yn = np.zeros(x.size, dtype=np.int32)
    lib().ebo_cell_indices_basic(x.size, p(x), p(y), p(flags), p(g.gridx_b), g.xnum,
                                 p(g.gridy_b), g.ynum, p(xn), p(yn))
    return xn, yn


def rk4_velocity_and_spin(order, timestep, g, x, y, xn, yn, flags, vx1, vy1, esp):
    n = x.size
    mvx, mvy, msp = np.empty(n), np.empty(n), np.empty(n)
    lib().ebo_rk4_velocity_and_spin(
        n, order, timestep, g.xnum, g.ynum, p(x), p(y), p(xn), p(yn), p(flags),
        p(g.gridx_b), p(g.gridy_b), p(g.gridx_vy), p(g.gridy_vx), p(g.xstp_b), p(g.ystp_b),
        p(g.xstp_vy), p(g.ystp_vx), p(vx1), p(vy1), p(esp), p(mvx), p(mvy), p(msp))
    return mvx, mvy, msp


def advect_markers_and_rotate_stress_tensor(order, timestep, g, x, y, xn, yn, flags,
                                            vx1, vy1, esp, sxx=None, sxy=None):
    """Advection.jl:109-114 on copies: returns (x, y, vx, vy, spin, sxx, sxy)."""
    mvx, mvy, msp = rk4_velocity_and_spin(order, timestep, g, x, y, xn, yn, flags, vx1, vy1, esp)
    x2, y2 = x.copy(), y.copy()
    lib().ebo_update_marker_location(x.size, timestep, p(flags), p(mvx), p(mvy), p(x2), p(y2))
    if sxx is not None:
        sxx, sxy = sxx.copy(), sxy.copy()
        lib().ebo_rotate_marker_stress(x.size, timestep, p(flags), p(msp), p(sxx), p(sxy))
    return x2, y2, mvx, mvy, msp, sxx, sxy


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's marker loops (RK4, location update)."""
    lib().ebo_set_threads(int(n))


def calculate_spin(g, vx1, vy1) -> np.ndarray:
    esp = np.zeros((g.ynum, g.xnum), order="F")
    lib().ebo_calculate_spin(g.xnum, g.ynum, p(g.xstp_vy), p(g.ystp_vx), p(vx1), p(vy1), p(esp))
    return esp


# ---------------------------------------------------------------- Stokes system
class _StokesIn(C.Structure):
    _fields_ = [("xnum", _I64), ("ynum", _I64),
                ("xstp", _P), ("ystp", _P), ("xstpc", _P), ("ystpc", _P),
                ("xstpavg", _D), ("ystpavg", _D),
                ("etan", _P), ("etas", _P), ("rho1_vy", _P),
                ("RX", _P), ("RY", _P), ("RC", _P),
                ("btopx", _P), ("btopy", _P), ("bbottomx", _P), ("bbottomy", _P),
                ("bleftx", _P), ("blefty", _P), ("brightx", _P), ("brighty", _P),
                ("bintern_zone", _P), ("bintern_velocity", _P),
                ("pressure_bc_mode", _I64), ("pressure_bc", _D),
                ("pscale", _D), ("timestep", _D), ("gravity_y", _D),
                ("iuse_interface_stabilization", _I64)]


def _stokes_in(prob):
    g, b = prob.grid, prob.bcs
    keep = [np.ascontiguousarray(b.bintern_zone, dtype=np.int64),
            np.ascontiguousarray(b.bintern_velocity, dtype=np.float64)]
    s = _StokesIn(g.xnum, g.ynum, p(g.xstp_b), p(g.ystp_b), p(g.xstp_vy), p(g.ystp_vx),
                  g.xstpavg, g.ystpavg, p(prob.etan0), p(prob.etas0), p(prob.rho1_vy),
                  p(prob.RX1), p(prob.RY1), p(prob.RC1),
                  p(b.btopx), p(b.btopy), p(b.bbottomx), p(b.bbottomy),
                  p(b.bleftx), p(b.blefty), p(b.brightx), p(b.brighty),
                  p(keep[0]), p(keep[1]), prob.pressure_bc_mode, prob.pressure_bc,
                  prob.pscale, prob.timestep, prob.gravity_y, prob.iuse_interface_stabilization)
    return s, keep


def stokes_build(prob):
    """K8: COO (Li, Lj 1-based int64; Lv) and R exactly as the reference builds them."""
    L = lib()
    L.ebo_stokes_build.restype = _I64
    L.ebo_stokes_build.argtypes = [C.POINTER(_StokesIn), _P, _P, _P, _P]
    s, keep = _stokes_in(prob)
    N = prob.N
    cap = 9 * N
    Li = np.empty(cap, dtype=np.int64)
    Lj = np.empty(cap, dtype=np.int64)
    Lv = np.empty(cap)
    R = np.empty(N)
    nnz = L.ebo_stokes_build(C.byref(s), p(Li), p(Lj), p(Lv), p(R))
    assert nnz <= cap
    return Li[:nnz].copy(), Lj[:nnz].copy(), Lv[:nnz].copy(), R


class _RhsIn(C.Structure):
    _fields_ = [("xnum", _I64), ("ynum", _I64), ("xstp_b", _P), ("ystp_b", _P), ("xstp_vy", _P),
                ("ystp_vx", _P), ("etas1", _P), ("mus1", _P), ("sxy1", _P), ("etan1", _P),
                ("mun1", _P), ("sxx1", _P), ("rho1", _P), ("rho1_vy", _P), ("dilatation", _P),
                ("eii_plastic_pressure", _P), ("timestep", C.c_double), ("gravity_x", C.c_double),
                ("gravity_y", C.c_double), ("iuse_interface_stabilization", _I64)]


def stokes_viscoelastic_rhs(g, f: dict, timestep, gravity_x, gravity_y, iuse_interface_stabilization):
    """K7 (StokesViscoelasticTerms.jl:26-108 then StokesContinuityRhs.jl:30-190):
    f holds etas1, mus1, sxy1, rho1 (ynum, xnum), etan1, mun1, sxx1 (ynum-1,
    xnum-1), rho1_vy (ynum, xnum+1) and optionally dilatation_grid and
    eii_plastic_pressure (ynum-1, xnum-1).  Returns etas0, sxy0, etan0, sxx0,
    RX1, RY1, RC1 (Fortran order)."""
    L = lib()
    L.ebo_stokes_rhs.restype = None
    L.ebo_stokes_rhs.argtypes = [C.POINTER(_RhsIn)] + [_P] * 7
    a = {k: (np.asfortranarray(v, dtype=np.float64) if v is not None else None) for k, v in f.items()}
    ax = [np.ascontiguousarray(v, dtype=np.float64) for v in (g.xstp_b, g.ystp_b, g.xstp_vy, g.ystp_vx)]
    r = _RhsIn(g.xnum, g.ynum, *[p(v) for v in ax], p(a["etas1"]), p(a["mus1"]), p(a["sxy1"]),
               p(a["etan1"]), p(a["mun1"]), p(a["sxx1"]), p(a["rho1"]), p(a.get("rho1_vy")),
               p(a.get("dilatation_grid")), p(a.get("eii_plastic_pressure")), float(timestep),
               float(gravity_x), float(gravity_y), int(iuse_interface_stabilization))
    ny, nx = g.ynum, g.xnum
    out = {"etas0": np.zeros((ny, nx), order="F"), "sxy0": np.zeros((ny, nx), order="F"),
           "etan0": np.zeros((ny - 1, nx - 1), order="F"), "sxx0": np.zeros((ny - 1, nx - 1), order="F"),
           "RX1": np.zeros((ny + 1, nx), order="F"), "RY1": np.zeros((ny, nx + 1), order="F"),
           "RC1": np.zeros((ny - 1, nx - 1), order="F")}
    L.ebo_stokes_rhs(C.byref(r), *[p(out[k]) for k in ("etas0", "sxy0", "etan0", "sxx0", "RX1", "RY1", "RC1")])
    return out


class _PlasticIn(C.Structure):
    _fields_ = [("xnum", _I64), ("ynum", _I64), ("gridx_b", _P), ("gridy_b", _P),
                ("plastic_yield", _P), ("etas1", _P), ("eta_flow", _P), ("mus1", _P),
                ("cohesion_grid", _P), ("fric_degrees_grid", _P), ("sxy2", _P), ("sxx2", _P), ("pr1", _P),
                ("timestep", C.c_double), ("timesum", C.c_double), ("ntimestep", _I64),
                ("viscosity_min", C.c_double), ("viscosity_max", C.c_double),
                ("iuse_fluid_pressure_for_yield", _I64), ("y_sealevel", C.c_double),
                ("no_yielding_in_mobile_wall", _I64), ("no_yielding_in_plate_extension", _I64),
                ("velocity_internal_x", C.c_double), ("x_left_mobile_wall", C.c_double),
                ("x_right_mobile_wall", C.c_double), ("y_top_mobile_wall", C.c_double),
                ("y_bottom_mobile_wall", C.c_double), ("plate_extension_thickness", C.c_double),
                ("plate_extension_width", C.c_double)]


PLASTIC_FIELDS = ("plastic_yield", "etas1", "eta_flow", "mus1", "cohesion_grid", "fric_degrees_grid", "sxy2",
                  "sxx2", "pr1")


def update_nodes_for_plastic_yielding(g, f: dict, par: dict, yield_error):
    """PlasticFailureNodes.update_nodes_for_plastic_yielding
    (rheology/plastic_failure/nodes/PlasticFailureNodes.jl:24-319).  f holds the
    PLASTIC_FIELDS grid arrays, par the scalars of _PlasticIn (missing ones 0).
    Returns (etavp_new, plastic_strain_rate, plastic_yield_new, yield_error,
    nnode_yield, global_yield_error); yield_error is updated in a copy."""
    L = lib()
    L.ebo_update_nodes_for_plastic_yielding.restype = None
    L.ebo_update_nodes_for_plastic_yielding.argtypes = [C.POINTER(_PlasticIn)] + [_P] * 4 + [
        C.POINTER(_I64), C.POINTER(C.c_double)]
    a = {k: np.asfortranarray(f[k], dtype=np.float64) for k in PLASTIC_FIELDS}
    gx = np.ascontiguousarray(g.gridx_b, dtype=np.float64)
    gy = np.ascontiguousarray(g.gridy_b, dtype=np.float64)
    s = _PlasticIn(g.xnum, g.ynum, p(gx), p(gy), *[p(a[k]) for k in PLASTIC_FIELDS])
    for k, _ in _PlasticIn._fields_[13:]:
        setattr(s, k, par.get(k, 0))
    ny, nx = g.ynum, g.xnum
    ev = np.zeros((ny, nx), order="F")
    psr = np.zeros((ny, nx), order="F")
    pyn = np.zeros((ny, nx), order="F")
    ye = np.asfortranarray(np.array(yield_error, dtype=np.float64, copy=True))
    n = _I64(0)
    err = C.c_double(0.0)
    L.ebo_update_nodes_for_plastic_yielding(C.byref(s), p(ev), p(psr), p(pyn), p(ye), C.byref(n), C.byref(err))
    return ev, psr, pyn, ye, int(n.value), float(err.value)


class _StrainIn(C.Structure):
    _fields_ = [("xnum", _I64), ("ynum", _I64), ("xstp_b", _P), ("ystp_b", _P), ("xstp_vy", _P),
                ("ystp_vx", _P), ("vx1", _P), ("vy1", _P), ("etas1", _P), ("mus1", _P), ("sxy1", _P),
                ("etan1", _P), ("mun1", _P), ("sxx1", _P), ("timestep", C.c_double)]


def grid_strain_rate_and_stress(g, vx1, vy1, f: dict, timestep):
    """GridStrainRate.jl:33-159 (exx, exy, eii), GridStress.forecast_viscoelastic_grid_stress!
    (sxy2, sxx2) and calculate_deviatoric_grid_stress_change! (dsxy, dsxx).  f
    holds etas1, mus1, sxy1 (ynum, xnum) and etan1, mun1, sxx1 (ynum-1, xnum-1)."""
    L = lib()
    L.ebo_grid_strain_rate_and_stress.restype = None
    L.ebo_grid_strain_rate_and_stress.argtypes = [C.POINTER(_StrainIn)] + [_P] * 7
    ax = [np.ascontiguousarray(v, dtype=np.float64) for v in (g.xstp_b, g.ystp_b, g.xstp_vy, g.ystp_vx)]
    a = {k: np.asfortranarray(f[k], dtype=np.float64) for k in ("etas1", "mus1", "sxy1", "etan1", "mun1", "sxx1")}
    v1 = np.asfortranarray(vx1, dtype=np.float64)
    v2 = np.asfortranarray(vy1, dtype=np.float64)
    s = _StrainIn(g.xnum, g.ynum, *[p(v) for v in ax], p(v1), p(v2), *[p(a[k]) for k in
                  ("etas1", "mus1", "sxy1", "etan1", "mun1", "sxx1")], float(timestep))
    ny, nx = g.ynum, g.xnum
    out = {"exx": np.zeros((ny - 1, nx - 1), order="F"), "exy": np.zeros((ny, nx), order="F"),
           "eii": np.zeros((ny - 1, nx - 1), order="F"), "sxy2": np.zeros((ny, nx), order="F"),
           "sxx2": np.zeros((ny - 1, nx - 1), order="F"), "dsxy": np.zeros((ny, nx), order="F"),
           "dsxx": np.zeros((ny - 1, nx - 1), order="F")}
    L.ebo_grid_strain_rate_and_stress(C.byref(s), *[p(out[k]) for k in ("exx", "exy", "eii", "sxy2", "sxx2",
                                                                      "dsxy", "dsxx")])
    return out


class _GatherIn(C.Structure):
    _fields_ = ([("xnum", _I64), ("ynum", _I64), ("marknum", _I64)] +
                [(k, _P) for k in ("gridx_b", "gridy_b", "xstp_b", "ystp_b", "gridx_vy", "xstp_vy",
                                   "gridy_vx", "ystp_vx", "x", "y", "xn", "yn", "inside", "exx", "exy",
                                   "pr1", "sxx2", "sxy2", "dsxx", "dsxy", "eta", "matid", "mat_mu")] +
                [("timestep", C.c_double)])


GATHER_OUT = ("marker_dx", "marker_dy", "marker_exx", "marker_exy", "marker_sr_ratio", "marker_pr")


def marker_strain_rate_and_pressure(g, m: dict, f: dict, timestep=0.0, prev: dict | None = None):
    """MarkerGridMapping.calc_normalized_distances_from_upper_left_basic_node!,
    MarkerStrainRate.calculate_marker_strain_rate!, MarkerStrainRateRatio.
    calculate_marker_strain_rate_ratio! (when f holds sxx2) and MarkerPressure.
    calculate_marker_pressure!, in the reference's order.  m holds x, y, xn, yn
    (1-based Int32), inside (Int8) and, for the ratio, eta, matid (Int16) and
    mat_mu; f holds the grid arrays.  prev: marker arrays the loops start from
    (zeros by default); outside markers keep them.  With marker_sxx and
    marker_sxy in m, StressCorrection.apply_marker_stress_correction_for_
    remaining_grid_stress! follows and the corrected stresses are returned."""
    L = lib()
    L.ebo_marker_strain_rate_and_pressure.restype = None
    L.ebo_marker_strain_rate_and_pressure.argtypes = [C.POINTER(_GatherIn)] + [_P] * 8
    keep = []

    def arr(v, dt):
        if v is None:
            return None
        a = np.asfortranarray(v, dtype=dt)
        keep.append(a)
        return p(a)

    n = int(np.asarray(m["x"]).size)
    ax = [arr(getattr(g, k), np.float64) for k in ("gridx_b", "gridy_b", "xstp_b", "ystp_b", "gridx_vy",
                                                  "xstp_vy", "gridy_vx", "ystp_vx")]
    ratio = "sxx2" in f
    stress = "marker_sxx" in m
    s = _GatherIn(g.xnum, g.ynum, n, *ax, arr(m["x"], np.float64), arr(m["y"], np.float64),
                  arr(m["xn"], np.int32), arr(m["yn"], np.int32), arr(m["inside"], np.int8),
                  arr(f["exx"], np.float64), arr(f["exy"], np.float64), arr(f["pr1"], np.float64),
                  arr(f.get("sxx2") if ratio else None, np.float64), arr(f.get("sxy2") if ratio else None, np.float64),
                  *[arr(f.get(k) if (ratio or stress) else None, np.float64) for k in ("dsxx", "dsxy")],
                  arr(m.get("eta") if ratio else None, np.float64),
                  arr(m.get("matid") if ratio else None, np.int16),
                  arr(m.get("mat_mu") if ratio else None, np.float64), float(timestep))
    prev = prev or {}
    out = {k: np.array(prev[k], dtype=np.float64, copy=True) if k in prev else np.zeros(n) for k in GATHER_OUT}
    for k in ("marker_sxx", "marker_sxy"):
        if stress:
            out[k] = np.array(m[k], dtype=np.float64, copy=True)
    L.ebo_marker_strain_rate_and_pressure(C.byref(s), *[p(out[k]) for k in GATHER_OUT],
                                          *[p(out.get(k)) for k in ("marker_sxx", "marker_sxy")])
    return out


def harmonic_normal_viscosity(g, etas1, plastics, plasticn, plastic_yield, etan1):
    """HarmonicAverage.normal_viscosity_from_harmonic_avg_shear_viscosity!: etan1
    (ynum-1, xnum-1) updated in a copy where a neighbouring node yields."""
    L = lib()
    L.ebo_harmonic_normal_viscosity.restype = None
    L.ebo_harmonic_normal_viscosity.argtypes = [_I64, _I64] + [_P] * 5
    a = [np.asfortranarray(v, dtype=np.float64) for v in (etas1, plastics, plasticn, plastic_yield)]
    out = np.asfortranarray(np.array(etan1, dtype=np.float64, copy=True))
    L.ebo_harmonic_normal_viscosity(g.xnum, g.ynum, *[p(v) for v in a], p(out))
    return out


def plastic_strain_rate_at_pressure_nodes(g, eii_basic):
    L = lib()
    L.ebo_plastic_strain_rate_at_pressure_nodes.restype = None
    L.ebo_plastic_strain_rate_at_pressure_nodes.argtypes = [_I64, _I64, _P, _P]
    a = np.asfortranarray(eii_basic, dtype=np.float64)
    out = np.zeros((g.ynum - 1, g.xnum - 1), order="F")
    L.ebo_plastic_strain_rate_at_pressure_nodes(g.xnum, g.ynum, p(a), p(out))
    return out


def stokes_unpack(prob, S):
    """K10: (vx1, vy1, pr1) from the solution vector with ghost BCs applied."""
    L = lib()
    L.ebo_stokes_unpack.restype = None
    L.ebo_stokes_unpack.argtypes = [C.POINTER(_StokesIn), _P, _P, _P, _P]
    g = prob.grid
    s, keep = _stokes_in(prob)
    vx1 = np.zeros((g.ynum + 1, g.xnum), order="F")
    vy1 = np.zeros((g.ynum, g.xnum + 1), order="F")
    pr1 = np.zeros((g.ynum - 1, g.xnum - 1), order="F")
    S = np.ascontiguousarray(S, dtype=np.float64)
    L.ebo_stokes_unpack(C.byref(s), p(S), p(vx1), p(vy1), p(pr1))
    return vx1, vy1, pr1


def coo_spmv(N, Li, Lj, Lv, x):
    L = lib()
    L.ebo_coo_spmv.restype = None
    L.ebo_coo_spmv.argtypes = [_I64, _I64, _P, _P, _P, _P, _P]
    y = np.empty(N)
    x = np.ascontiguousarray(x, dtype=np.float64)
    L.ebo_coo_spmv(N, Li.size, p(Li), p(Lj), p(Lv), p(x), p(y))
    return y


def coo_residual(N, Li, Lj, Lv, x, R):
    """K11: |L*x - R| in COO triplet order (StokesResiduals.jl:63-82)."""
    L = lib()
    L.ebo_coo_residual.restype = None
    L.ebo_coo_residual.argtypes = [_I64, _I64, _P, _P, _P, _P, _P, _P]
    res = np.empty(N)
    L.ebo_coo_residual(N, Li.size, p(Li), p(Lj), p(Lv), p(np.ascontiguousarray(x)), p(R), p(res))
    return res


def direct_solve(N, Li, Lj, Lv, R, refine: int = 3):
    """FP64 sparse LU (scipy SuperLU) standing in for MUMPS
    (SystemSolver.jl:67-91), plus FP64 iterative refinement steps."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    A = sp.csc_matrix((Lv, (Li - 1, Lj - 1)), shape=(N, N))
    lu = spla.splu(A, permc_spec="COLAMD", options={"SymmetricMode": False})
    S = lu.solve(R)
    for _ in range(refine):
        r = R - A @ S
        S = S + lu.solve(r)
    return S


# ---------------------------------------------------------------- scatter
class _ScatterIn(C.Structure):
    _fields_ = [("marknum", _I64), ("xnum", _I64), ("ynum", _I64), ("x", _P), ("y", _P),
                ("flags", _P), ("gridx_b", _P), ("gridy_b", _P), ("xstp_b", _P), ("ystp_b", _P),
                ("gridx_vy", _P), ("xstp_vy", _P), ("rho", _P), ("cohesion", _P), ("fric", _P),
                ("eta_flow", _P), ("eta", _P), ("sxy", _P), ("sxx", _P), ("meltfrac", _P),
                ("pfailure", _P), ("dilatation", _P), ("matid", _P), ("mat_mu", _P)]


_SC_OUT = ("rho1", "rho1_vy", "etas1", "etan1", "eta_flow", "mus1", "mun1", "sxy1", "sxx1",
           "plastics", "plasticn", "cohesion", "fric_deg", "dilatation", "meltfrac",
           "wtnodes", "wtetas", "wtetan", "wtnodes_vy", "xn", "yn")


class _ScatterOut(C.Structure):
    _fields_ = [(k, _P) for k in _SC_OUT]


def scatter_stokes(g, m: dict, prev: dict):
    """Serial restatement of the Stokes scatter.  m: marker arrays (numpy);
    prev: previous grid values keyed like the GPU field names.  Returns the
    new fields (same keys) plus weights and indices."""
    L = lib()
    L.ebo_scatter_stokes.restype = C.c_int
    L.ebo_scatter_stokes.argtypes = [C.POINTER(_ScatterIn), C.POINTER(_ScatterOut)]
    n = m["x"].size
    flags = m.get("inside_flags")
    if flags is None:
        flags = inside_flags(m["x"], m["y"], g.limits)
    zeros = {k: np.zeros(n) for k in ("rho", "cohesion", "fric", "eta_flow", "eta", "sxy", "sxx", "meltfrac")}
    arr = {k: np.ascontiguousarray(m.get(k, zeros.get(k)), dtype=np.float64) for k in zeros}
    pf = np.ascontiguousarray(m.get("pfailure", np.zeros(n, np.float32)), dtype=np.float32)
    dl = np.ascontiguousarray(m.get("dilatation", np.zeros(n, np.float32)), dtype=np.float32)
    mid = np.ascontiguousarray(m.get("matid", np.ones(n, np.int16)), dtype=np.int16)
    mmu = np.ascontiguousarray(m.get("mat_mu", np.ones(1)), dtype=np.float64)
    sin = _ScatterIn(n, g.xnum, g.ynum, p(m["x"]), p(m["y"]), p(flags), p(g.gridx_b), p(g.gridy_b),
                     p(g.xstp_b), p(g.ystp_b), p(g.gridx_vy), p(g.xstp_vy),
                     *[p(arr[k]) for k in ("rho", "cohesion", "fric", "eta_flow", "eta", "sxy", "sxx",
                                           "meltfrac")], p(pf), p(dl), p(mid), p(mmu))
    names = {"rho1": "rho1", "rho1_vy": "rho1_vy", "etas1": "etas1", "etan1": "etan1",
             "eta_flow": "eta_flow", "mus1": "mus1", "mun1": "mun1", "sxy1": "sxy1", "sxx1": "sxx1",
             "plastics": "plastics", "plasticn": "plasticn", "cohesion": "cohesion_grid",
             "fric_deg": "fric_degrees_grid", "dilatation": "dilatation_grid",
             "meltfrac": "extractable_meltfrac_grid", "wtnodes": "wtnodes", "wtetas": "wtetas",
             "wtetan": "wtetan", "wtnodes_vy": "wtnodes_vy"}
    ny, nx = g.ynum, g.xnum
    basic, pres, vyg = (ny, nx), (ny - 1, nx - 1), (ny, nx + 1)
    shape = {"rho1": basic, "rho1_vy": vyg, "etas1": basic, "etan1": pres, "eta_flow": basic, "mus1": basic,
             "mun1": pres, "sxy1": basic, "sxx1": pres, "plastics": basic, "plasticn": pres,
             "cohesion_grid": basic, "fric_degrees_grid": basic, "dilatation_grid": pres,
             "extractable_meltfrac_grid": pres}
    out = {}
    for ok, gk in names.items():   # fields without a previous value start (and back up) at zero
        a = prev.get(gk)
        if gk in shape:
            out[gk] = np.asfortranarray(a.copy()) if a is not None else np.zeros(shape[gk], order="F")
    shapes = {"wtnodes": basic, "wtetas": basic, "wtetan": pres, "wtnodes_vy": vyg}
    for k, s in shapes.items():
        out[k] = np.zeros(s, order="F")
    xn = np.zeros(n, np.int32)
    yn = np.zeros(n, np.int32)
    so = _ScatterOut(*[p(out[names[k]]) for k in _SC_OUT[:19]], p(xn), p(yn))
    err = L.ebo_scatter_stokes(C.byref(sin), C.byref(so))
    if err:
        raise ValueError("zero shear modulus")
    out["marker_xn"], out["marker_yn"] = xn, yn
    return out


def stokes_grid_residuals(prob, vx1, vy1, pr1):
    """compute_stokes_residuals! (StokesResiduals.jl:94-137): resx1, resy1, resc1."""
    L = lib()
    L.ebo_stokes_grid_residuals.restype = None
    L.ebo_stokes_grid_residuals.argtypes = [C.POINTER(_StokesIn), _P, _P, _P, _P, _P, _P]
    s, keep = _stokes_in(prob)
    g = prob.grid
    vx1, vy1, pr1 = (np.asfortranarray(a, dtype=np.float64) for a in (vx1, vy1, pr1))
    rx = np.zeros((g.ynum + 1, g.xnum), order="F")
    ry = np.zeros((g.ynum, g.xnum + 1), order="F")
    rc = np.zeros((g.ynum - 1, g.xnum - 1), order="F")
    L.ebo_stokes_grid_residuals(C.byref(s), p(vx1), p(vy1), p(pr1), p(rx), p(ry), p(rc))
    return rx, ry, rc


def get_L2_norms(v, v_old):
    """MathTools.get_L2_norms (src/utils/MathTools.jl:64-73)."""
    v, v_old = np.asarray(v, np.float64).ravel(order="F"), np.asarray(v_old, np.float64).ravel(order="F")
    d = np.linalg.norm(np.abs(v_old - v))
    n = np.linalg.norm(v)
    return d, (d / n if n != 0.0 else 1e38)


def get_inf_norms(v, v_old):
    """MathTools.get_inf_norms (src/utils/MathTools.jl:75-84)."""
    v, v_old = np.asarray(v, np.float64).ravel(order="F"), np.asarray(v_old, np.float64).ravel(order="F")
    d = float(np.max(np.abs(v_old - v))) if v.size else 0.0
    n = float(np.max(np.abs(v))) if v.size else 0.0
    return d, (d / n if n != 0.0 else 1e38)


def calculate_new_timestep(timestep_old, marker_cell_displ_max, xstpavg, ystpavg, vx1, vy1):
    """time_step/update/Update.jl:43-61, 71-98, 198-218: the velocity-limited
    step and whether it changed."""
    def vmax(a):
        return max(abs(float(np.max(a))), abs(float(np.min(a))))

    def limit(dt, size, v):
        if v > 0:
            dt = min(dt, marker_cell_displ_max * size / v)
        return dt
    dt = limit(timestep_old, xstpavg, vmax(vx1))
    dt = limit(dt, ystpavg, vmax(vy1))
    return dt, dt != timestep_old


# ---------------------------------------------------------------- marker rheology
class _RheoIn(C.Structure):
    _fields_ = [("marknum", _I64), ("flags", _P), ("matid", _P), ("strain_plastic", _P), ("fric_ini", _P),
                ("melt_damage", _P), ("sxx", _P), ("sxy", _P), ("exx", _P), ("exy", _P), ("TK", _P),
                ("pr", _P), ("GII", _P), ("sr_ratio", _P), ("nmat", _I64), ("mat_flow_type", _P),
                ("mat_flow", _P), ("mat_plastic", _P), ("mat_mu", _P), ("timestep", _D),
                ("temperature_top", _D), ("viscosity_min", _D), ("viscosity_max", _D),
                ("powerlaw_stress_min", _D), ("gas_constant", _D), ("iuse_viscous_strain_soft", _I64),
                ("vsoftfac", _D), ("itype_global", _I64), ("ntimestep", _I64)]


RHEO_MARKER = ("strain_plastic", "fric_ini", "sxx", "sxy", "exx", "exy", "TK", "pr", "GII", "sr_ratio")


def marker_rheology(m: dict, mats: dict, par: dict) -> dict:
    """Rheology.update_marker_rheology! (failure properties, composite flow
    viscosity) then update_marker_viscoplastic_viscosity! on copies.  m:
    marker arrays (RHEO_MARKER, matid int16, flags int8, optional
    melt_damage, eta); mats: mat_flow_type (nmat,), mat_flow (nmat, 12),
    mat_plastic (nmat, 6), mat_mu (nmat,); par: the scalars of _RheoIn."""
    L = lib()
    for name in ("ebo_failure_properties", "ebo_flow_viscosity"):
        getattr(L, name).restype = None
        getattr(L, name).argtypes = [C.POINTER(_RheoIn), _P, _P]
    L.ebo_viscoplastic_viscosity.restype = None
    L.ebo_viscoplastic_viscosity.argtypes = [C.POINTER(_RheoIn), _P, _P]
    n = m["matid"].size
    keep = {k: np.ascontiguousarray(m[k], dtype=np.float64) for k in RHEO_MARKER}
    md = m.get("melt_damage")
    md = None if md is None else np.ascontiguousarray(md, dtype=np.float64)
    flags = np.ascontiguousarray(m.get("flags", np.ones(n, np.int8)), dtype=np.int8)
    mid = np.ascontiguousarray(m["matid"], dtype=np.int16)
    mft = np.ascontiguousarray(mats["mat_flow_type"], dtype=np.int64)
    mf = np.asfortranarray(mats["mat_flow"], dtype=np.float64)
    mp = np.asfortranarray(mats["mat_plastic"], dtype=np.float64)
    mmu = np.ascontiguousarray(mats["mat_mu"], dtype=np.float64)
    s = _RheoIn(n, p(flags), p(mid), p(keep["strain_plastic"]), p(keep["fric_ini"]), p(md), p(keep["sxx"]),
                p(keep["sxy"]), p(keep["exx"]), p(keep["exy"]), p(keep["TK"]), p(keep["pr"]), p(keep["GII"]),
                p(keep["sr_ratio"]), mft.size, p(mft), p(mf), p(mp), p(mmu), par["timestep"],
                par["temperature_top"], par["viscosity_min"], par["viscosity_max"], par["powerlaw_stress_min"],
                par["gas_constant"], int(par.get("iuse_viscous_strain_soft", 0)), par.get("vsoftfac", 1.0),
                int(par.get("itype_global", 1)), int(par.get("ntimestep", 0)))
    out = {k: np.zeros(n) for k in ("cohesion", "fric", "eta_flow", "preexp")}
    for k in out:   # outside markers keep their values
        if k in m:
            out[k] = np.array(m[k], dtype=np.float64)
    out["eta"] = np.array(m.get("eta", np.zeros(n)), dtype=np.float64)
    L.ebo_failure_properties(C.byref(s), p(out["cohesion"]), p(out["fric"]))
    L.ebo_flow_viscosity(C.byref(s), p(out["eta_flow"]), p(out["preexp"]))
    L.ebo_viscoplastic_viscosity(C.byref(s), p(out["eta_flow"]), p(out["eta"]))
    return out


def marker_viscoplastic_from_grid(g, m: dict, plastic_yield, etas1):
    """GridViscoplasticityToMarkers.calc_marker_viscoplastic_viscosity! on
    copies: (eta, pfailure).  m: x, y, flags, marker_xn/yn (1-based),
    eta_flow, eta, pfailure (float32).  The distances are
    MarkerGridMapping.jl:211-214's (x - gridx_b[xn]) / xstp_b[xn]."""
    L = lib()
    L.ebo_marker_viscoplastic_from_grid.restype = None
    L.ebo_marker_viscoplastic_from_grid.argtypes = [_I64, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P]
    n = m["x"].size
    xn = np.ascontiguousarray(m["marker_xn"], dtype=np.int32)
    yn = np.ascontiguousarray(m["marker_yn"], dtype=np.int32)
    flags = np.ascontiguousarray(m["flags"], dtype=np.int8)
    ix = np.clip(xn, 1, g.xnum - 1) - 1
    iy = np.clip(yn, 1, g.ynum - 1) - 1
    dx = np.ascontiguousarray((m["x"] - g.gridx_b[ix]) / g.xstp_b[ix])
    dy = np.ascontiguousarray((m["y"] - g.gridy_b[iy]) / g.ystp_b[iy])
    eta = np.array(m["eta"], dtype=np.float64)
    pf = np.array(m["pfailure"], dtype=np.float32)
    L.ebo_marker_viscoplastic_from_grid(n, p(flags), p(xn), p(yn), p(dx), p(dy), g.ynum,
                                        p(np.asfortranarray(plastic_yield, dtype=np.float64)),
                                        p(np.asfortranarray(etas1, dtype=np.float64)),
                                        p(np.ascontiguousarray(m["eta_flow"], dtype=np.float64)), p(eta), p(pf))
    return eta, pf


def update_marker_strain(m: dict, timestep: float, healing: float):
    """Advection.update_marker_strain! on copies: (GII, strain_plastic,
    strain_rate_plastic)."""
    L = lib()
    L.ebo_update_marker_strain.restype = None
    L.ebo_update_marker_strain.argtypes = [_I64, _P, _D, _D] + [_P] * 10
    n = m["exx"].size
    f = lambda k: np.ascontiguousarray(m[k], dtype=np.float64)
    gii, ps, psr = (np.array(m[k], dtype=np.float64) for k in ("GII", "strain_plastic", "strain_rate_plastic"))
    L.ebo_update_marker_strain(n, p(np.ascontiguousarray(m["flags"], dtype=np.int8)), timestep, healing,
                               p(f("exx")), p(f("exy")), p(f("sxx")), p(f("sxy")), p(f("eta")), p(f("eta_flow")),
                               p(np.ascontiguousarray(m["pfailure"], dtype=np.float32)), p(gii), p(ps), p(psr))
    return gii, ps, psr


# ---------------------------------------------------------------- heat path
class _HeatGrid(C.Structure):
    _fields_ = [("xnum", _I64), ("ynum", _I64)] + [(k, _P) for k in (
        "gridx_b", "gridy_b", "xstp_b", "ystp_b", "xstp_pr", "ystp_pr", "gridx_vy", "xstp_vy", "gridy_vx",
        "ystp_vx")]


_HEAT_SIGS = {
    "ebo_scatter_basic_inclusive": (None, [C.c_void_p, _I64, _P, _P, _P, _P, _P, C.c_int, _P, _P, _P]),
    "ebo_marker_heat_production": (None, [_I64, _P, _P, _P, _P, _P]),
    "ebo_apply_thermal_bc": (None, [_I64, _I64, _P, _P, _P, _P, _P]),
    "ebo_heat_rhs": (None, [_I64, _I64] + [_P] * 9 + [_D, _D, C.c_int, C.c_int, _P]),
    "ebo_heat_build": (_I64, [C.c_void_p, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P, _P]),
    "ebo_heat_residuals": (None, [C.c_void_p, _P, _P, _P, _P, _P, _D, _P]),
    "ebo_grid_to_marker_temperature": (None, [C.c_void_p, _I64, _P, _P, _P, _P, _P, _P, C.c_int, _P]),
    "ebo_subgrid_temperature_markers": (None, [C.c_void_p, _I64] + [_P] * 8 + [_D] * 4 + [_P, _P]),
    "ebo_sticky_temperature_correction": (None, [_I64, _P, _P, C.c_int16, C.c_int16, _D, _P]),
    "ebo_subgrid_stress_markers": (None, [C.c_void_p, _I64] + [_P] * 10 + [_D, _D] + [_P] * 4),
}


def _heat_lib():
    L = lib()
    for name, (res, args) in _HEAT_SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _fm(a):
    return np.asfortranarray(a, dtype=np.float64)


class _HG:
    """Heat grid struct plus the arrays it points at (kept alive)."""

    def __init__(self, g):
        self.keep = {k: _f64(getattr(g, k)) for k in ("gridx_b", "gridy_b", "xstp_b", "ystp_b", "xstp_pr",
                                                      "ystp_pr", "gridx_vy", "xstp_vy", "gridy_vx", "ystp_vx")}
        self.s = _HeatGrid(g.xnum, g.ynum, *[p(self.keep[k]) for k in (
            "gridx_b", "gridy_b", "xstp_b", "ystp_b", "xstp_pr", "ystp_pr", "gridx_vy", "xstp_vy", "gridy_vx",
            "ystp_vx")])

    @property
    def ref(self):
        return C.byref(self.s)


def _marker_cells(g, m):
    flags = np.ascontiguousarray(m["flags"], dtype=np.int8)
    if "marker_xn" in m:
        xn = np.ascontiguousarray(m["marker_xn"], dtype=np.int32)
        yn = np.ascontiguousarray(m["marker_yn"], dtype=np.int32)
    else:
        xn, yn = cell_indices_basic(m["x"], m["y"], flags, g)
    return flags, xn, yn


def scatter_basic_inclusive(g, m: dict, vals: list, backups: list):
    """Inclusive basic-grid scatter of per-marker values (BilinearNumerator.jl:31-141
    with the averages of BilinearAverage.jl:276-294).  backups: the values kept
    where a node receives no weight.  Returns (averages, wtnodes)."""
    L = _heat_lib()
    hg = _HG(g)
    flags, xn, yn = _marker_cells(g, m)
    x, y = _f64(m["x"]), _f64(m["y"])
    vals = [_f64(v) for v in vals]
    outs = [np.array(b, dtype=np.float64, order="F", copy=True) for b in backups]
    wt = np.zeros((g.ynum, g.xnum), order="F")
    vp = (C.c_void_p * max(1, len(vals)))(*[p(v) for v in vals])
    op = (C.c_void_p * max(1, len(outs)))(*[p(o) for o in outs])
    L.ebo_scatter_basic_inclusive(hg.ref, x.size, p(x), p(y), p(flags), p(xn), p(yn), len(vals),
                                  C.cast(vp, _P), C.cast(op, _P), p(wt))
    return outs, wt


HEAT_FIELDS = ("tk1", "kt1", "rhocp1", "hr1", "ha1")


def marker_heat_production(m: dict, mat_hr, serp=None):
    """Per-marker heat production (BilinearNumerator.jl:107-113)."""
    L = _heat_lib()
    flags = np.ascontiguousarray(m["flags"], dtype=np.int8)
    n = flags.size
    hr = np.empty(n)
    L.ebo_marker_heat_production(n, p(flags), p(np.ascontiguousarray(m["matid"], dtype=np.int16)), p(_f64(mat_hr)),
                                 p(None if serp is None else np.ascontiguousarray(serp, dtype=np.float64)), p(hr))
    return hr


def scatter_heat(g, m: dict, prev: dict, mat_hr, serp=None, bc: dict | None = None):
    """MarkerToHeatArrays.interpolate_markers_to_transport_arrays!
    (MarkerToHeatArrays.jl:29-35): backups (tk0 from tk2, the others from the
    previous suffix-1 arrays, BackupTransportArrays.jl:154-163), the inclusive
    scatter of TK, kt, rhocp, heat production and ha, the averages, then the
    thermal boundary conditions on tk1 when bc is given.  prev keys: tk2,
    kt1, rhocp1, hr1, ha1."""
    hr = marker_heat_production(m, mat_hr, serp)
    vals = [m["TK"], m["kt"], m["rhocp"], hr, m["ha"]]
    backs = [prev["tk2"], prev["kt1"], prev["rhocp1"], prev["hr1"], prev["ha1"]]
    outs, wt = scatter_basic_inclusive(g, m, vals, backs)
    out = dict(zip(HEAT_FIELDS, outs))
    out["wtnodes"] = wt
    if bc is not None:
        out["tk1"] = apply_thermal_bc(g, bc, out["tk1"])
    return out


def apply_thermal_bc(g, bc: dict, tk1):
    L = _heat_lib()
    t = np.array(tk1, dtype=np.float64, order="F", copy=True)
    L.ebo_apply_thermal_bc(g.xnum, g.ynum, p(_fm(bc["btopt"])), p(_fm(bc["bbottomt"])), p(_fm(bc["bleftt"])),
                           p(_fm(bc["brightt"])), p(t))
    return t


def heat_rhs(g, f: dict, gravity_x, gravity_y, iuse_adiabatic, iuse_shear):
    """HeatSolver.calculate_heat_rhs_terms_grid! (HeatSolver.jl:107-118)."""
    L = _heat_lib()
    RT1 = np.empty((g.ynum, g.xnum), order="F")
    z = lambda k, s: _fm(f[k]) if k in f else np.zeros(s, order="F")
    ny, nx = g.ynum, g.xnum
    L.ebo_heat_rhs(nx, ny, p(_fm(f["hr1"])), p(z("ha1", (ny, nx))), p(z("rho1", (ny, nx))),
                   p(z("vx1", (ny + 1, nx))), p(z("vy1", (ny, nx + 1))), p(z("sxy2", (ny, nx))),
                   p(z("etas1", (ny, nx))), p(z("sxx2", (ny - 1, nx - 1))), p(z("etan1", (ny - 1, nx - 1))),
                   gravity_x, gravity_y, int(iuse_adiabatic), int(iuse_shear), p(RT1))
    return RT1


def heat_build(g, bc: dict, RT1, tk0, rhocp1, kt1, timestep):
    """HeatBuildManager.build_system_of_equations: (N, Li, Lj, Lv, R), 1-based."""
    L = _heat_lib()
    hg = _HG(g)
    N = g.xnum * g.ynum
    Li = np.empty(5 * N, np.int64)
    Lj = np.empty(5 * N, np.int64)
    Lv = np.empty(5 * N)
    R = np.empty(N)
    nz = L.ebo_heat_build(hg.ref, p(_fm(bc["btopt"])), p(_fm(bc["bbottomt"])), p(_fm(bc["bleftt"])),
                          p(_fm(bc["brightt"])), p(_fm(RT1)), p(_fm(tk0)), p(_fm(rhocp1)), p(_fm(kt1)),
                          timestep, p(Li), p(Lj), p(Lv), p(R))
    return N, Li[:nz].copy(), Lj[:nz].copy(), Lv[:nz].copy(), R


def heat_solve(g, bc: dict, RT1, tk0, rhocp1, kt1, timestep):
    """discretize_and_solve_conductive_heat_equation (SystemSolver.jl:32-66):
    tk2 from the direct solve of the restated system (reload_solu_temp keeps
    the column-major node order)."""
    N, Li, Lj, Lv, R = heat_build(g, bc, RT1, tk0, rhocp1, kt1, timestep)
    S = direct_solve(N, Li, Lj, Lv, R)
    return np.asfortranarray(S.reshape((g.ynum, g.xnum), order="F"))


def heat_residuals(g, RT1, rhocp1, tk0, tk2, kt1, timestep):
    L = _heat_lib()
    hg = _HG(g)
    rest = np.empty((g.ynum, g.xnum), order="F")
    L.ebo_heat_residuals(hg.ref, p(_fm(RT1)), p(_fm(rhocp1)), p(_fm(tk0)), p(_fm(tk2)), p(_fm(kt1)), timestep,
                         p(rest))
    return rest


def heat_loop(g, bc: dict, RT1, tk1, rhocp1, kt1, timestep, max_temp_change):
    """HeatLoop.run_conductive_heat_solver_loop! (heat_loop/HeatLoop.jl:12-78)
    starting from tk0 = tk1 (HeatSolver.initialize_old_temperature_array!).
    Returns (tk2, rest, log of (timestep_heat, dtkmax, resolved))."""
    tk0 = np.array(tk1, order="F", copy=True)
    dt_loop = timestep
    total = 0.0
    log = []
    tk2 = rest = None
    while total < timestep:
        tk2 = heat_solve(g, bc, RT1, tk0, rhocp1, kt1, dt_loop)
        rest = heat_residuals(g, RT1, rhocp1, tk0, tk2, kt1, dt_loop)
        dtkmax = float(np.max(np.abs(tk2 - tk0)))
        resolved = dtkmax > max_temp_change
        if resolved:
            dt_loop = dt_loop * max_temp_change / dtkmax
            tk2 = heat_solve(g, bc, RT1, tk0, rhocp1, kt1, dt_loop)
            rest = heat_residuals(g, RT1, rhocp1, tk0, tk2, kt1, dt_loop)
        log.append((dt_loop, dtkmax, resolved))
        total = total + dt_loop
        dt_loop = min(dt_loop, timestep - total)
        tk0 = tk2.copy(order="F")
    return tk2, rest, log


def grid_to_marker_temperature(g, m: dict, grid, mode: int, marker_TK):
    """mode 0: GridTemperatureToMarkers.interpolate!; mode 1: add the grid
    change (MarkerTemperature.add_remaining_grid_change_to_marker_temperature!)."""
    L = _heat_lib()
    hg = _HG(g)
    flags, xn, yn = _marker_cells(g, m)
    t = np.array(marker_TK, dtype=np.float64, copy=True)
    L.ebo_grid_to_marker_temperature(hg.ref, t.size, p(_f64(m["x"])), p(_f64(m["y"])), p(flags), p(xn), p(yn),
                                     p(_fm(grid)), int(mode), p(t))
    return t


def subgrid_temperature_markers(g, m: dict, tk1, kt1, rhocp1, timestep, coef):
    L = _heat_lib()
    hg = _HG(g)
    flags, xn, yn = _marker_cells(g, m)
    t = np.array(m["TK"], dtype=np.float64, copy=True)
    dT = np.empty(t.size)
    L.ebo_subgrid_temperature_markers(hg.ref, t.size, p(_f64(m["x"])), p(_f64(m["y"])), p(flags), p(xn), p(yn),
                                      p(_fm(tk1)), p(_fm(kt1)), p(_fm(rhocp1)), timestep, coef, g.xstpavg,
                                      g.ystpavg, p(t), p(dT))
    return t, dT


def sticky_temperature_correction(m: dict, matid_air, matid_water, temperature_top):
    L = _heat_lib()
    t = np.array(m["TK"], dtype=np.float64, copy=True)
    L.ebo_sticky_temperature_correction(t.size, p(np.ascontiguousarray(m["flags"], dtype=np.int8)),
                                        p(np.ascontiguousarray(m["matid"], dtype=np.int16)), int(matid_air),
                                        int(matid_water), temperature_top, p(t))
    return t


def update_marker_temp_for_grid_and_subgrid_changes(g, bc: dict, m: dict, f: dict, timestep, max_temp_change,
                                                    subgrid_coef, gravity_x=0.0, gravity_y=0.0,
                                                    iuse_adiabatic=0, iuse_shear=0):
    """HeatSolver.update_marker_temp_for_grid_and_subgrid_changes! (HeatSolver.jl:77-101):
    f holds tk1, kt1, rhocp1, hr1 (plus ha1, rho1, vx1, vy1, sxy2, etas1, sxx2,
    etan1 for the optional right-side terms).  Returns a dict with marker TK,
    RT1, tk2, rest, dtk1 (remaining change), dtkn and the heat-loop log."""
    RT1 = heat_rhs(g, f, gravity_x, gravity_y, iuse_adiabatic, iuse_shear)
    tk2, rest, log = heat_loop(g, bc, RT1, f["tk1"], f["rhocp1"], f["kt1"], timestep, max_temp_change)
    dtk1 = np.asfortranarray(tk2 - f["tk1"])
    TK = np.array(m["TK"], dtype=np.float64, copy=True)
    dtkn = None
    if subgrid_coef > 0:
        TK, dT = subgrid_temperature_markers(g, m, f["tk1"], f["kt1"], f["rhocp1"], timestep, subgrid_coef)
        (dtkn,), _ = scatter_basic_inclusive(g, m, [dT], [np.zeros((g.ynum, g.xnum), order="F")])
        dtk1 = np.asfortranarray(dtk1 - dtkn)
    TK = grid_to_marker_temperature(g, m, dtk1, 1, TK)
    return {"TK": TK, "RT1": RT1, "tk2": tk2, "rest": rest, "dtk1": dtk1, "dtkn": dtkn, "log": log}


def subgrid_stress_markers(g, m: dict, sxy1, sxx1, timestep, coef):
    L = _heat_lib()
    hg = _HG(g)
    flags, xn, yn = _marker_cells(g, m)
    sxy = np.array(m["sxy"], dtype=np.float64, copy=True)
    sxx = np.array(m["sxx"], dtype=np.float64, copy=True)
    n = sxy.size
    dxy, dxx = np.empty(n), np.empty(n)
    L.ebo_subgrid_stress_markers(hg.ref, n, p(_f64(m["x"])), p(_f64(m["y"])), p(flags), p(xn), p(yn),
                                 p(np.ascontiguousarray(m["matid"], dtype=np.int16)), p(_f64(m["mat_mu"])),
                                 p(_f64(m["eta"])), p(_fm(sxy1)), p(_fm(sxx1)), timestep, coef, p(sxy), p(sxx),
                                 p(dxy), p(dxx))
    return sxy, sxx, dxy, dxx


def update_subgrid_stress(g, m: dict, f: dict, timestep, coef):
    """SubGridStress.Update.update_subgrid_stress! (subgrid_stress/update/Update.jl:55-71):
    relaxed marker stresses, their exclusive (shear, basic grid) and central
    (normal, pressure grid) scatter into dsxyn, dsxxn with zero where no weight
    (BilinearAverage.jl:234-251), and the remaining changes dsxy - dsxyn,
    dsxx - dsxxn (RemainingStress.jl:16-23)."""
    sxy, sxx, dxy, dxx = subgrid_stress_markers(g, m, f["sxy1"], f["sxx1"], timestep, coef)
    ny, nx = g.ynum, g.xnum
    sm = {"x": m["x"], "y": m["y"], "inside_flags": np.ascontiguousarray(m["flags"], dtype=np.int8),
          "sxy": dxy, "sxx": dxx}
    prev = {"sxy1": np.zeros((ny, nx), order="F"), "sxx1": np.zeros((ny - 1, nx - 1), order="F")}
    sc = scatter_stokes(g, sm, prev)
    dsxyn, dsxxn = sc["sxy1"], sc["sxx1"]
    return {"sxy": sxy, "sxx": sxx, "dsxyn": dsxyn, "dsxxn": dsxxn,
            "dsxy": np.asfortranarray(f["dsxy"] - dsxyn), "dsxx": np.asfortranarray(f["dsxx"] - dsxxn),
            "wtetas": sc["wtetas"], "wtetan": sc["wtetan"]}


# ---------------------------------------------------------------- marker plasticity
class _MarkerPlasticIn(C.Structure):
    _fields_ = [("marknum", _I64), ("flags", _P), ("y", _P), ("matid", _P), ("mat_mu", _P), ("cohesion", _P),
                ("fric", _P), ("exx", _P), ("exy", _P), ("sr_ratio", _P), ("pr", _P), ("forecast", _I64),
                ("ntimestep", _I64), ("timestep", _D), ("y_sealevel", _D), ("viscosity_min", _D),
                ("viscosity_max", _D), ("yield_stress_min", _D), ("yield_stress_max", _D),
                ("iuse_fluid_pressure_for_yield", _I64), ("no_yielding_in_mobile_wall", _I64),
                ("matid_sticky_air", _I64), ("matid_mobile_wall", _I64), ("matid_plate_extension", _I64)]


MARKER_PLASTIC_SCALARS = ("forecast", "ntimestep", "timestep", "y_sealevel", "viscosity_min", "viscosity_max",
                          "yield_stress_min", "yield_stress_max", "iuse_fluid_pressure_for_yield",
                          "no_yielding_in_mobile_wall", "matid_sticky_air", "matid_mobile_wall",
                          "matid_plate_extension")


def update_marker_plastic_yielding(m: dict, par: dict):
    """MarkerPlasticity.update_plastic_yielding! on copies: (sxx, sxy, eta,
    pfailure).  m: flags, y, matid, mat_mu, cohesion, fric, exx, exy,
    sr_ratio, pr, sxx, sxy, eta, pfailure; par: MARKER_PLASTIC_SCALARS."""
    L = lib()
    L.ebo_update_marker_plastic_yielding.restype = None
    L.ebo_update_marker_plastic_yielding.argtypes = [C.POINTER(_MarkerPlasticIn), _P, _P, _P, _P]
    f = lambda k: np.ascontiguousarray(m[k], dtype=np.float64)
    keep = {k: f(k) for k in ("y", "cohesion", "fric", "exx", "exy", "sr_ratio", "pr")}
    flags = np.ascontiguousarray(m["flags"], dtype=np.int8)
    matid = np.ascontiguousarray(m["matid"], dtype=np.int16)
    mu = np.ascontiguousarray(m["mat_mu"], dtype=np.float64)
    a = _MarkerPlasticIn(flags.size, p(flags), p(keep["y"]), p(matid), p(mu),
                         *[p(keep[k]) for k in ("cohesion", "fric", "exx", "exy", "sr_ratio", "pr")],
                         *[par[k] for k in MARKER_PLASTIC_SCALARS])
    sxx, sxy, eta = (np.array(m[k], dtype=np.float64) for k in ("sxx", "sxy", "eta"))
    pf = np.array(m["pfailure"], dtype=np.float32)
    L.ebo_update_marker_plastic_yielding(C.byref(a), p(sxx), p(sxy), p(eta), p(pf))
    return sxx, sxy, eta, pf
