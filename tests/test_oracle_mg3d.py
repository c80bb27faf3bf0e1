"""Pins of the restated 3-D multigrid oracle (oracle/mg3d.c, oracle/mg3.py).

What fixes each function independently of its own formulas:
  - residuals: manufactured fields whose staggered differences are exact
    (linear velocity, pressure and viscosity on non-uniform axes; quadratic
    velocity with constant viscosity on uniform axes) against calculus, and
    the diagonal coefficient against its closed form;
  - ghost pass: face-interior free-slip rules (normal zero, tangential mirror)
    and the level-size rule against the reference's examples;
  - transfers: partition of unity and exact reproduction of linear fields;
  - smoother and V-cycle: the fixed point equals a sparse direct solve of the
    operator assembled column by column, for both sweep orders.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import mg3

RNG = np.random.default_rng(1234)


def nonuniform_axis(n, size, rng):
    h = rng.uniform(0.6, 1.4, n - 1)
    b = np.concatenate([[0.0], np.cumsum(h)]) * size / h.sum()
    return mg3.axis_from_basic(b)


def node_coords(lv, field):
    kinds = {"vx": "vbv", "vy": "bvv", "vz": "vvb", "pr": "ppp", "etaxy": "bbs",
             "etaxz": "sbb", "etayz": "bsb"}[field]
    key = {"b": "b", "v": "v", "p": "p", "s": "s"}
    y = lv.ay[key[kinds[0]]]
    x = lv.ax[key[kinds[1]]]
    z = lv.az[key[kinds[2]]]
    return np.meshgrid(y, x, z, indexing="ij")


def fill(lv, field, f):
    Y, X, Z = node_coords(lv, field)
    return np.asfortranarray(f(X, Y, Z))


def interior(field, a):
    return a[1:-1, 1:-1, 1:-1] if field != "pr" else a


def test_level_sizes_reference_rule():
    # LevelManager.calculate_multigrid_levels_3d: halve with ceil while the
    # smallest dimension stays >= 4, at most 10 levels
    assert [s[0] for s in mg3.level_sizes(257, 257, 257)] == [257, 129, 65, 33, 17, 9, 5]
    assert [s[0] for s in mg3.level_sizes(97, 97, 97)] == [97, 49, 25, 13, 7, 4]
    assert [s[0] for s in mg3.level_sizes(17, 17, 17)] == [17, 9, 5]
    assert mg3.level_sizes(9, 40, 12) == [(9, 40, 12), (5, 20, 6)]


def test_axis_kinds_geometry():
    a = nonuniform_axis(9, 3.0, RNG)
    b = a["b"]
    assert np.allclose(a["v"][1:-1], 0.5 * (b[1:] + b[:-1]), rtol=0, atol=1e-15)
    assert np.isclose(a["v"][0], b[0] - (b[1] - b[0]) / 2) and np.isclose(a["v"][-1], b[-1] + (b[-1] - b[-2]) / 2)
    assert np.allclose(a["stp_v"][1:-1], 0.5 * (b[2:] - b[:-2]))
    assert np.allclose(a["p"], 0.5 * (b[1:] + b[:-1]))
    assert np.allclose(a["s"], 0.5 * (b[1:] + b[:-1]))


def _manufactured_level(uniform, uniform_x=False):
    nx, ny, nz = 7, 8, 6
    if uniform:
        lv = mg3.Level(nx, ny, nz, 1.3, 0.9, 1.1)
    else:
        axes = [nonuniform_axis(n, s, RNG) for n, s in ((nx, 1.3), (ny, 0.9), (nz, 1.1))]
        if uniform_x:
            axes[0] = mg3.axis_uniform(nx, 1.3)
        lv = mg3.Level(nx, ny, nz, 0, 0, 0, axes=tuple(axes))
    return lv


def _residual_arrays(lv):
    lv.rx[...] = lv.ry[...] = lv.rz[...] = lv.rc[...] = np.nan
    mg3._sigs().ebo_mg3_residuals(mg3.C.byref(lv.cs()))
    return {"vx": lv.rx, "vy": lv.ry, "vz": lv.rz, "pr": lv.rc}


@pytest.mark.parametrize("seed,uniform_x", [(0, True), (1, True), (2, False)])
def test_residuals_linear_fields_nonuniform(seed, uniform_x):
    """Linear v, p, eta: stress divergence is grad(eta) . (G + G^T) exactly.

    The reference's y-Stokes stencil (Residuals.jl:311-313) takes the spacing
    between the vy nodes j-1 and j from xstp_b[j-1] and the spacing between
    the shear nodes from xstp_vy[j-1]; the two coincide on the uniform x axes
    the multigrid builds, so the vy equation is checked against calculus only
    when x is uniform (y and z stay non-uniform)."""
    rng = np.random.default_rng(seed)
    lv = _manufactured_level(uniform=False, uniform_x=uniform_x)
    G = rng.normal(size=(3, 3))      # G[a, b] = d v_a / d x_b, axes (x, y, z)
    c = rng.normal(size=3)
    gp = rng.normal(size=3)
    ge = rng.uniform(0.1, 0.5, size=3)
    eta = lambda X, Y, Z: 2.0 + ge[0] * X + ge[1] * Y + ge[2] * Z
    comp = {"vx": 0, "vy": 1, "vz": 2}
    for f in ("vx", "vy", "vz"):
        a = comp[f]
        getattr(lv, f)[...] = fill(lv, f, lambda X, Y, Z, a=a: c[a] + G[a, 0] * X + G[a, 1] * Y + G[a, 2] * Z)
    lv.pr[...] = fill(lv, "pr", lambda X, Y, Z: gp[0] * X + gp[1] * Y + gp[2] * Z)
    lv.etan[...] = fill(lv, "pr", eta)
    for f in ("etaxy", "etaxz", "etayz"):
        getattr(lv, f)[...] = fill(lv, f, eta)
    for R in (lv.RX, lv.RY, lv.RZ, lv.RC):
        R[...] = rng.normal(size=R.shape)
    res = _residual_arrays(lv)
    S = G + G.T
    for f, R in (("vx", lv.RX), ("vy", lv.RY), ("vz", lv.RZ)):
        if f == "vy" and not uniform_x:
            continue
        a = comp[f]
        divsig = sum(ge[b] * S[a, b] for b in range(3))
        expect = R + gp[a] - divsig
        np.testing.assert_allclose(interior(f, res[f]), interior(f, expect), rtol=1e-10, atol=1e-10)
        assert np.all(res[f][0] == 0) and np.all(res[f][-1] == 0)  # ghost rows zero
    np.testing.assert_allclose(res["pr"], lv.RC - np.trace(G), rtol=1e-10, atol=1e-10)


def test_residuals_quadratic_fields_uniform():
    """Quadratic v, constant eta on uniform axes: second differences are exact."""
    rng = np.random.default_rng(7)
    lv = _manufactured_level(uniform=True)
    eta0 = 3.0
    for a in (lv.etan, lv.etaxy, lv.etaxz, lv.etayz):
        a[...] = eta0
    # v_a = sum_{b<=c} Q[a,b,c] x_b x_c ; Hessians H_a
    Q = rng.normal(size=(3, 3, 3))
    H = np.zeros((3, 3, 3))
    for a in range(3):
        for b in range(3):
            for cc in range(b, 3):
                H[a, b, cc] += Q[a, b, cc]
                H[a, cc, b] += Q[a, b, cc]

    def vfun(a):
        def f(X, Y, Z):
            P = (X, Y, Z)
            return sum(Q[a, b, cc] * P[b] * P[cc] for b in range(3) for cc in range(b, 3))
        return f
    for a, f in enumerate(("vx", "vy", "vz")):
        getattr(lv, f)[...] = fill(lv, f, vfun(a))
    pq = rng.normal(size=3)
    lv.pr[...] = fill(lv, "pr", lambda X, Y, Z: pq[0] * X * X + pq[1] * Y * Z + pq[2] * Z)
    res = _residual_arrays(lv)
    for a, f in enumerate(("vx", "vy", "vz")):
        Y, X, Z = node_coords(lv, f)
        dp = [2 * pq[0] * X, pq[1] * Z, pq[1] * Y + pq[2]][a]
        # div(eta (grad v + grad v^T))_a = eta (lap v_a + d_a div v)
        lap = np.trace(H[a])
        ddiv = sum(H[b, b, a] for b in range(3))
        expect = dp - eta0 * (lap + ddiv)
        np.testing.assert_allclose(interior(f, res[f]), interior(f, expect), rtol=1e-9, atol=1e-9)


def test_diagonal_coefficient_closed_form():
    lv = _manufactured_level(uniform=True)
    eta0 = 2.5
    for a in (lv.etan, lv.etaxy, lv.etaxz, lv.etayz):
        a[...] = eta0
    hx, hy, hz = lv.ax["stp_b"][0], lv.ay["stp_b"][0], lv.az["stp_b"][0]
    L = lv.cs()
    coef = mg3.C.c_double()
    fn = mg3._sigs()
    for name, (w2, w1a, w1b) in (("vx", (hx, hy, hz)), ("vy", (hy, hx, hz)), ("vz", (hz, hx, hy))):
        f = getattr(fn, f"ebo_mg3_{name}_residual")
        f.restype = mg3.C.c_double
        f.argtypes = [mg3.C.POINTER(mg3._Level)] + [mg3.C.c_int64] * 3 + [mg3.C.c_void_p]
        f(mg3.C.byref(L), 3, 3, 3, mg3.C.addressof(coef))
        expect = -eta0 * (4.0 / w2 ** 2 + 2.0 / w1a ** 2 + 2.0 / w1b ** 2)
        assert coef.value == pytest.approx(expect, rel=1e-13)


def test_ghost_pass_face_rules():
    lv = mg3.Level(6, 7, 5, 1.0, 1.0, 1.0)
    for f in ("vx", "vy", "vz"):
        getattr(lv, f)[...] = RNG.normal(size=getattr(lv, f).shape)
    mg3._sigs().ebo_mg3_set_bc(mg3.C.byref(lv.cs()))
    vx, vy, vz = lv.vx, lv.vy, lv.vz
    s = np.s_[1:-1]
    # vx: tangential mirror at y and z faces, normal zero at x faces
    assert np.array_equal(vx[0, s, s], vx[1, s, s]) and np.array_equal(vx[-1, s, s], vx[-2, s, s])
    assert np.array_equal(vx[s, s, 0], vx[s, s, 1]) and np.array_equal(vx[s, s, -1], vx[s, s, -2])
    assert np.all(vx[s, 0, s] == 0) and np.all(vx[s, -1, s] == 0)
    # vy: mirror at x and z faces, zero at y faces
    assert np.array_equal(vy[s, 0, s], vy[s, 1, s]) and np.array_equal(vy[s, -1, s], vy[s, -2, s])
    assert np.array_equal(vy[s, s, 0], vy[s, s, 1]) and np.array_equal(vy[s, s, -1], vy[s, s, -2])
    assert np.all(vy[0, s, s] == 0) and np.all(vy[-1, s, s] == 0)
    # vz: mirror at y and x faces, zero at z faces
    assert np.array_equal(vz[0, s, s], vz[1, s, s]) and np.array_equal(vz[-1, s, s], vz[-2, s, s])
    assert np.array_equal(vz[s, 0, s], vz[s, 1, s]) and np.array_equal(vz[s, -1, s], vz[s, -2, s])
    assert np.all(vz[s, s, 0] == 0) and np.all(vz[s, s, -1] == 0)


def _hierarchy(dims=(9, 11, 10), sizes=(1.0, 1.2, 1.1), seed=0):
    rng = np.random.default_rng(seed)
    H = mg3.Hierarchy(*dims, *sizes)
    L0 = H.levels[0]
    for k in ("etan", "etaxy", "etaxz", "etayz"):
        getattr(L0, k)[...] = 10 ** rng.uniform(0, 1, getattr(L0, k).shape)
    L0.RX[...] = rng.normal(size=L0.RX.shape)
    L0.RZ[...] = rng.normal(size=L0.RZ.shape)
    H.restrict_viscosity()
    return H


def test_transfers_partition_of_unity_and_linear_exactness():
    H = _hierarchy(dims=(13, 11, 12))
    f, c = H.levels[0], H.levels[1]
    # constant fine residual -> the same constant on interior coarse nodes
    for a in (f.rx, f.ry, f.rz):
        a[...] = 2.5
    f.rc[...] = 2.5 / f.etan   # continuity is scaled by etan before restriction
    H.restrict(0)
    np.testing.assert_allclose(c.RX[1:-1, 1:-1, 1:-1], 2.5, rtol=1e-14)
    np.testing.assert_allclose(c.RY[1:-1, 1:-1, 1:-1], 2.5, rtol=1e-14)
    np.testing.assert_allclose(c.RZ[1:-1, 1:-1, 1:-1], 2.5, rtol=1e-14)
    np.testing.assert_allclose(c.RC * c.etan, 2.5, rtol=1e-13)
    # prolongation reproduces linear coarse fields on every fine node
    g = RNG.normal(size=4)
    lin = lambda X, Y, Z: g[0] + g[1] * X + g[2] * Y + g[3] * Z
    for fld in ("vx", "vy", "vz", "pr"):
        getattr(c, fld)[...] = fill(c, fld, lin)
        getattr(f, fld)[...] = 0.0
    H.prolong(1)
    for fld in ("vx", "vy", "vz", "pr"):
        np.testing.assert_allclose(getattr(f, fld), fill(f, fld, lin), rtol=1e-12, atol=1e-12)
    # viscosity restriction keeps a constant
    for k in ("etan", "etaxy", "etaxz", "etayz"):
        getattr(f, k)[...] = 7.0
    H.restrict_viscosity()
    for k in ("etan", "etaxy", "etaxz", "etayz"):
        np.testing.assert_allclose(getattr(c, k), 7.0, rtol=1e-14)


def _direct_solution(L0):
    A, masks = mg3.assemble(L0)
    b = mg3.pack({"vx": L0.RX, "vy": L0.RY, "vz": L0.RZ, "pr": L0.RC}, masks)
    b[b.size - int(masks["pr"].sum())] = 0.0   # gauge p(1,1,1) = 0
    return mg3.unpack(spla.spsolve(A.tocsc(), b), L0, masks), masks


@pytest.mark.parametrize("serial", [False, True])
def test_vcycle_fixed_point_is_direct_solve(serial):
    H = _hierarchy()
    L0 = H.levels[0]
    xs, masks = _direct_solution(L0)
    sm = mg3.smoothing_counts(len(H.levels), 3, 2)
    for _ in range(110):
        H.vcycle(sm, serial=serial)
    for k in ("vx", "vy", "vz", "pr"):
        m = masks[k]
        err = np.linalg.norm(getattr(L0, k)[m] - xs[k][m]) / np.linalg.norm(xs[k][m])
        assert err < 1e-9, (k, err)


def test_operator_apply_matches_residual_definition():
    """A x = R - residual(x) on the unknowns, gauge row p(1,1,1)/etan(1,1,1)."""
    H = _hierarchy(dims=(7, 6, 8))
    L0 = H.levels[0]
    x = {k: RNG.normal(size=L0.sh[k]) for k in ("vx", "vy", "vz", "pr")}
    y = mg3.apply(L0, x)
    for k in ("vx", "vy", "vz", "pr"):
        getattr(L0, k)[...] = x[k]
    mg3._sigs().ebo_mg3_set_bc(mg3.C.byref(L0.cs()))
    mg3._sigs().ebo_mg3_residuals(mg3.C.byref(L0.cs()))
    for k, r, R in (("vx", L0.rx, L0.RX), ("vy", L0.ry, L0.RY), ("vz", L0.rz, L0.RZ)):
        np.testing.assert_allclose(interior(k, y[k]), interior(k, R - r), rtol=1e-9, atol=1e-9)
    expect = L0.RC - L0.rc
    expect.flat[0] = x["pr"].flat[0] / L0.etan.flat[0]
    np.testing.assert_allclose(y["pr"], expect, rtol=1e-9, atol=1e-9)


def test_mean_residuals_closed_form():
    """Interior RMS with ghost slots excluded: alternating +-c gives log10(c/scale)."""
    H = _hierarchy(dims=(7, 9, 8))
    L0 = H.levels[0]
    ny, nx, nz = L0.ynum, L0.xnum, L0.znum
    c = {"rx": 3.0, "ry": 0.25, "rz": 7.5, "rc": 0.04}
    for k in ("rx", "ry", "rz", "rc"):
        a = getattr(L0, k)
        a[...] = 1e30  # ghost slots must not count
    sl = {"rx": (slice(1, ny), slice(1, nx - 1), slice(1, nz)),
          "ry": (slice(1, ny - 1), slice(1, nx), slice(1, nz)),
          "rz": (slice(1, ny), slice(1, nx), slice(1, nz - 1)),
          "rc": (slice(None), slice(None), slice(None))}
    for k, s in sl.items():
        a = getattr(L0, k)
        sign = np.where(RNG.uniform(size=a[s].shape) < 0.5, -1.0, 1.0)
        a[s] = c[k] * sign
    ss, cs = 2.0, 0.5
    out = H.mean_residuals(ss, cs)
    expect = [np.log10(c["rx"] / ss), np.log10(c["ry"] / ss), np.log10(c["rz"] / ss),
              np.log10(c["rc"] / cs)]
    np.testing.assert_allclose(out, expect, rtol=0, atol=1e-14)


def _originals(H):
    return {k: getattr(H.levels[0], k).copy() for k in ("etan", "etaxy", "etaxz", "etayz")}


def test_run_continuation_schedule_and_stopping_rule():
    """Viscosity starts constant at the minimum, the maximum grows by
    (max/min)^(1/njumps) every nvcycles_viscosity_jump cycles, the full field
    returns after njumps jumps, and the stop comes at the first cycle past
    nvcycles_viscosity_jump * njumps when the criterion is loose."""
    H = _hierarchy(seed=3)
    eta0 = _originals(H)
    vmin = min(a.min() for a in eta0.values())
    vmax = max(a.max() for a in eta0.values())
    n, conv, hist = H.run(1, 1e300, nviscosity_jumps=3.0, nvcycles_viscosity_jump=2, resmax=0.0)
    assert (n, conv) == (1, False) and hist.shape == (1, 4)
    for k in eta0:
        assert np.all(getattr(H.levels[0], k) == vmin)
        assert np.all(getattr(H.levels[1], k) == vmin)
    H = _hierarchy(seed=3)
    H.run(2, 1e300, nviscosity_jumps=3.0, nvcycles_viscosity_jump=2, resmax=0.0)
    step = 10 ** (np.log10(vmax / vmin) / 3.0)
    top = max(getattr(H.levels[0], k).max() for k in eta0)
    bottom = min(getattr(H.levels[0], k).min() for k in eta0)
    np.testing.assert_allclose([bottom, top], [vmin, vmin * step], rtol=1e-13)
    H = _hierarchy(seed=3)
    n, conv, hist = H.run(20, 1e300, nviscosity_jumps=3.0, nvcycles_viscosity_jump=2, resmax=0.0)
    assert (n, conv) == (7, True) and hist.shape == (7, 4)
    for k in eta0:
        np.testing.assert_allclose(getattr(H.levels[0], k), eta0[k], rtol=1e-12)
    # resmax above every residual: a jump on every cycle, full field after njumps cycles
    H = _hierarchy(seed=3)
    n, conv, _ = H.run(3, 1e300, nviscosity_jumps=3.0, nvcycles_viscosity_jump=50, resmax=1e300)
    assert (n, conv) == (3, False)
    for k in eta0:
        np.testing.assert_allclose(getattr(H.levels[0], k), eta0[k], rtol=1e-12)


def test_run_converges_to_direct_solve():
    H = _hierarchy(seed=5)
    L0 = H.levels[0]
    xs, masks = _direct_solution(L0)   # full-contrast system
    n, conv, hist = H.run(300, 1e-11, nviscosity_jumps=2.0, nvcycles_viscosity_jump=5,
                          finest_sweeps=3)
    assert conv and 10 < n < 300
    assert np.all(10.0 ** hist[-1] < 1e-11)
    for k in ("vx", "vy", "vz", "pr"):
        m = masks[k]
        err = np.linalg.norm(getattr(L0, k)[m] - xs[k][m]) / np.linalg.norm(xs[k][m])
        assert err < 1e-8, (k, err)


def test_multimulti_schedule_and_direct_solution():
    """Multimulti (use_multimulti = true): with one cycle per jump and 1.9999
    jumps the global updates fall on cycles 3, 6, 9 (the maximum oversteps
    the full value at each ramp's second jump, derived from the rule), the
    run cannot stop before two updates; a converged run leaves the direct
    solution of the full-contrast system in level 0."""
    kw = dict(nviscosity_jumps=1.9999, nvcycles_viscosity_jump=1, finest_sweeps=2)
    for cycles, updates in ((2, 0), (3, 1), (6, 2), (9, 3)):
        H = _hierarchy(seed=3)
        n, conv, hist, gh = H.run_multimulti(cycles, 0.0, **kw)
        assert (n, conv, gh.shape) == (cycles, False, (updates, 4)), (cycles, gh.shape)
    H = _hierarchy(seed=3)
    n, conv, _, gh = H.run_multimulti(50, 1e30, **kw)
    assert (n, conv, gh.shape[0]) == (6, True, 2)
    H = _hierarchy(seed=5)
    xs, masks = _direct_solution(H.levels[0])
    n, conv, hist, gh = H.run_multimulti(600, 1e-11, nviscosity_jumps=2.9999, nvcycles_viscosity_jump=3,
                                         finest_sweeps=3)
    assert conv and n < 600 and gh.shape[0] >= 2
    for k in ("vx", "vy", "vz", "pr"):
        m = masks[k]
        err = np.linalg.norm(getattr(H.level0, k)[m] - xs[k][m]) / np.linalg.norm(xs[k][m])
        assert err < 1e-8, (k, err)
