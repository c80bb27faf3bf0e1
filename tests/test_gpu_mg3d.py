"""GPU parity of the reference's 3-D multigrid (FP64 reference-order API) and
of the 3-D Stokes solve, against the oracle (oracle/mg3d.c) through the C ABI.

Bars: ghost pass, smoother, restriction, prolongation, viscosity restriction,
V-cycles and the operator apply are bit-exact in FP64; the solve matches a
sparse direct solve of the oracle's assembled operator within 1e-8 relative
L2 per field when run to rtol 1e-11 (FP32 and FP64 preconditioners), and on
the sphere-inclusion workload converges to ||b - A x||/||b|| <= 1e-8 with the
residual recomputed independently.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from oracle import mg3
from earthbox_gpu import workloads
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.stokes3d import Multigrid3DGPU, Stokes3DSolverGPU

pytestmark = pytest.mark.gpu

CASES = [
    dict(dims=(9, 11, 10), sizes=(1.0, 1.2, 1.1), seed=0),
    dict(dims=(19, 17, 21), sizes=(2.0, 1.5, 1.8), seed=1),
    dict(dims=(6, 5, 7), sizes=(1.0, 1.0, 1.0), seed=2),
]
VISC = ("etan", "etaxy", "etaxz", "etayz")
SOL = ("vx", "vy", "vz", "pr")
RHS = ("RX", "RY", "RZ", "RC")
RES = ("rx", "ry", "rz", "rc")


def _pair(ctx, dims, sizes, seed):
    """Oracle hierarchy and device hierarchy with the same level-0 viscosity."""
    rng = np.random.default_rng(seed)
    H = mg3.Hierarchy(*dims, *sizes)
    G = Multigrid3DGPU(ctx, *dims, *sizes)
    assert G.nlevels == len(H.levels)
    assert G.shapes == [(lv.xnum, lv.ynum, lv.znum) for lv in H.levels]
    L0 = H.levels[0]
    for k in VISC:
        getattr(L0, k)[...] = 10 ** rng.uniform(-1, 2, getattr(L0, k).shape)
        G.set(0, k, to_device(getattr(L0, k)))
    H.restrict_viscosity()
    G.restrict_viscosity()
    return H, G, rng


def _fill(H, G, level, names, rng, scale=1.0):
    lv = H.levels[level]
    for k in names:
        a = getattr(lv, k)
        a[...] = rng.normal(size=a.shape) * scale
        G.set(level, k, to_device(a))


def _same(H, G, level, names):
    lv = H.levels[level]
    for k in names:
        got = to_host(G.get(level, k))
        exp = getattr(lv, k)
        assert np.array_equal(got, exp), (level, k, np.abs(got - exp).max())


@pytest.mark.parametrize("case", CASES)
def test_viscosity_restriction_bit_exact(gpu_ctx, case):
    print(f"mg3d viscosity restriction {case}", flush=True)
    H, G, _ = _pair(gpu_ctx, **case)
    for k in range(len(H.levels)):
        _same(H, G, k, VISC)


@pytest.mark.parametrize("case", CASES)
def test_ghost_pass_bit_exact(gpu_ctx, case):
    print(f"mg3d ghost pass {case}", flush=True)
    H, G, rng = _pair(gpu_ctx, **case)
    for level in range(len(H.levels)):
        _fill(H, G, level, ("vx", "vy", "vz"), rng)
        mg3._sigs().ebo_mg3_set_bc(mg3.C.byref(H.levels[level].cs()))
        G.set_bc(level)
        _same(H, G, level, ("vx", "vy", "vz"))


@pytest.mark.parametrize("fused", [0, 10 ** 9])
@pytest.mark.parametrize("case", CASES)
def test_smoother_bit_exact(gpu_ctx, case, fused):
    """Both smoother paths: one launch per phase (fused limit 0) and the
    single-workgroup launch per smoother call (every level fused)."""
    print(f"mg3d smoother {case} fused limit {fused}", flush=True)
    H, G, rng = _pair(gpu_ctx, **case)
    G.set_fused_limit(fused)
    for level in range(len(H.levels)):
        _fill(H, G, level, SOL + RHS, rng)
        H.smoother(level, 3, 0.9, 0.3, 0.37)
        G.smoother(level, 3, 0.9, 0.3, 0.37)
        _same(H, G, level, SOL + RES)


@pytest.mark.parametrize("case", CASES[:2])
def test_restrict_and_prolong_bit_exact(gpu_ctx, case):
    print(f"mg3d transfers {case}", flush=True)
    H, G, rng = _pair(gpu_ctx, **case)
    for level in range(len(H.levels) - 1):
        _fill(H, G, level, RES, rng)
        H.restrict(level)
        G.restrict(level)
        _same(H, G, level + 1, RHS)
        _fill(H, G, level + 1, SOL, rng)
        _fill(H, G, level, SOL, rng)
        H.prolong(level + 1, 1.0, 0.7)
        G.prolong(level + 1, 1.0, 0.7)
        _same(H, G, level, SOL)


@pytest.mark.parametrize("fused", [0, 12000])
@pytest.mark.parametrize("case", CASES[:2])
def test_vcycles_bit_exact(gpu_ctx, case, fused):
    print(f"mg3d vcycles {case} fused limit {fused}", flush=True)
    H, G, rng = _pair(gpu_ctx, **case)
    G.set_fused_limit(fused)
    _fill(H, G, 0, RHS, rng)
    for k in SOL:
        G.set(0, k, to_device(getattr(H.levels[0], k)))
    sm = mg3.smoothing_counts(len(H.levels), 2, 2, 8)
    for _ in range(3):
        H.vcycle(sm, 0.9, 0.3, 1.0, 1.0, 0.0)
    G.vcycles(3, finest_sweeps=2, multiplier=2, cap=8, theta_s=0.9, theta_c=0.3)
    _same(H, G, 0, SOL + RES)


def _smooth_viscosity(shape, rng, decades):
    """A smooth positive field spanning about `decades` decades (rough fields
    can restrict to negative coarse viscosity through the extrapolating
    boundary weights)."""
    ph = rng.uniform(0, 2 * np.pi, 3)
    i, j, k = np.meshgrid(*[np.linspace(0, 1, n) for n in shape], indexing="ij")
    g = np.sin(3 * i + ph[0]) * np.cos(2.5 * j + ph[1]) * np.sin(2 * k + ph[2])
    return np.asfortranarray(10 ** (decades * 0.5 * (1 + g)))


def _run_pair(ctx, dims, sizes, seed, decades):
    rng = np.random.default_rng(seed)
    H = mg3.Hierarchy(*dims, *sizes)
    G = Multigrid3DGPU(ctx, *dims, *sizes)
    L0 = H.levels[0]
    for k in VISC:
        getattr(L0, k)[...] = _smooth_viscosity(getattr(L0, k).shape, rng, decades)
        G.set(0, k, to_device(getattr(L0, k)))
    for k in RHS + SOL:
        a = getattr(L0, k)
        a[...] = rng.normal(size=a.shape) if k in ("RX", "RY", "RZ") else 0.0
        G.set(0, k, to_device(a))
    return H, G


RUN_KW = dict(nviscosity_jumps=2.0, nvcycles_viscosity_jump=3, resmax=1e-6, stokesscale=1.5,
              continscale=0.5, finest_sweeps=2, multiplier=2, cap=8)


@pytest.mark.parametrize("prnorm", [0.0, 2.0])
@pytest.mark.parametrize("case,criterion", [(CASES[0], 1e-8), (CASES[1], 1e-7)])
def test_run_matches_oracle(gpu_ctx, case, criterion, prnorm):
    """Stand-alone driver with viscosity continuation: same stopping cycle and
    flag, residual history within 1e-7 (log10 units) and the same solution to
    1e-9 relative.  Only the power law of the viscosity scaling and the order
    of the RMS sums differ from the oracle (last-bit effects)."""
    print(f"mg3d run {case} criterion {criterion} prnorm {prnorm}", flush=True)
    H, G = _run_pair(gpu_ctx, **case, decades=2.0)
    kw = dict(RUN_KW, pressure_bc=prnorm)
    n_o, c_o, h_o = H.run(200, criterion, **kw)
    n_g, c_g, h_g = G.run(200, criterion, **kw)
    assert (n_g, c_g) == (n_o, c_o) and c_o and n_o > 6
    np.testing.assert_allclose(h_g, h_o, rtol=0, atol=1e-7)
    L0 = H.levels[0]
    for k in SOL:
        got, exp = to_host(G.get(0, k)), getattr(L0, k)
        assert np.linalg.norm(got - exp) <= 1e-9 * np.linalg.norm(exp), k
    for k in VISC:
        np.testing.assert_allclose(to_host(G.get(0, k)), getattr(L0, k), rtol=1e-13)
    # a second run restarts the continuation from the finest-level viscosity
    # it now holds and stops when the cycle budget is spent
    n2, c2, h2 = G.run(5, criterion, **kw)
    assert n2 == 5 and not c2 and h2.shape == (5, 4)
    with pytest.raises(Exception):
        G.run(-1, criterion)


@pytest.mark.parametrize("case,criterion", [(CASES[0], 1e-8), (CASES[1], 1e-6)])
def test_run_multimulti_matches_oracle(gpu_ctx, case, criterion):
    """Multimulti driver: same cycle count, stop flag and global updates as
    the oracle, principal and global residual histories within 1e-6 (log10
    units; the global residuals are level 0's residuals, a difference of
    nearly equal terms), global solution to 1e-9 relative."""
    print(f"mg3d multimulti {case}", flush=True)
    H, G = _run_pair(gpu_ctx, **case, decades=2.0)
    kw = dict(RUN_KW, nviscosity_jumps=2.9999, nvcycles_viscosity_jump=3)
    kw.pop("resmax")
    n_o, c_o, h_o, g_o = H.run_multimulti(300, criterion, **kw)
    n_g, c_g, h_g, g_g = G.run_multimulti(300, criterion, **kw)
    print(f"  oracle {n_o} cycles, {len(g_o)} updates; device {n_g} cycles, {len(g_g)} updates", flush=True)
    assert (n_g, c_g, len(g_g)) == (n_o, c_o, len(g_o)) and c_o and len(g_o) >= 2
    np.testing.assert_allclose(h_g, h_o, rtol=0, atol=1e-6)
    np.testing.assert_allclose(g_g, g_o, rtol=0, atol=1e-6)
    for k in SOL:
        got, exp = to_host(G.get(0, k)), getattr(H.level0, k)
        assert np.linalg.norm(got - exp) <= 1e-9 * np.linalg.norm(exp), k


def test_run_negative_coarse_viscosity_is_an_error(gpu_ctx):
    """Both the oracle and the device refuse the fractional power of a
    negative restricted viscosity (the reference raises DomainError)."""
    case = CASES[1]
    H, G = _run_pair(gpu_ctx, **case, decades=3.0)
    with pytest.raises(ValueError):
        H.run(20, 1e-8, **RUN_KW)
    with pytest.raises(Exception, match="negative"):
        G.run(20, 1e-8, **RUN_KW)


def _system(ctx, dims, sizes, seed, contrast=2.0):
    rng = np.random.default_rng(seed)
    lv = mg3.Level(*dims, *sizes)
    for k in VISC:
        getattr(lv, k)[...] = 10 ** rng.uniform(0, contrast, getattr(lv, k).shape)
    sol = Stokes3DSolverGPU(ctx)
    sol.set_system(*dims, *sizes, *[to_device(getattr(lv, k)) for k in VISC])
    return lv, sol, rng


@pytest.mark.parametrize("case", CASES)
def test_operator_apply_bit_exact(gpu_ctx, case):
    print(f"stokes3d apply {case}", flush=True)
    lv, sol, rng = _system(gpu_ctx, **case)
    x = {k: np.asfortranarray(rng.normal(size=lv.sh[k])) for k in SOL}
    y = mg3.apply(lv, x)
    dx = sol.pack({k: to_device(x[k]) for k in SOL})
    dy = sol.vector()
    sol.apply(dx, dy)
    got = sol.split(dy)
    for k in SOL:
        assert np.array_equal(to_host(got[k]), y[k]), k


@pytest.mark.parametrize("precision,flexible,restart", [(1, 1, 30), (0, 1, 30), (1, 0, 60), (0, 0, 60)])
@pytest.mark.parametrize("case,prnorm", [(CASES[0], 0.0), (CASES[1], 1.5)])
def test_solve_matches_direct(gpu_ctx, case, prnorm, precision, flexible, restart):
    print(f"stokes3d solve {case} precision {precision} flexible {flexible}", flush=True)
    lv, sol, rng = _system(gpu_ctx, **case)
    R = {k: np.asfortranarray(rng.normal(size=lv.sh[s])) for k, s in zip(RHS, SOL)}
    A, masks = mg3.assemble(lv)
    b = mg3.pack({"vx": R["RX"], "vy": R["RY"], "vz": R["RZ"], "pr": R["RC"]}, masks)
    b[b.size - int(masks["pr"].sum())] = prnorm / lv.etan.flat[0]
    xs = mg3.unpack(spla.spsolve(A.tocsc(), b), lv, masks)
    rhs = sol.pack({s: to_device(R[k]) for k, s in zip(RHS, SOL)})
    x, st = sol.solve(rhs, prnorm=prnorm, precision=precision, rtol=1e-11, max_iterations=400,
                      flexible=flexible, restart=restart)
    print(st, flush=True)
    assert st["converged"], st
    got = sol.split(x)
    for k in SOL:
        m = masks[k]
        g = to_host(got[k])[m]
        err = np.linalg.norm(g - xs[k][m]) / np.linalg.norm(xs[k][m])
        assert err <= 1e-8, (k, err)


@pytest.mark.parametrize("n", [17, 33])
def test_solve_spheres_workload(gpu_ctx, n):
    """Inclusions of contrast 1e2-10^2.5 (the range the multigrid preconditioner
    handles; contrasts of 1e4 and more stall, see the benchmark record)."""
    print(f"stokes3d spheres n={n}", flush=True)
    w = workloads.spheres3d(n=n, seed=3, log_contrast=(2.0, 2.5), device=gpu_ctx.device)
    sol = Stokes3DSolverGPU(gpu_ctx)
    sol.set_system(*w["dims"], w["etan"], w["etaxy"], w["etaxz"], w["etayz"])
    rhs = sol.pack({"vx": w["RX"], "vy": w["RY"], "vz": w["RZ"], "pr": w["RC"]})
    xs = {}
    for precision in (1, 0):
        x, st = sol.solve(rhs, precision=precision, rtol=1e-8, max_iterations=600)
        print(precision, st, flush=True)
        assert st["converged"], st
        # independent residual: b (rhs on the unknowns, gauge 0) - A x
        ax = sol.vector()
        sol.apply(x.clone(), ax)
        b = rhs.clone()
        parts = sol.split(b)
        for k in ("vx", "vy", "vz"):
            p = parts[k]
            p[0].zero_(); p[-1].zero_(); p[:, 0].zero_(); p[:, -1].zero_(); p[:, :, 0].zero_(); p[:, :, -1].zero_()
        parts["pr"][0, 0, 0] = 0.0
        rel = (torch.linalg.vector_norm(b - ax) / torch.linalg.vector_norm(b)).item()
        assert rel <= 1.05e-8, rel
        xs[precision] = x
    d = (torch.linalg.vector_norm(xs[1] - xs[0]) / torch.linalg.vector_norm(xs[0])).item()
    assert d < 1e-6, d


def _sinker_device_hierarchy(ctx, S):
    G = Multigrid3DGPU(ctx, *S["dims"])
    for k in VISC + RHS + SOL:
        G.set(0, k, to_device(S[k]))
    return G


def test_sinker_smoke_run_matches_oracle(gpu_ctx):
    """The reference's 17^3 sinker smoke run (StokesSinker.jl:34-57: 30
    V-cycles, contrast 1e6, continuation every 15 cycles) through the
    stand-alone driver: same cycle count and flag as the oracle, solution to
    1e-9 relative, history to 1e-3 (log10; inside a continuation step the
    iteration amplifies last-bit differences of the coarse viscosity power
    law, see the 2-D test)."""
    print("mg3d sinker smoke run 17^3", flush=True)
    S = workloads.stokes_sinker3d(17)
    run = dict(workloads.SINKER_RUNS["3d_smoke"])
    n = run.pop("n")
    nv, crit = run.pop("nvcycles"), run.pop("convergence_criterion")
    kw = dict(stokesscale=S["stokesscale"], continscale=S["continscale"], pressure_bc=S["pressure_bc"],
              **workloads.SINKER_RELAXATION, **run)
    H = mg3.Hierarchy(n, n, n, 1e5, 1e5, 1e5)
    for k in VISC + RHS + SOL:
        getattr(H.levels[0], k)[...] = S[k]
    G = _sinker_device_hierarchy(gpu_ctx, S)
    n_o, c_o, h_o = H.run(nv, crit, **kw)
    n_g, c_g, h_g = G.run(nv, crit, **kw)
    assert (n_g, c_g) == (n_o, c_o) == (30, False)
    np.testing.assert_allclose(h_g, h_o, rtol=0, atol=1e-3)
    for k in SOL:
        got, exp = to_host(G.get(0, k)), getattr(H.levels[0], k)
        assert np.linalg.norm(got - exp) <= 1e-9 * np.linalg.norm(exp), k


def _sinker_stats_error(parts):
    import json
    import pathlib
    ref = json.loads((pathlib.Path(__file__).parent / "golden" / "stokes_sinker3d_97_stats.json").read_text())
    err = {}
    for k in ("vx", "vy", "vz", "pr"):
        a = parts[k]
        scale = abs(ref["pr_max"]) if k == "pr" else abs(ref[k + "_min"])
        err[k + "_min"] = abs(float(a.min()) - ref[k + "_min"]) / scale
        err[k + "_max"] = abs(float(a.max()) - ref[k + "_max"]) / abs(ref[k + "_max"])
    err["pr_avg"] = abs(float(parts["pr"].mean()) - ref["pr_avg"]) / ref["pr_avg"]
    return err


@pytest.mark.parametrize("precision", [1, 0])
def test_sinker97_solve_matches_reference_statistics(gpu_ctx, precision):
    """The reference's 97^3 Stokes sinker (contrast 1e6) solved to 1e-10 by
    GMRES(200) with the V-cycle preconditioner reproduces the final-solution
    statistics the reference's authors recorded (StokesSinker.jl:60-64):
    velocity extrema, pressure maximum and mean within 5e-8 relative,
    pressure minimum within 5e-8 of the pressure range (measured 7e-9 and
    2e-8).  This pins the 3-D operator, ghost passes and the demo's inputs
    against the reference's own output."""
    print(f"stokes3d sinker 97^3 precision {precision}", flush=True)
    S = workloads.stokes_sinker3d(97)
    sol = Stokes3DSolverGPU(gpu_ctx)
    d = lambda a: to_device(a)
    sol.set_system(*S["dims"], d(S["etan"]), d(S["etaxy"]), d(S["etaxz"]), d(S["etayz"]))
    rhs = sol.pack({"vx": d(S["RX"]), "vy": d(S["RY"]), "vz": d(S["RZ"]), "pr": d(S["RC"])})
    x, st = sol.solve(rhs, prnorm=S["pressure_bc"], precision=precision, restart=200, flexible=0,
                      rtol=1e-10, max_iterations=600)
    print(st, flush=True)
    assert st["converged"], st
    err = _sinker_stats_error({k: to_host(v) for k, v in sol.split(x).items()})
    print(err, flush=True)
    assert max(err.values()) < 5e-8, err


def test_sinker97_reference_driver_approaches_reference_statistics(gpu_ctx):
    """The reference's own 97^3 demo settings (StokesSinker.jl:64-82: 250
    V-cycles, five finest-level sweeps, continuation every 15 cycles over
    11.475 jumps, criterion 1e-20) through the stand-alone driver: all 250
    cycles run and the result is within 1e-3 of the recorded statistics
    (measured 7e-4: the plain V-cycle iteration is not fully converged after
    250 cycles at contrast 1e6)."""
    print("mg3d sinker 97^3 reference driver", flush=True)
    S = workloads.stokes_sinker3d(97)
    run = dict(workloads.SINKER_RUNS["3d"])
    run.pop("n")
    nv, crit = run.pop("nvcycles"), run.pop("convergence_criterion")
    G = _sinker_device_hierarchy(gpu_ctx, S)
    n, conv, hist = G.run(nv, crit, stokesscale=S["stokesscale"], continscale=S["continscale"],
                          pressure_bc=S["pressure_bc"], **workloads.SINKER_RELAXATION, **run)
    assert (n, conv) == (250, False) and np.all(np.isfinite(hist))
    err = _sinker_stats_error({k: to_host(G.get(0, k)) for k in SOL})
    print(err, flush=True)
    assert max(err.values()) < 1e-3, err
