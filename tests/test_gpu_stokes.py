"""GPU parity of the Stokes-continuity operator, RHS, unpack, residual and
solve, and of the reference multigrid, against the oracle (through the C ABI).

Bars: operator apply, RHS, unpack and |L S - R| bit-exact (same coefficients,
same COO summation order); solve: ||R - L S||/||R|| <= 1e-8 at the default
tolerance and velocity/pressure within 1e-8 relative L2 of the FP64 SuperLU
solve when run to rtol 1e-12; reference multigrid bit-exact in FP64.
"""
import numpy as np
import pytest
import torch

import oracle
import oracle.mg as omg
from earthbox_gpu import workloads
from earthbox_gpu.context import GPUContext
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.stokes import MultigridGPU, StokesSolverGPU, bc_mirror_factors
from stokes_cases import CASES, random_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES)
def test_apply_rhs_residual_unpack_bit_exact(gpu_ctx, case):
    print(f"stokes operator case {case}", flush=True)
    pr = random_case(**case)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    rng = np.random.default_rng(case["seed"])
    S = rng.standard_normal(pr.N) * np.tile([1e-9, 1e-9, 1e3], pr.N // 3)
    sol = StokesSolverGPU(gpu_ctx)
    sol.set_system(pr)
    dS = to_device(S)
    LS = to_host(sol.apply(dS))
    Rg = to_host(sol.rhs())
    res = to_host(sol.stokes_calc_nonlinear_system_residual(dS))
    vx1, vy1, pr1 = sol.process_stokes_solution(dS)
    gpu_ctx.synchronize()
    assert np.array_equal(LS, oracle.coo_spmv(pr.N, Li, Lj, Lv, S))
    assert np.array_equal(Rg, R)
    assert np.array_equal(res, oracle.coo_residual(pr.N, Li, Lj, Lv, S, R))
    ex, ey, ep = oracle.stokes_unpack(pr, S)
    assert np.array_equal(to_host(vx1), ex)
    assert np.array_equal(to_host(vy1), ey)
    assert np.array_equal(to_host(pr1), ep)


# default preconditioner (coarsest level solved directly) and the all-smoothed
# hierarchy of the reference's sweep rule
# default (exact coarsest level, additive tiles below the finest level), the
# four-colour tile launches, and smoothing on every level
PRECOND = [dict(), dict(additive_tiles=0), dict(coarse_unknowns=0, finest_sweeps=4, sweep_cap=48)]


@pytest.mark.parametrize("pc", PRECOND)
@pytest.mark.parametrize("precision", [1, 0])
@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[3], CASES[6]])
def test_solve_matches_direct(gpu_ctx, case, precision, pc):
    print(f"stokes solve case {case} precision {precision} {pc}", flush=True)
    pr = random_case(**case)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    Sd = oracle.direct_solve(pr.N, Li, Lj, Lv, R)
    sol = StokesSolverGPU(gpu_ctx)
    sol.set_system(pr)
    S, st = sol.solve_system(precision=precision, rtol=1e-12, max_iterations=3000,
                             bc_d=bc_mirror_factors(pr), **pc)
    print(st, flush=True)
    assert st["converged"], st
    Sg = to_host(S)
    rel = np.linalg.norm(R - oracle.coo_spmv(pr.N, Li, Lj, Lv, Sg)) / np.linalg.norm(R)
    assert rel <= 1.0001e-12
    vx, vy, p = oracle.stokes_unpack(pr, Sg)
    ex, ey, ep = oracle.stokes_unpack(pr, Sd)
    v = np.concatenate([vx.ravel(), vy.ravel()])
    ev = np.concatenate([ex.ravel(), ey.ravel()])
    assert np.linalg.norm(v - ev) <= 1e-8 * np.linalg.norm(ev)
    assert np.linalg.norm(p - ep) <= 1e-8 * np.linalg.norm(ep)


@pytest.mark.parametrize("pc", PRECOND)
@pytest.mark.parametrize("precision", [1, 0])
def test_falling_block_time_to_tolerance(gpu_ctx, precision, pc):
    """The 101 x 101 falling block (1e6 viscosity contrast): FGMRES reaches
    1e-8 relative residual; the forward error of that solve is reported."""
    pr = workloads.falling_block(101, 101)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    sol = StokesSolverGPU(gpu_ctx)
    sol.set_system(pr)
    S, st = sol.solve_system(precision=precision, rtol=1e-8, **pc)
    print(f"falling block precision={precision}: {st}", flush=True)
    assert st["converged"] and st["relres"] <= 1e-8
    Sg = to_host(S)
    rel = np.linalg.norm(R - oracle.coo_spmv(pr.N, Li, Lj, Lv, Sg)) / np.linalg.norm(R)
    assert rel <= 1.0001e-8


def test_reference_multigrid_bit_exact(gpu_ctx):
    """The device multigrid reproduces the oracle restatement of the
    reference's V-cycle bit for bit in FP64: coarse viscosities, two V-cycles
    and the final residuals."""
    pr = workloads.falling_block(45, 37, size=1.0e5, eta_block=1e23)
    g = pr.grid
    H = omg.Hierarchy(g.xnum, g.ynum, g.xsize, g.ysize)
    f = H.levels[0]
    f.etas[...] = pr.etas0
    f.etan[...] = pr.etan0
    f.RX[...] = pr.RX1
    f.RY[...] = pr.RY1
    f.RC[...] = pr.RC1
    H.restrict_viscosity()
    M = MultigridGPU(gpu_ctx, g.xnum, g.ynum, g.xsize, g.ysize)
    assert M.nlevels == len(H.levels)
    for name in ("etas", "etan", "RX", "RY", "RC"):
        M.set(0, name, to_device(getattr(f, name)))
    M.restrict_viscosity()
    gpu_ctx.synchronize()
    for k, lev in enumerate(H.levels):
        assert np.array_equal(to_host(M.get(k, "etas")), lev.etas), k
        assert np.array_equal(to_host(M.get(k, "etan")), lev.etan), k
    sm = omg.smoothing_counts(len(H.levels), 2, cap=48)
    for _ in range(2):
        H.vcycle(sm, 0.9, 0.3)
    M.vcycles(2, finest_sweeps=2, multiplier=2, cap=48, theta_s=0.9, theta_c=0.3)
    gpu_ctx.synchronize()
    for name in ("vx", "vy", "pr", "rx", "ry", "rc"):
        assert np.array_equal(to_host(M.get(0, name)), getattr(f, name)), name
    for k in range(1, len(H.levels)):
        for name in ("vx", "vy", "pr", "RX", "RY", "RC"):
            assert np.array_equal(to_host(M.get(k, name)), getattr(H.levels[k], name)), (k, name)


def _mg_setup(gpu_ctx, nx=45, ny=37, seed=3):
    pr = workloads.falling_block(nx, ny, size=1.0e5, eta_block=1e23)
    g = pr.grid
    M = MultigridGPU(gpu_ctx, g.xnum, g.ynum, g.xsize, g.ysize)
    M.set(0, "etas", to_device(pr.etas0))
    M.set(0, "etan", to_device(pr.etan0))
    rng = np.random.default_rng(seed)
    for name in ("RX", "RY", "RC", "vx", "vy", "pr"):
        a = rng.standard_normal(M.field_shape(0, name))
        a *= 1e-9 if name in ("vx", "vy") else (1e3 if name == "pr" else 1.0)
        M.set(0, name, to_device(np.asfortranarray(a)))
    M.restrict_viscosity()
    for k in range(1, M.nlevels):
        for name in ("RX", "RY", "RC"):
            M.set(k, name, to_device(np.asfortranarray(
                rng.standard_normal(M.field_shape(k, name)))))
    return M


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("bc_d", [(1, 1, 1, 1), (-1, -1, -1, -1)])
def test_tiled_smoother_single_tile_equals_global(gpu_ctx, level, bc_d):
    """One LDS tile per level: the tiled sweeps, gauge shift and residuals
    equal the global phased smoother bit for bit."""
    print(f"tiled smoother single tile level={level} bc_d={bc_d}", flush=True)
    names = ("vx", "vy", "pr", "rx", "ry", "rc")
    outs = []
    for tiled in (False, True):
        M = _mg_setup(gpu_ctx)
        if tiled:
            M.smoother_tiled(level, 5, tile=0, theta_s=0.9, theta_c=0.7, prnorm=12.5, bc_d=bc_d)
        else:
            M.smoother(level, 5, theta_s=0.9, theta_c=0.7, prnorm=12.5, bc_d=bc_d)
        gpu_ctx.synchronize()
        outs.append({n: to_host(M.get(level, n)) for n in names})
    for n in names:
        assert np.array_equal(outs[0][n], outs[1][n]), n


def test_tiled_smoother_multi_tile_deterministic(gpu_ctx):
    """Four-colour tiles (edge 8, 2 sweeps per launch): two runs are bit
    identical, and the residual norms stay within a factor 2 of those of the
    global smoother after the same number of sweeps (the block Gauss-Seidel
    variant smooths like the pointwise one)."""
    print("tiled smoother multi tile", flush=True)
    names = ("vx", "vy", "pr", "rx", "ry", "rc")
    res = []
    for _ in range(2):
        M = _mg_setup(gpu_ctx)
        M.smoother_tiled(0, 8, local_sweeps=2, tile=8, theta_s=0.9, theta_c=0.7)
        gpu_ctx.synchronize()
        res.append({n: to_host(M.get(0, n)) for n in names})
    for n in names:
        assert np.array_equal(res[0][n], res[1][n]), n
    M = _mg_setup(gpu_ctx)
    M.smoother(0, 8, theta_s=0.9, theta_c=0.7)
    gpu_ctx.synchronize()
    for n in ("rx", "ry", "rc"):
        a, b = np.linalg.norm(res[0][n]), np.linalg.norm(to_host(M.get(0, n)))
        assert 0.5 * b <= a <= 2.0 * b, (n, a, b)


@pytest.mark.parametrize("smoother", [0, 1])
def test_solve_smoother_variants_match_direct(gpu_ctx, smoother):
    """Both preconditioner smoothers (global phased, LDS-tiled) reach the
    direct solution on the falling block with a no-slip box."""
    print(f"stokes solve smoother={smoother}", flush=True)
    pr = random_case(**CASES[1])
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    Sd = oracle.direct_solve(pr.N, Li, Lj, Lv, R)
    sol = StokesSolverGPU(gpu_ctx)
    sol.set_system(pr)
    S, st = sol.solve_system(precision=1, rtol=1e-12, max_iterations=3000, smoother=smoother,
                             bc_d=bc_mirror_factors(pr))
    print(st, flush=True)
    assert st["converged"], st
    Sg = to_host(S)
    assert np.linalg.norm(Sg - Sd) <= 1e-8 * np.linalg.norm(Sd)


@pytest.mark.parametrize("case", CASES)
def test_grid_residuals_bit_exact(gpu_ctx, case):
    """compute_stokes_residuals! on random fields: bit-exact vs the oracle."""
    print(f"grid residuals case {case}", flush=True)
    pr = random_case(**case)
    g = pr.grid
    rng = np.random.default_rng(case["seed"] + 50)
    vx = np.asfortranarray(rng.standard_normal((g.ynum + 1, g.xnum)) * 1e-9)
    vy = np.asfortranarray(rng.standard_normal((g.ynum, g.xnum + 1)) * 1e-9)
    p = np.asfortranarray(rng.standard_normal((g.ynum - 1, g.xnum - 1)) * 1e6)
    ex, ey, ec = oracle.stokes_grid_residuals(pr, vx, vy, p)
    sol = StokesSolverGPU(gpu_ctx)
    sol.set_system(pr)
    rx, ry, rc = sol.compute_stokes_residuals(to_device(vx), to_device(vy), to_device(p))
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(rx), ex)
    assert np.array_equal(to_host(ry), ey)
    assert np.array_equal(to_host(rc), ec)


def test_change_norms_and_timestep(gpu_ctx):
    from earthbox_gpu.stokes import calculate_new_timestep, change_norms
    from earthbox_gpu.grid import uniform_grid
    print("change norms and time step", flush=True)
    rng = np.random.default_rng(4)
    a, ao = rng.standard_normal(100003), rng.standard_normal(100003)
    b, bo = rng.standard_normal(777), rng.standard_normal(777)
    da, dao, db, dbo = (to_device(t) for t in (a, ao, b, bo))
    out = change_norms(gpu_ctx, [da, db], [dao, dbo])
    v, vo = np.concatenate([a, b]), np.concatenate([ao, bo])
    l2, rl2 = oracle.get_L2_norms(v, vo)
    inf, rinf = oracle.get_inf_norms(v, vo)
    assert out[0] == pytest.approx(l2, rel=1e-13) and out[1] == pytest.approx(rl2, rel=1e-13)
    assert out[2] == inf and out[3] == rinf
    assert change_norms(gpu_ctx, da)[0] == pytest.approx(np.linalg.norm(a), rel=1e-13)
    g = uniform_grid(41, 23, 4.0e5, 2.2e5)
    gpu_ctx.set_grid(g)
    vx1 = np.asfortranarray(rng.standard_normal((g.ynum + 1, g.xnum)) * 1e-9)
    vy1 = np.asfortranarray(rng.standard_normal((g.ynum, g.xnum + 1)) * 2e-9)
    got = calculate_new_timestep(gpu_ctx, to_device(vx1), to_device(vy1), 1e16, 0.5)
    assert got == oracle.calculate_new_timestep(1e16, 0.5, g.xstpavg, g.ystpavg, vx1, vy1)


def test_system_invalidated_when_grid_changes(gpu_ctx):
    """The operator reads the context's grid arrays, so setting another grid
    invalidates the system instead of leaving it pointing at freed arrays."""
    from earthbox_gpu._lib import EBGError
    print("stale system after grid change", flush=True)
    pr = random_case(**CASES[0])
    sol = StokesSolverGPU(gpu_ctx)
    sol.set_system(pr)
    S = to_device(np.ones(pr.N))
    sol.apply(S)
    gpu_ctx.set_grid(random_case(**CASES[1]).grid)
    with pytest.raises(EBGError, match="grid changed"):
        sol.apply(S)
    sol.set_system(pr)  # re-setting the system makes it usable again
    sol.apply(S)
    gpu_ctx.synchronize()


@pytest.mark.parametrize("eta_block,kw,hist_tol", [
    (1e26, dict(nvcycles=40, criterion=1e-5, nviscosity_jumps=12.0, nvcycles_viscosity_jump=17), 1e-3),
    (1e22, dict(nvcycles=300, criterion=1e-9, nviscosity_jumps=2.0, nvcycles_viscosity_jump=5), 1e-7),
])
def test_reference_multigrid_run_matches_oracle(gpu_ctx, eta_block, kw, hist_tol):
    """The reference's stand-alone 2-D driver on its own sinker demo (33^2):
    same cycle count and stop flag as the oracle, solution and scaled
    finest-level viscosity to 1e-9 / 1e-13 relative, residual history within
    hist_tol (log10 units).  V-cycles are bit-exact; only the power law of the
    coarse-level viscosity scaling (last bit) and the RMS summation order
    differ.  At contrast 1e6 the reference's own iteration amplifies such
    differences inside a continuation step: moving one coarse viscosity of
    the oracle by one ulp changes its history by up to 2e-2 there, so that
    case is held to 1e-3 (the device measures 5e-5)."""
    print(f"mg2d run sinker eta_block {eta_block} {kw}", flush=True)
    n = 33
    S = workloads.stokes_sinker2d(n, viscosity_block=eta_block)
    H = omg.Hierarchy(n, n, 1e5, 1e5)
    f = H.levels[0]
    M = MultigridGPU(gpu_ctx, n, n, 1e5, 1e5)
    for k in ("etas", "etan", "RX", "RY", "RC", "vx", "vy", "pr"):
        getattr(f, k)[...] = S[k]
        M.set(0, k, to_device(S[k]))
    kw = dict(kw)
    nv, crit = kw.pop("nvcycles"), kw.pop("criterion")
    common = dict(stokesscale=S["stokesscale"], continscale=S["continscale"], finest_sweeps=5,
                  cap=48, theta_s=0.9, theta_c=0.3, pressure_bc=0.0, **kw)
    n_o, c_o, h_o = H.run(nv, crit, **common)
    n_g, c_g, h_g = M.run(nv, crit, **common)
    assert (n_g, c_g) == (n_o, c_o), (n_g, c_g, n_o, c_o)
    np.testing.assert_allclose(h_g, h_o, rtol=0, atol=hist_tol)
    for k in ("vx", "vy", "pr"):
        got, exp = to_host(M.get(0, k)), getattr(f, k)
        assert np.linalg.norm(got - exp) <= 1e-9 * np.linalg.norm(exp), k
    for k in ("etas", "etan"):
        np.testing.assert_allclose(to_host(M.get(0, k)), getattr(f, k), rtol=1e-13)


@pytest.mark.parametrize("eta_block,crit,hist_tol", [(1e22, 1e-9, 1e-7), (1e26, 1e-10, 1e-2)])
def test_reference_multigrid_multimulti_matches_oracle(gpu_ctx, eta_block, crit, hist_tol):
    """The multimulti driver (the reference's run_stokes_sinker_2d_multimulti.jl
    settings: 3 cycles per jump, 5.9999 jumps) on the 33^2 sinker: same
    cycle count, stop flag and number of global updates as the oracle,
    principal residual history within hist_tol and global residuals within
    1e-6 (log10 units; 5e-2 at contrast 1e6), global solution to 1e-9
    relative.  As for the plain driver, only the
    coarse-level power law and the RMS summation order differ."""
    print(f"mg2d multimulti sinker eta_block {eta_block} criterion {crit}", flush=True)
    n = 33
    S = workloads.stokes_sinker2d(n, viscosity_block=eta_block)
    H = omg.Hierarchy(n, n, 1e5, 1e5)
    f = H.levels[0]
    M = MultigridGPU(gpu_ctx, n, n, 1e5, 1e5)
    for k in ("etas", "etan", "RX", "RY", "RC", "vx", "vy", "pr"):
        getattr(f, k)[...] = S[k]
        M.set(0, k, to_device(S[k]))
    common = dict(stokesscale=S["stokesscale"], continscale=S["continscale"], finest_sweeps=5, cap=48,
                  nviscosity_jumps=5.9999, nvcycles_viscosity_jump=3)
    n_o, c_o, h_o, g_o = H.run_multimulti(600, crit, **common)
    n_g, c_g, h_g, g_g = M.run_multimulti(600, crit, **common)
    print(f"  oracle {n_o} cycles, {len(g_o)} global updates; device {n_g} cycles, {len(g_g)} updates", flush=True)
    assert (n_g, c_g, len(g_g)) == (n_o, c_o, len(g_o))
    assert c_o and len(g_o) >= 2
    np.testing.assert_allclose(h_g, h_o, rtol=0, atol=hist_tol)
    for k in ("vx", "vy", "pr"):
        got, exp = to_host(M.get(0, k)), getattr(H.level0, k)
        err = np.linalg.norm(got - exp) / np.linalg.norm(exp)
        print(f"  global solution {k}: relative L2 difference {err:.2e}", flush=True)
        assert err <= 1e-9, k
    # the global residuals are level 0's residuals at 1e-4 .. 1e-10 of the
    # scale, a difference of nearly equal terms: a solution difference of
    # 1e-12 moves them by percent at contrast 1e6
    print(f"  global residuals: max difference {np.abs(g_g - g_o).max():.2e} (log10)", flush=True)
    np.testing.assert_allclose(g_g, g_o, rtol=0, atol=1e-6 if hist_tol < 1e-3 else 5e-2)


@pytest.mark.parametrize("bc_d", [(1.0, 1.0, 1.0, 1.0), (-1.0, -1.0, -1.0, -1.0), (1.0, -1.0, -1.0, 1.0)])
def test_direct_level_solve_zeroes_the_reference_residuals(gpu_ctx, bc_d):
    """The dense direct solve the preconditioner uses on its coarsest level
    solves the reference smoother's equations: after it, the oracle's ghost
    pass and residuals (BoundaryConditions.jl, Residuals.jl restated) vanish
    on every unknown except the continuity row of cell (1,1), which the gauge
    p(1,1) = 0 replaces."""
    print(f"mg2d direct level solve bc_d={bc_d}", flush=True)
    pr = workloads.falling_block(45, 37, size=1.0e5, eta_block=1e23)
    g = pr.grid
    H = omg.Hierarchy(g.xnum, g.ynum, g.xsize, g.ysize, bc_d=bc_d)
    f = H.levels[0]
    f.etas[...] = pr.etas0
    f.etan[...] = pr.etan0
    H.restrict_viscosity()
    M = MultigridGPU(gpu_ctx, g.xnum, g.ynum, g.xsize, g.ysize)
    M.set(0, "etas", to_device(f.etas))
    M.set(0, "etan", to_device(f.etan))
    M.restrict_viscosity()
    rng = np.random.default_rng(4)
    L = omg._sigs()
    for level in (1, 2):
        lv = H.levels[level]
        for name in ("RX", "RY", "RC"):
            a = getattr(lv, name)
            a[...] = rng.standard_normal(a.shape) * (1e-12 if name == "RC" else 1.0)
            M.set(level, name, to_device(a))
        M.direct_solve(level, bc_d)
        gpu_ctx.synchronize()
        for name in ("vx", "vy", "pr"):
            getattr(lv, name)[...] = to_host(M.get(level, name))
        assert lv.pr[0, 0] == 0.0
        L.ebo_mg_set_bc(omg.C.byref(lv.cs()), *bc_d)
        L.ebo_mg_residuals(omg.C.byref(lv.cs()))
        ny, nx = lv.ynum, lv.xnum
        rx = lv.rx[1:ny, 1:nx - 1]
        ry = lv.ry[1:ny - 1, 1:nx]
        rc = lv.rc.copy()
        rc[0, 0] = 0.0
        # the level operator is badly scaled (viscosity 1e21-1e23 against unit
        # pressure; condition ~1e40): an FP64 LU solve of the same equations in
        # numpy leaves residuals of ~1e-7 of the right-hand side; a wrong
        # operator would leave residuals of its order
        assert np.abs(rx).max() <= 1e-4 * np.abs(lv.RX).max(), (level, np.abs(rx).max())
        assert np.abs(ry).max() <= 1e-4 * np.abs(lv.RY).max(), (level, np.abs(ry).max())
        assert np.abs(rc).max() <= 1e-4 * np.abs(lv.RC).max(), (level, np.abs(rc).max())


@pytest.mark.parametrize("pc", [dict(), dict(additive_tiles=0), dict(coarse_unknowns=0, finest_sweeps=4, sweep_cap=48)])
@pytest.mark.parametrize("precision", [1, 0])
def test_captured_vcycle_bit_identical(gpu_ctx, precision, pc):
    """The V-cycle replayed from its captured HIP graph against the same
    V-cycle launched kernel by kernel: 60 FGMRES iterations (60 replays with
    the Krylov kernels in between) give a bit-identical iterate and residual
    history, on two systems in a row on the same hierarchy (new viscosities,
    the graph replayed again); the 101 x 101 falling block converges to 1e-8
    identically."""
    print(f"captured V-cycle precision={precision} {pc}", flush=True)
    out = []
    try:
        for graphs in (True, False):
            gpu_ctx.set_stokes_graphs(graphs)
            res = []
            for seed in (3, 4):
                pr = workloads.layered_random(129, 65, 1.5e6 * 128 / 1000, 6.0e5 * 64 / 500, seed=seed,
                                              amplitude=0.5)
                sol = StokesSolverGPU(gpu_ctx)
                sol.set_system(pr)
                S, st = sol.solve_system(precision=precision, rtol=1e-14, max_iterations=60,
                                         bc_d=bc_mirror_factors(pr), **pc)
                gpu_ctx.synchronize()
                res.append((to_host(S), st["iterations"], st["relres"]))
            pr = workloads.falling_block(101, 101)
            sol = StokesSolverGPU(gpu_ctx)
            sol.set_system(pr)
            S, st = sol.solve_system(precision=precision, rtol=1e-8, **pc)
            gpu_ctx.synchronize()
            assert st["converged"], st
            res.append((to_host(S), st["iterations"], st["relres"]))
            out.append(res)
    finally:
        gpu_ctx.set_stokes_graphs(True)
    for (a, ia, ra), (b, ib, rb) in zip(out[0], out[1]):
        assert ia == ib and ra == rb
        assert np.array_equal(a, b)


def test_arnoldi_status_one_step_behind(gpu_ctx):
    """The host reads each Arnoldi step's status one step late, with the next
    step already queued: the reported count is still the first step that
    meets the tolerance (one fewer iteration does not converge and stops at
    exactly that count), the step queued past convergence leaves the solution
    untouched (bit-identical to a solve capped at the converged count), and
    the reported residuals equal the ones recomputed from the solutions."""
    print("lagged Arnoldi status", flush=True)
    pr = workloads.layered_random(129, 65, 1.5e6 * 128 / 1000, 6.0e5 * 64 / 500, seed=5, amplitude=0.5)
    sol = StokesSolverGPU(gpu_ctx)
    sol.set_system(pr)
    kw = dict(precision=1, rtol=1e-8, restart=100, bc_d=bc_mirror_factors(pr))
    S, st = sol.solve_system(**kw)
    print(f"  {st['iterations']} iterations, {st['restarts']} cycles", flush=True)
    assert st["converged"], st
    n = st["iterations"]
    S1, st1 = sol.solve_system(max_iterations=n - 1, **kw)
    assert not st1["converged"] and st1["iterations"] == n - 1 and st1["relres"] > 1e-8
    S2, st2 = sol.solve_system(max_iterations=n, **kw)
    assert st2["converged"] and st2["iterations"] == n
    assert np.array_equal(to_host(S2), to_host(S))
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    import scipy.sparse as sp
    L = sp.csr_matrix((Lv, (Li - 1, Lj - 1)), shape=(pr.N, pr.N))
    for x, s in ((to_host(S), st), (to_host(S1), st1)):
        rel = np.linalg.norm(R - L @ x) / np.linalg.norm(R)
        assert abs(rel - s["relres"]) <= 1e-6 * s["relres"], (rel, s["relres"])


def test_arnoldi_short_restart_cycles(gpu_ctx):
    """Restart cycles of 10 steps with the device-side Hessenberg matrix and
    rotations: each cycle minimises the residual over its space, so the true
    residual after 10, 20, 30 and 40 iterations never grows (up to rounding);
    every capped solve stops at exactly its cap, and the residual it reports
    is the one recomputed from its solution."""
    print("Arnoldi restart cycles of 10", flush=True)
    pr = workloads.layered_random(129, 65, 1.5e6 * 128 / 1000, 6.0e5 * 64 / 500, seed=5, amplitude=0.5)
    sol = StokesSolverGPU(gpu_ctx)
    sol.set_system(pr)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    import scipy.sparse as sp
    L = sp.csr_matrix((Lv, (Li - 1, Lj - 1)), shape=(pr.N, pr.N))
    prev = np.inf
    for cap in (10, 20, 30, 40):
        S, st = sol.solve_system(precision=1, rtol=1e-14, restart=10, max_iterations=cap,
                                 bc_d=bc_mirror_factors(pr))
        rel = np.linalg.norm(R - L @ to_host(S)) / np.linalg.norm(R)
        print(f"  {cap} iterations: relres {st['relres']:.3e}", flush=True)
        assert st["iterations"] == cap and st["restarts"] == cap // 10
        assert abs(rel - st["relres"]) <= 1e-6 * st["relres"]
        assert st["relres"] <= prev * (1 + 1e-9)
        prev = st["relres"]
