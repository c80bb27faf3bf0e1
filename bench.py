#!/usr/bin/env python3
"""Benchmark of the EarthBox per-timestep hot path on one MI355X.

Workload (1001 x 501 nodes over 1500 x 600 km, 6 x 6 markers per cell =
18,000,000 markers, FP64): the layered random-viscosity Stokes system of that
grid is solved first (FP64 FGMRES to 1e-8), and its velocity field with the
time step of half a cell of displacement (the Courant limit 0.5) drives the
markers.  One step = one fused RK4 gather-advect pass of all markers
(velocity at 4 stages, position update) from the markers' initial positions:
every timed step advects its own copy of those positions, prepared in HBM
before the timed region.  Tolerance arithmetic by default (positions within
1e-10 of the domain size; the exact, bit-identical kernel is timed beside it).
Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import platform
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "earthbox.jl_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
RK4_BYTES_PER_MARKER = 32.0  # read x, y; write x, y (FP64)
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on the
# 1001 x 501-node, 6 x 6 markers-per-cell workload (18M markers), each counter
# scaled by its reading for a plain 512 MiB copy, keyed by kernel and workload.
RK4_MEASURED_TRAFFIC = {
    ("rk4_tol_kernel", 1001, 501, 6): 640.4e6,      # 352.4 MB read + 288.0 MB written
    ("rk4_advect_kernel", 1001, 501, 6): 641.3e6,   # 353.4 MB read + 288.0 MB written
}


def progress(msg):
    """One progress line on stderr (stdout carries only the JSON result)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--xnum", type=int, default=1001)
    ap.add_argument("--ynum", type=int, default=501)
    ap.add_argument("--markers-per-cell", type=int, default=6)
    ap.add_argument("--rk4-arithmetic", choices=("tolerance", "exact"), default="tolerance")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three plain 512 MiB device copies (the reference kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--no-stokes", action="store_true")
    ap.add_argument("--no-scatter", action="store_true")
    ap.add_argument("--stokes-max-iter", type=int, default=3000)
    ap.add_argument("--no-stokes3d", action="store_true")
    ap.add_argument("--stokes3d-n", type=int, default=257)
    ap.add_argument("--stokes3d-max-iter", type=int, default=300)
    ap.add_argument("--no-sinker", action="store_true")
    ap.add_argument("--no-picard", action="store_true")
    ap.add_argument("--no-grid-kernels", action="store_true")
    ap.add_argument("--no-marker-gathers", action="store_true")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--no-config5", action="store_true")
    ap.add_argument("--config5-steps", type=int, default=20)
    ap.add_argument("--no-heat", action="store_true")
    ap.add_argument("--no-cnn", action="store_true")
    ap.add_argument("--stokes-amplitude", type=float, default=1.0,
                    help="decades of the random log-viscosity perturbation of the 1001 x 501 system")
    return ap.parse_args()


def bench_stokes3d(ctx, n, max_iter):
    """3-D inclusion problem on n^3 nodes (unit cube, 8 seeded spheres of
    radius 0.05-0.12, density 1 +- 1 %, body force in z, free slip): time to
    ||b - A x|| / ||b|| <= 1e-8 (FP64) with the FP32 and the FP64 V-cycle as
    preconditioner, at the inclusion contrast range 1e2-1e6 and at contrast 1e2."""
    import torch
    from earthbox_gpu import workloads
    from earthbox_gpu.stokes3d import Stokes3DSolverGPU
    out = {"rtol": 1e-8, "n": n}
    # contrast 1e2: FGMRES(30); contrast 1e2-1e6: GMRES(200) keeping only the
    # basis (the stiff spheres' near-rigid modes need the long restart)
    for tag, lc, cap, kry in (("contrast_1e2", (2.0, 2.0 + 1e-9), max_iter, dict(restart=30, flexible=1)),
                              ("contrast_1e2_1e6", (2.0, 6.0), 2 * max_iter, dict(restart=200, flexible=0))):
        w = workloads.spheres3d(n=n, seed=3, log_contrast=lc, device=ctx.device)
        sol = Stokes3DSolverGPU(ctx)
        sol.set_system(*w["dims"], w["etan"], w["etaxy"], w["etaxz"], w["etayz"])
        rhs = sol.pack({"vx": w["RX"], "vy": w["RY"], "vz": w["RZ"], "pr": w["RC"]})
        import math
        leg = {"sphere_log10_contrast": [round(math.log10(float(v)), 3) for v in w["spheres"]["eta"]],
               "krylov": kry}
        for name, prec in (("mixed_fp32_mg", 1), ("plain_fp64_mg", 0)):
            sol.solve(rhs, precision=prec, rtol=1e-8, max_iterations=2, **kry)   # build + warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, st = sol.solve(rhs, precision=prec, rtol=1e-8, max_iterations=cap, **kry)
            torch.cuda.synchronize()
            leg[name] = {"seconds": time.perf_counter() - t0, "iterations": st["iterations"],
                         "relres": st["relres"], "converged": st["converged"]}
            del x
        if leg["mixed_fp32_mg"]["converged"] and leg["plain_fp64_mg"]["converged"]:
            leg["speedup_mixed_vs_plain"] = leg["plain_fp64_mg"]["seconds"] / leg["mixed_fp32_mg"]["seconds"]
        out[tag] = leg
        del sol, rhs, w
        torch.cuda.empty_cache()
    return out


def _sinker_stats_error(parts):
    """Largest relative deviation from the solution statistics recorded in the
    reference's 97^3 sinker demo (pressure minimum relative to the maximum)."""
    ref = json.loads((ROOT / "tests" / "golden" / "stokes_sinker3d_97_stats.json").read_text())
    err = 0.0
    for k in ("vx", "vy", "vz", "pr"):
        a = parts[k]
        scale = abs(ref["pr_max"]) if k == "pr" else abs(ref[k + "_min"])
        err = max(err, abs(float(a.min()) - ref[k + "_min"]) / scale,
                  abs(float(a.max()) - ref[k + "_max"]) / abs(ref[k + "_max"]))
    return max(err, abs(float(parts["pr"].mean()) - ref["pr_avg"]) / ref["pr_avg"])


def bench_sinker(ctx):
    """The reference's own multigrid demo (Stokes sinker, contrast 1e6):
    3-D 97^3 through the stand-alone V-cycle driver with the demo's settings
    (250 cycles) and through GMRES(200) + V-cycle to 1e-10; 2-D 97^2 through
    the driver with its settings (up to 500 cycles, criterion 1e-5).  Each
    3-D result carries its largest deviation from the statistics the
    reference recorded for that demo."""
    import torch
    from earthbox_gpu import workloads
    from earthbox_gpu.device import to_device, to_host
    from earthbox_gpu.stokes import MultigridGPU
    from earthbox_gpu.stokes3d import Multigrid3DGPU, Stokes3DSolverGPU
    out = {}
    S = workloads.stokes_sinker3d(97)
    run = dict(workloads.SINKER_RUNS["3d"])
    run.pop("n")
    nv, crit = run.pop("nvcycles"), run.pop("convergence_criterion")
    kw = dict(stokesscale=S["stokesscale"], continscale=S["continscale"], pressure_bc=S["pressure_bc"],
              **workloads.SINKER_RELAXATION, **run)
    G = Multigrid3DGPU(ctx, *S["dims"])
    names = ("etan", "etaxy", "etaxz", "etayz", "RX", "RY", "RZ", "RC", "vx", "vy", "vz", "pr")
    for k in names:
        G.set(0, k, to_device(S[k]))
    G.run(2, crit, **kw)   # warm: first-run allocations and code paths
    for k in names:
        G.set(0, k, to_device(S[k]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n, conv, hist = G.run(nv, crit, **kw)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    out["mg3d_97_reference_driver"] = {
        "seconds": sec, "cycles": n, "ms_per_cycle": 1e3 * sec / n, "converged": conv,
        "final_log10_residuals": [float(v) for v in hist[-1]],
        "max_rel_dev_from_reference_stats": _sinker_stats_error({k: to_host(G.get(0, k)) for k in
                                                                 ("vx", "vy", "vz", "pr")})}
    del G
    sol = Stokes3DSolverGPU(ctx)
    d = lambda a: to_device(a)
    sol.set_system(*S["dims"], d(S["etan"]), d(S["etaxy"]), d(S["etaxz"]), d(S["etayz"]))
    rhs = sol.pack({"vx": d(S["RX"]), "vy": d(S["RY"]), "vz": d(S["RZ"]), "pr": d(S["RC"])})
    for name, prec in (("mixed_fp32_mg", 1), ("plain_fp64_mg", 0)):
        kry = dict(restart=200, flexible=0, rtol=1e-10)
        sol.solve(rhs, precision=prec, max_iterations=2, **kry)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, st = sol.solve(rhs, prnorm=S["pressure_bc"], precision=prec, max_iterations=600, **kry)
        torch.cuda.synchronize()
        out[f"stokes3d_97_gmres_{name}"] = {
            "seconds": time.perf_counter() - t0, "iterations": st["iterations"], "relres": st["relres"],
            "converged": st["converged"],
            "max_rel_dev_from_reference_stats": _sinker_stats_error({k: to_host(v) for k, v in
                                                                     sol.split(x).items()})}
        del x
    del sol, rhs
    torch.cuda.empty_cache()
    S2 = workloads.stokes_sinker2d(97)
    run = dict(workloads.SINKER_RUNS["2d"])
    run.pop("n")
    nv, crit = run.pop("nvcycles"), run.pop("convergence_criterion")
    kw2 = dict(stokesscale=S2["stokesscale"], continscale=S2["continscale"], pressure_bc=S2["pressure_bc"],
               **workloads.SINKER_RELAXATION, **run)
    M = MultigridGPU(ctx, 97, 97, 1e5, 1e5)
    for rep in range(2):   # warm-up run, then the timed one
        for k in ("etas", "etan", "RX", "RY", "RC", "vx", "vy", "pr"):
            M.set(0, k, to_device(S2[k]))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n, conv, hist = M.run(nv if rep else 2, crit, **kw2)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
    out["mg2d_97_reference_driver"] = {"seconds": sec, "cycles": n, "ms_per_cycle": 1e3 * sec / n,
                                       "converged": conv, "final_log10_residuals": [float(v) for v in hist[-1]]}
    # the same 2-D system through the FGMRES solve (its equations are the
    # multigrid's: free slip, no stabilization), both answers judged by the
    # reference's grid residuals (StokesResiduals.jl:94-137)
    import numpy as np
    from earthbox_gpu.grid import uniform_grid
    from earthbox_gpu.stokes import StokesSolverGPU
    from earthbox_gpu.stokes_problem import StokesProblem, free_slip_bcs
    g2 = uniform_grid(97, 97, 1e5, 1e5)
    prob = StokesProblem(grid=g2, etas0=S2["etas"], etan0=S2["etan"], rho1_vy=np.zeros((97, 98), order="F"),
                         RX1=S2["RX"], RY1=S2["RY"], RC1=S2["RC"], bcs=free_slip_bcs(g2), pressure_bc_mode=0,
                         pressure_bc=S2["pressure_bc"], timestep=0.0, gravity_y=workloads.SINKER["gravity"],
                         iuse_interface_stabilization=0)
    ctx.set_grid(g2)
    sol2 = StokesSolverGPU(ctx)
    sol2.set_system(prob)
    rnorm = float(np.sqrt((S2["RX"] ** 2).sum() + (S2["RY"] ** 2).sum()))

    def momentum_relres(vx1, vy1, pr1):
        rx, ry, _ = sol2.compute_stokes_residuals(vx1, vy1, pr1)
        return float(torch.sqrt((rx ** 2).sum() + (ry ** 2).sum())) / rnorm

    out["mg2d_97_reference_driver"]["momentum_relres"] = momentum_relres(M.get(0, "vx"), M.get(0, "vy"),
                                                                         M.get(0, "pr"))
    for name, prec in (("mixed_fp32_mg", 1), ("plain_fp64_mg", 0)):
        sol2.solve_system(rtol=1e-8, max_iterations=3, precision=prec)   # build + warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol2.set_system(prob)
        S, st = sol2.solve_system(rtol=1e-8, max_iterations=2000, precision=prec)
        torch.cuda.synchronize()
        sec2 = time.perf_counter() - t0
        out[f"stokes2d_97_fgmres_{name}"] = {
            "seconds_with_setup": sec2, "iterations": st["iterations"], "relres": st["relres"],
            "converged": st["converged"], "momentum_relres": momentum_relres(*sol2.process_stokes_solution(S)),
            "speedup_vs_reference_driver": sec / sec2}
    return out


def cpu_sinker_driver(cycles: int = 3):
    """CPU time per V-cycle of the oracle's serial C restatement of the
    stand-alone 3-D driver on the 97^3 sinker (after the first, set-up
    cycle), for the driver leg of the sinker benchmark."""
    from earthbox_gpu import workloads
    from oracle import mg3
    S = workloads.stokes_sinker3d(97)
    run = dict(workloads.SINKER_RUNS["3d"])
    run.pop("n")
    run.pop("nvcycles")
    crit = run.pop("convergence_criterion")
    kw = dict(stokesscale=S["stokesscale"], continscale=S["continscale"], pressure_bc=S["pressure_bc"],
              **workloads.SINKER_RELAXATION, **run)

    def timed(nv):
        H = mg3.Hierarchy(97, 97, 97, 1e5, 1e5, 1e5)
        for k in ("etan", "etaxy", "etaxz", "etayz", "RX", "RY", "RZ", "RC", "vx", "vy", "vz", "pr"):
            getattr(H.levels[0], k)[...] = S[k]
        t0 = time.perf_counter()
        H.run(nv, crit, **kw)
        return time.perf_counter() - t0
    per = (timed(1 + cycles) - timed(1)) / cycles
    return {"seconds_per_cycle": per, "seconds_250_cycles_extrapolated": 250 * per, "cores": 1,
            "kind": "port", "sample": f"{cycles} V-cycles of the 97^3 sinker (oracle C restatement)"}


def bench_stokes(ctx, nx, ny, max_iter, amplitude=1.0):
    """Time to ||R - L S|| / ||R|| <= 1e-8 (FP64) of the layered random
    log-viscosity system (18-25, 1500 x 600 km, perturbation of `amplitude`
    decades) for three preconditioners of the FP64 FGMRES(100): the FP64
    V-cycle, the FP32 V-cycle, and the FP32 V-cycle with the learned
    correction judged by the FP64 residual.  Times include the per-system
    set-up (new fields, viscosity restriction, coarsest-level inverse); the
    solve alone is reported beside them."""
    import torch
    from earthbox_gpu import workloads
    from earthbox_gpu.stokes import StokesSolverGPU, bc_mirror_factors
    prob = workloads.layered_random(nx, ny, 1.5e6 * (nx - 1) / 1000, 6.0e5 * (ny - 1) / 500, seed=2,
                                    amplitude=amplitude)
    sol = StokesSolverGPU(ctx)
    sol.set_system(prob)
    bcd = bc_mirror_factors(prob)
    out = {"system": f"layered random log10(eta) in [18,25] (+-{amplitude} decades), {nx}x{ny} nodes, "
                     f"N={prob.N}", "rtol": 1e-8}
    from earthbox_gpu import learned
    net = learned.LearnedCorrection(ctx)
    if learned.WEIGHTS_FILE.exists():
        net.load()
        out["learned_weights"] = learned.WEIGHTS_FILE.name
    legs = (("plain_fp64_mg", dict(precision=0)), ("mixed_fp32_mg", dict(precision=1)),
            ("learned_mixed", dict(precision=1, learned=1)))
    from earthbox_gpu.device import to_device
    dev = {k: to_device(getattr(prob, k)) for k in ("etas0", "etan0", "rho1_vy", "RX1", "RY1", "RC1")}
    for name, kw in legs:
        sol.solve_system(rtol=1e-8, max_iterations=3, bc_d=bcd, **kw)  # build + warm
        # per-system set-up: a new system (device fields) and the preconditioner
        # it needs (viscosity restriction, coarsest-level inverse), no iterations
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol.set_system(prob, fields=dev)
        sol.solve_system(rtol=1e-8, max_iterations=0, bc_d=bcd, **kw)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        S, st = sol.solve_system(rtol=1e-8, max_iterations=max_iter, bc_d=bcd, **kw)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[name] = {"seconds": el, "iterations": st["iterations"], "relres": st["relres"],
                     "converged": st["converged"], "learned_accepted": st["learned_accepted"],
                     "learned_evaluated": st["learned_evaluated"],
                     "setup_seconds": setup, "seconds_with_setup": el + setup}
    p, m, lm = out["plain_fp64_mg"], out["mixed_fp32_mg"], out["learned_mixed"]
    if p["converged"] and m["converged"]:
        out["speedup_mixed_vs_plain"] = p["seconds_with_setup"] / m["seconds_with_setup"]
        out["speedup_mixed_vs_plain_solve_only"] = p["seconds"] / m["seconds"]
    if p["converged"] and lm["converged"]:
        out["speedup_learned_mixed_vs_plain"] = p["seconds_with_setup"] / lm["seconds_with_setup"]
        out["speedup_learned_mixed_vs_plain_solve_only"] = p["seconds"] / lm["seconds"]
    return out


def config2_velocity(ctx, nx, ny):
    """The velocity field of the layered random-viscosity system on the
    benchmark grid (1.0-decade perturbation, FP64 FGMRES with the FP32 V-cycle
    to 1e-8), unpacked with its ghost BCs, and the time step of half a cell of
    displacement (Update.jl:43-61 with marker_cell_displ_max = 0.5)."""
    from earthbox_gpu import workloads
    from earthbox_gpu.stokes import StokesSolverGPU, bc_mirror_factors, calculate_new_timestep
    prob = workloads.layered_random(nx, ny, 1.5e6 * (nx - 1) / 1000, 6.0e5 * (ny - 1) / 500, seed=2,
                                    amplitude=1.0)
    sol = StokesSolverGPU(ctx)
    sol.set_system(prob)
    S, st = sol.solve_system(rtol=1e-8, max_iterations=3000, bc_d=bc_mirror_factors(prob), precision=1)
    vx1, vy1, _ = sol.process_stokes_solution(S)
    dt, _ = calculate_new_timestep(ctx, vx1, vy1, 1.0e30, 0.5)
    field = {"system": "layered random log10(eta) in [18,25] (+-1.0 decades)", "relres": st["relres"],
             "iterations": st["iterations"], "dt_s": dt}
    return prob.grid, vx1, vy1, dt, field


def bench_config1(ctx):
    """The 101 x 101 falling block (contrast 1e6): GPU time to 1e-8 and 1e-10
    (set-up included) with each preconditioner, the forward error of each
    solution against the FP64 sparse direct solve (SuperLU + 3 refinement
    steps, standing in for MUMPS; velocity, pressure outside the block and
    inside it), and that direct solve's time on one host core."""
    import numpy as np
    import torch
    import oracle
    from earthbox_gpu import workloads
    from earthbox_gpu.device import to_host
    from earthbox_gpu.stokes import StokesSolverGPU
    pr = workloads.falling_block(101, 101)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    t0 = time.perf_counter()
    Sd = oracle.direct_solve(pr.N, Li, Lj, Lv, R)
    t_direct = time.perf_counter() - t0
    ex, ey, ep = oracle.stokes_unpack(pr, Sd)
    ev = np.concatenate([ex.ravel(), ey.ravel()])
    blk = pr.etan0 > 1e26
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out = {"system": "falling block 101x101 (1e21 / 1e27 Pa s, 3200 / 3300 kg/m3), N=30,000",
           "direct_solve_cpu": {"seconds": t_direct, "kind": "port",
                                "method": "scipy SuperLU (COLAMD) + 3 FP64 refinement steps, 1 core"}}
    sol = StokesSolverGPU(ctx)
    for name, prec in (("mixed_fp32_mg", 1), ("plain_fp64_mg", 0)):
        for rtol in (1e-8, 1e-10):
            sol.set_system(pr)
            sol.solve_system(precision=prec, rtol=rtol, max_iterations=2)   # build + warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sol.set_system(pr)
            S, st = sol.solve_system(precision=prec, rtol=rtol, max_iterations=3000)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            vx, vy, p = oracle.stokes_unpack(pr, to_host(S))
            v = np.concatenate([vx.ravel(), vy.ravel()])
            out[f"{name}_rtol_{rtol:g}"] = {
                "seconds_with_setup": el, "iterations": st["iterations"], "relres": st["relres"],
                "converged": st["converged"], "forward_error_velocity": rel(v, ev),
                "forward_error_pressure_outside_block": rel(p[~blk], ep[~blk]),
                "forward_error_pressure_inside_block": rel(p[blk], ep[blk])}
    return out


def bench_config5(ctx, steps):
    """The rifting model on 2001 x 801 nodes with 5 x 5 markers per cell (40M
    markers; four creeping layers with a geotherm, strain-weakening
    Drucker-Prager, +-0.5 cm/yr): `steps` time steps of the device driver
    (rheology, scatter, Picard loop over nodes with up to 10 solves to 1e-8,
    grid -> marker updates, RK4 with stress rotation, strain) for each solver
    setting, from the same initial state.  Solve times include the set-up of
    each new system."""
    import numpy as np
    import torch
    from earthbox_gpu import learned, timeloop, workloads
    tmpl, m, mats, par = workloads.rift2d_markers()
    out = {"system": f"rift2d_markers 2001x801 nodes (500 x 160 km), {m['x'].size} markers, "
                     f"{steps} time steps, <= {par['nglobal']} Picard iterations, solves to 1e-8",
           "steps": steps}
    net = learned.LearnedCorrection(ctx)
    if learned.WEIGHTS_FILE.exists():
        net.load()
    legs = (("mixed_fp32_mg", dict(precision=1)), ("plain_fp64_mg", dict(precision=0)),
            ("learned_mixed", dict(precision=1, learned=1)))
    for name, kw in legs:
        model = timeloop.upload(ctx, tmpl, m, mats, par)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        logs = []
        for k in range(steps):
            logs.append(timeloop.time_step(ctx, model, solve_opts=dict(rtol=1e-8, **kw)))
            progress(f"config5 {name} step {k + 1}/{steps}: {logs[-1]['seconds']:.2f} s, "
                     f"{logs[-1]['picard_iterations']} Picard, Krylov {logs[-1]['krylov_iterations']}")
        el = time.perf_counter() - t0
        solves = [t for lg in logs for t in lg["solve_seconds_with_setup"]]
        its = [k for lg in logs for k in lg["krylov_iterations"]]
        out[name] = {"seconds": el, "seconds_per_step": el / steps, "solves": len(solves),
                     "solve_seconds_with_setup_total": float(np.sum(solves)),
                     "solve_seconds_with_setup_mean": float(np.mean(solves)),
                     "krylov_iterations_mean": float(np.mean(its)), "krylov_iterations_max": int(np.max(its)),
                     "max_relres": max(lg["max_relres"] for lg in logs),
                     "picard_iterations": [lg["picard_iterations"] for lg in logs],
                     "learned_accepted": sum(lg["learned_accepted"] for lg in logs),
                     "simulated_years": model.timesum / 3.15576e7}
        del model
        torch.cuda.empty_cache()
    p, mx, lm = out["plain_fp64_mg"], out["mixed_fp32_mg"], out["learned_mixed"]
    out["speedup_mixed_vs_plain"] = p["solve_seconds_with_setup_mean"] / mx["solve_seconds_with_setup_mean"]
    out["speedup_learned_mixed_vs_plain"] = p["solve_seconds_with_setup_mean"] / lm["solve_seconds_with_setup_mean"]
    return out


def bench_heat(ctx, reps=5):
    """The rifting model's heat step on 2001 x 801 nodes with 40M markers:
    the heat-field scatter with the thermal boundary conditions, one heat
    solve (multigrid-preconditioned FGMRES to 1e-12) and the whole heat step
    (right side with shear heating off, adaptive heat loop with a 20 K change
    limit, marker temperature update) at the model time step; then the same
    solve on the 501 x 201 version of the grid beside the oracle's sparse
    direct solve of that system (SuperLU, one host core), the reference's
    direct-solve seam."""
    import numpy as np
    import torch
    import oracle
    from earthbox_gpu import heat, workloads
    from earthbox_gpu.device import to_device, to_host
    out = {}
    for nx, ny, nm, with_markers in ((2001, 801, 5, True), (501, 201, 5, False)):
        tmpl, m, mats, par = workloads.rift2d_markers(nx, ny, nm)
        therm, tmats, bc, hpar = workloads.rift2d_thermal(tmpl, m)
        g = tmpl.grid
        ctx.set_grid(g)
        dm = {k: to_device(v) for k, v in ({"x": m["x"], "y": m["y"], "TK": m["TK"], "matid": m["matid"]} | therm).items()}
        dm["inside_flags"] = torch.empty(m["x"].size, dtype=torch.int8, device=dm["x"].device)
        ctx.get_marker_inside_flags(dm["x"], dm["y"], out=dm["inside_flags"])
        dm["marker_xn"] = torch.zeros(m["x"].size, dtype=torch.int32, device=dm["x"].device)
        dm["marker_yn"] = torch.zeros_like(dm["marker_xn"])
        del m
        z = np.zeros((ny, nx), order="F")
        fields = {k: to_device(z) for k in heat.HEAT_FIELDS}
        fields.update(marker_xn=dm["marker_xn"], marker_yn=dm["marker_yn"])
        dbc = {k: to_device(v) for k, v in bc.items()}
        dmats = {"mat_hr": to_device(tmats["mat_hr"])}
        key = f"{nx}x{ny}"
        r = {"markers": int(dm["x"].numel())}
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        heat.interpolate_marker_heat_parameters_to_grid(ctx, dm, dmats, fields, dbc)
        if with_markers:
            torch.cuda.synchronize()
            ev[0].record()
            for _ in range(reps):
                heat.interpolate_marker_heat_parameters_to_grid(ctx, dm, dmats, fields, dbc)
            ev[1].record()
            torch.cuda.synchronize()
            r["scatter_ms"] = ev[0].elapsed_time(ev[1]) / reps
        RT1 = heat.calculate_heat_rhs_terms_grid(ctx, fields)
        dt = par["timestep"]
        tk2, st = heat.discretize_and_solve_conductive_heat_equation(ctx, RT1, fields["tk1"], fields["rhocp1"],
                                                                     fields["kt1"], dbc, dt)
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            tk2, st = heat.discretize_and_solve_conductive_heat_equation(ctx, RT1, fields["tk1"], fields["rhocp1"],
                                                                         fields["kt1"], dbc, dt, tk2)
        ev[1].record()
        torch.cuda.synchronize()
        r.update(solve_ms=ev[0].elapsed_time(ev[1]) / reps, solve_iterations=st["iterations"],
                 solve_relres=st["relres"])
        if with_markers:
            t0 = time.perf_counter()
            for _ in range(reps):
                TK = dm["TK"].clone()
                hr = heat.update_marker_temp_for_grid_and_subgrid_changes(
                    ctx, {**dm, "TK": TK}, fields, dbc, dt, hpar["max_temp_change"], 0.0, 1)
            torch.cuda.synchronize()
            r["heat_step_ms"] = (time.perf_counter() - t0) / reps * 1e3
            r["heat_loop"] = hr["stats"]
        else:
            Rt = {k: to_host(fields[k]) for k in ("tk1", "kt1", "rhocp1")}
            N, Li, Lj, Lv, R = oracle.heat_build(g, bc, to_host(RT1), Rt["tk1"], Rt["rhocp1"], Rt["kt1"], dt)
            t0 = time.perf_counter()
            oracle.direct_solve(N, Li, Lj, Lv, R, refine=0)
            r["oracle_direct_solve_ms"] = (time.perf_counter() - t0) * 1e3
            r["speedup_vs_oracle_direct_solve"] = r["oracle_direct_solve_ms"] / r["solve_ms"]
        progress(f"heat {key}: {r}")
        out[key] = r
        del dm, fields
        torch.cuda.empty_cache()
    return out


def bench_rk4_config5(ctx, steps=10):
    """RK4 gather-advect of 40M markers (5 x 5 per cell) on the rifting
    model's 2001 x 801-node grid, both arithmetics; kernel time from HIP
    events on the library's stream."""
    import numpy as np
    import torch
    from earthbox_gpu import workloads
    from earthbox_gpu.device import empty, to_device
    from earthbox_gpu.grid import uniform_grid
    g = uniform_grid(2001, 801, 5.0e5, 1.6e5)
    x, y = workloads.marker_positions(g, 5, 5, 5)
    vx1, vy1 = workloads.random_velocity_fields(g, 13, vmax=1e-9)
    ctx.set_grid(g)
    dvx1, dvy1 = to_device(vx1), to_device(vy1)
    desp = empty((g.ynum, g.xnum), np.float64)
    ctx.calculate_spin(dvx1, dvy1, desp)
    dx, dy = to_device(x), to_device(y)
    dt = 0.5 * g.xstpavg / 1e-9
    n = x.size
    out = {"markers": n, "grid": "2001x801 (500 x 160 km)"}
    stream = torch.cuda.current_stream()
    for name, mode in (("tolerance", ctx.RK4_TOLERANCE), ("exact", ctx.RK4_EXACT)):
        ctx.set_rk4_arithmetic(mode)
        ctx.advect_markers_and_rotate_stress_tensor(4, dt, None, dx, dy, None, None, dvx1, dvy1, desp)
        torch.cuda.synchronize()
        s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s0.record(stream)
        for _ in range(steps):
            ctx.advect_markers_and_rotate_stress_tensor(4, dt, None, dx, dy, None, None, dvx1, dvy1, desp)
        s1.record(stream)
        torch.cuda.synchronize()
        ms = s0.elapsed_time(s1) / steps
        out[name] = {"kernel_ms": ms, "Mmarkers_per_s": n / (ms * 1e-3) / 1e6,
                     "frac_of_8TBs": RK4_BYTES_PER_MARKER * n / (ms * 1e-3) / 8e12}
    ctx.set_rk4_arithmetic(ctx.RK4_TOLERANCE)
    del dx, dy
    torch.cuda.empty_cache()
    return out


def bench_scatter(ctx, g, x, y, reps=3):
    """One marker->grid Stokes scatter of all markers (material-table shear
    modulus, per-marker density and viscosity)."""
    import numpy as np
    import torch
    from earthbox_gpu.device import to_device
    from earthbox_gpu.scatter import interpolate_marker_stokes_parameters_to_grid, new_grid_fields
    n = x.size
    ctx.set_grid(g)
    rng = np.random.default_rng(5)
    m = {"x": to_device(x), "y": to_device(y),
         "rho": to_device(3000 + 300 * rng.random(n)),
         "eta": to_device(10 ** rng.uniform(19, 23, n)),
         "matid": to_device(rng.integers(1, 4, n).astype(np.int16)),
         "mat_mu": to_device(np.array([3e10, 6e10, 1e10]))}
    f = new_grid_fields(g.xnum, g.ynum, 1.0)
    # algorithmic bytes: x, y, rho, eta (8 B each) and matid (2 B) per marker,
    # and the grid outputs written once (rho1, etas1, mus1, wtnodes, wtetas;
    # etan1, mun1, wtetan; rho1_vy, wtnodes_vy)
    nb, npr, nvy = g.xnum * g.ynum, (g.xnum - 1) * (g.ynum - 1), g.ynum * (g.xnum + 1)
    alg = 34 * n + 8 * (5 * nb + 3 * npr + 2 * nvy)
    out = {"markers": n, "algorithmic_bytes": alg}
    for name, order in (("serial_order", ctx.SCATTER_SERIAL_ORDER), ("cell_order", ctx.SCATTER_CELL_ORDER)):
        ctx.set_scatter_order(order)
        interpolate_marker_stokes_parameters_to_grid(ctx, m, f)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            interpolate_marker_stokes_parameters_to_grid(ctx, m, f)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / reps
        out[name] = {"ms": el * 1e3, "Mmarkers_per_s": n / el / 1e6,
                     "frac_of_8TBs": alg / el / 1e9 / HBM_PEAK_GBS}
    ctx.set_scatter_order(ctx.SCATTER_SERIAL_ORDER)
    out["ms"] = out["serial_order"]["ms"]
    out["speedup_cell_vs_serial"] = out["serial_order"]["ms"] / out["cell_order"]["ms"]
    return out


def bench_grid_kernels(ctx, reps=20):
    """Per-call time of the grid kernels around the Stokes solve on the
    2001 x 801 rift state, with HIP events on the library's stream, and the
    achieved bandwidth on each call's algorithmic bytes (every input array
    read once, every output written once; FP64)."""
    import numpy as np
    import torch
    from earthbox_gpu import workloads
    from earthbox_gpu.device import to_device
    from earthbox_gpu.plastic import (interpolate_basic_viscoplastic_viscosity_to_pressure_grid,
                                      plastic_strain_rate_at_pressure_nodes, update_nodes_for_plastic_yielding)
    from earthbox_gpu.viscoelastic import calc_viscoelastic_terms_and_rhs, update_grid_strain_rate_and_stress
    tmpl, state, par = workloads.rift2d(2001, 801)
    g = tmpl.grid
    ctx.set_grid(g)
    nb, npr = g.xnum * g.ynum, (g.xnum - 1) * (g.ynum - 1)
    nvx, nvy = (g.ynum + 1) * g.xnum, g.ynum * (g.xnum + 1)
    ds = {k: to_device(v) for k, v in state.items()}
    rng = np.random.default_rng(1)
    vx1 = to_device(rng.uniform(-1e-9, 1e-9, (g.ynum + 1, g.xnum)))
    vy1 = to_device(rng.uniform(-1e-9, 1e-9, (g.ynum, g.xnum + 1)))
    ve = calc_viscoelastic_terms_and_rhs(ctx, ds, par["timestep"], 0.0, tmpl.gravity_y, 0)
    ss = update_grid_strain_rate_and_stress(ctx, {"vx1": vx1, "vy1": vy1, **ds}, par["timestep"])
    pr1 = to_device(rng.uniform(0, 5e8, (g.ynum - 1, g.xnum - 1)))
    pf = {k: ds[k] for k in ("plastic_yield", "etas1", "eta_flow", "mus1", "cohesion_grid", "fric_degrees_grid")}
    pf.update(sxy2=ss["sxy2"], sxx2=ss["sxx2"], pr1=pr1)
    pout = update_nodes_for_plastic_yielding(ctx, pf, ds["yield_error"], par)[0]
    calls = {
        # inputs: 4 basic + 3 pressure + vy density + 2 pressure; outputs: 2 basic + 2 pressure + RX1, RY1, RC1
        "viscoelastic_rhs": (lambda: calc_viscoelastic_terms_and_rhs(ctx, ds, par["timestep"], 0.0, tmpl.gravity_y,
                                                                     0, out=ve),
                             8 * (4 * nb + 5 * npr + nvy + 2 * nb + 3 * npr + nvx + nvy)),
        # inputs: vx1, vy1, 3 basic, 3 pressure; outputs: exy, sxy2, dsxy, exx, eii, sxx2, dsxx
        "strain_rate_and_stress": (lambda: update_grid_strain_rate_and_stress(ctx, {"vx1": vx1, "vy1": vy1, **ds},
                                                                              par["timestep"], out=ss),
                                   8 * (nvx + nvy + 3 * nb + 3 * npr + 3 * nb + 4 * npr)),
        # inputs: 7 basic + 2 pressure (+ yield_error read and written); outputs: 3 basic
        "plastic_yielding": (lambda: update_nodes_for_plastic_yielding(ctx, pf, ds["yield_error"], par, out=pout),
                             8 * (7 * nb + 2 * npr + 2 * nb + 3 * nb)),
        "harmonic_normal_viscosity": (lambda: interpolate_basic_viscoplastic_viscosity_to_pressure_grid(
            ctx, ds["etas1"], ds["plastics"], ds["plasticn"], ds["plastic_yield"], ds["etan1"]),
            8 * (3 * nb + 2 * npr + npr)),
        "plastic_rate_at_pressure_nodes": (lambda: plastic_strain_rate_at_pressure_nodes(
            ctx, pout["plastic_strain_rate"], ds["eii_plastic_pressure"]), 8 * (nb + npr)),
    }
    stream = torch.cuda.current_stream()
    out = {"system": "rift2d 2001x801 nodes", "nodes": nb}
    for name, (fn, nbytes) in calls.items():
        fn()
        torch.cuda.synchronize()
        s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s0.record(stream)
        for _ in range(reps):
            fn()
        s1.record(stream)
        torch.cuda.synchronize()
        ms = s0.elapsed_time(s1) / reps
        out[name] = {"ms": ms, "algorithmic_bytes": nbytes, "GB_per_s": nbytes / (ms * 1e-3) / 1e9,
                     "frac_of_8TBs": nbytes / (ms * 1e-3) / 8e12}
    return out


def bench_cnn(ctx, reps=10):
    """The learned correction's network on the MFMA units: a forward pass on a
    512 x 1024 cell image (the size of the config-2 solve's inputs) and a
    training step (forward, backward, weight and bias gradients) on a batch of
    32 crops of 64 x 128 cells, timed with HIP events; achieved TFLOP/s on the
    convolutions' multiply-adds (padded channels counted, as the MFMA units
    compute them) against the MI355X's dense BF16 peak (2.5 PFLOP/s).  The
    backward pass is counted as twice the forward (input and weight
    gradients)."""
    import numpy as np
    import torch
    from earthbox_gpu import learned
    net = learned.LearnedCorrection(ctx, seed=5)
    flop_px = sum(2 * 9 * cin * cout / 4 ** lv for cin, cout, _, _, lv, _ in learned.LAYERS)
    rng = np.random.default_rng(3)
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s0.record(stream)
        for _ in range(reps):
            fn()
        s1.record(stream)
        torch.cuda.synchronize()
        return s0.elapsed_time(s1) / reps

    x = torch.from_numpy(rng.standard_normal((1, 512, 1024, learned.N_IN)).astype(np.float32)).cuda()
    ms_f = timed(lambda: net.forward(x))
    flops_f = flop_px * 512 * 1024
    xb = torch.from_numpy(rng.standard_normal((32, 64, 128, learned.N_IN)).astype(np.float32)).cuda()
    tb = torch.from_numpy(rng.standard_normal((32, 64, 128, learned.N_OUT)).astype(np.float32)).cuda()
    ms_t = timed(lambda: net.loss_and_grad(xb, tb))
    flops_t = 3 * flop_px * 32 * 64 * 128
    peak = 2500.0
    return {"network": f"4-level U-Net, {len(learned.LAYERS)} 3x3 convolutions, 32 channels, BF16 MFMA "
                       f"(v_mfma_f32_16x16x32_bf16 where 32 input channels, 16x16x16 otherwise)",
            "forward_512x1024": {"ms": ms_f, "gflop": flops_f / 1e9, "tflops": flops_f / (ms_f * 1e-3) / 1e12,
                                 "frac_of_bf16_peak": flops_f / (ms_f * 1e-3) / 1e12 / peak},
            "train_step_32x64x128": {"ms": ms_t, "gflop": flops_t / 1e9,
                                     "tflops": flops_t / (ms_t * 1e-3) / 1e12,
                                     "frac_of_bf16_peak": flops_t / (ms_t * 1e-3) / 1e12 / peak}}


def bench_picard(ctx, nglobal=10, tol=1e-4):
    """The Picard loop over nodes on the synthetic 2001 x 801 visco-elasto-plastic
    rift state (four layers, cohesion and friction, +-0.5 cm/yr extension, one
    100-kyr step): every Picard iteration forms the viscoelastic right-hand
    side, solves to 1e-8 in FP64 (set-up of the preconditioner for the new
    viscosity included), processes the solution on the grid and updates the
    nodes for plastic yielding, until LoopTermination stops the loop."""
    import torch
    from earthbox_gpu import learned, workloads
    from earthbox_gpu.device import to_device
    from earthbox_gpu.picard import picard_loop_nodes
    tmpl, state, par = workloads.rift2d(2001, 801)
    ctx.set_grid(tmpl.grid)
    out = {"system": "rift2d 2001x801 nodes (500 x 160 km), 4 layers, plastic, +-0.5 cm/yr",
           "picard_tolerance": tol, "nglobal": nglobal, "solve_rtol": 1e-8}
    net = learned.LearnedCorrection(ctx)
    if learned.WEIGHTS_FILE.exists():
        net.load()
    legs = (("mixed_fp32_mg", dict(precision=1)), ("plain_fp64_mg", dict(precision=0)),
            ("learned_mixed", dict(precision=1, learned=1)))
    for name, kw in legs:
        opts = dict(rtol=1e-8, **kw)
        picard_loop_nodes(ctx, tmpl, {k: to_device(v) for k, v in state.items()}, par, 1, tol, solve_opts=opts)
        ds = {k: to_device(v) for k, v in state.items()}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, log = picard_loop_nodes(ctx, tmpl, ds, par, nglobal, tol, solve_opts=opts, warm_start=True)
        el = time.perf_counter() - t0
        out[name] = {"seconds": el, "picard_iterations": len(log),
                     "krylov_iterations": [r["iterations"] for r in log],
                     "max_relres": max(r["relres"] for r in log),
                     "nnode_yield": [r["nnode_yield"] for r in log],
                     "final_dvxy_rel_L2": log[-1]["dvxy_rel_L2"],
                     "final_global_yield_error": log[-1]["global_yield_error"]}
    p, m, lm = out["plain_fp64_mg"], out["mixed_fp32_mg"], out["learned_mixed"]
    out["speedup_mixed_vs_plain"] = p["seconds"] / m["seconds"]
    out["speedup_learned_mixed_vs_plain"] = p["seconds"] / lm["seconds"]
    return out


def _host_cpu():
    """CPU model, logical CPUs and the OpenMP thread setting of this host."""
    model = platform.processor() or "unknown"
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "sched_affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(g, x, y, vx1, vy1, esp, dt, sample_markers: int):
    """The oracle's RK4 restatement (threaded over markers as the reference's
    Threads.@threads loop, VelocityAndSpin.jl:77) timed on this host: on all
    the cores this process may use, and on one core.  Test-infrastructure
    checker used only as the reported CPU baseline."""
    import oracle
    host = _host_cpu()
    all_cores = int(os.environ.get("OMP_NUM_THREADS") or host["sched_affinity"] or os.cpu_count() or 1)
    xs, ys = x[:sample_markers].copy(), y[:sample_markers].copy()
    flags = oracle.inside_flags(xs, ys, g.limits)
    xn, yn = oracle.cell_indices_basic(xs, ys, flags, g)

    def rate(threads, budget):
        oracle.set_threads(threads)
        oracle.advect_markers_and_rotate_stress_tensor(4, dt, g, xs, ys, xn, yn, flags, vx1, vy1, esp)
        times = []
        t_end = time.perf_counter() + budget
        while time.perf_counter() < t_end or len(times) < 3:
            t0 = time.perf_counter()
            oracle.advect_markers_and_rotate_stress_tensor(4, dt, g, xs, ys, xn, yn, flags,
                                                           vx1, vy1, esp)
            times.append(time.perf_counter() - t0)
        return sample_markers / float(np.median(times)) / 1e6, len(times)

    import numpy as np
    r_all, n_all = rate(all_cores, 8.0)
    r_one, n_one = rate(1, 8.0)
    oracle.set_threads(all_cores)
    return {"value": r_all, "unit": "Mmarkers/s", "cores": all_cores, "kind": "port",
            "single_core_value": r_one, "host": host,
            "sample": f"RK4 advect of the first {sample_markers} markers of the same workload "
                      f"(oracle C restatement, OpenMP over markers): median of {n_all} runs on "
                      f"{all_cores} threads, {n_one} runs on 1 thread"}


def bench_marker_gathers(ctx, g, dx, dy, reps=10):
    """Per-call time of the fused grid -> marker gathers after the solve
    (marker_exx, marker_exy, marker_sr_ratio, marker_pr) on the RK4 workload's
    markers, with HIP events on the library's stream.  Algorithmic bytes: per
    marker x, y, xn, yn, eta, matid read and four FP64 outputs written (66 B),
    plus each of the seven grid arrays read once."""
    import numpy as np
    import torch
    from earthbox_gpu.device import to_device
    from earthbox_gpu.marker_gather import update_markers_after_stokes
    ctx.set_grid(g)
    n = int(dx.numel())
    xn = torch.zeros(n, dtype=torch.int32, device=dx.device)
    yn = torch.zeros(n, dtype=torch.int32, device=dx.device)
    ctx.calc_upper_left_basic_grid_cell_indices_for_markers(dx, dy, None, xn, yn)
    rng = np.random.default_rng(3)
    nb, npr = g.xnum * g.ynum, (g.xnum - 1) * (g.ynum - 1)
    pshape, bshape = (g.ynum - 1, g.xnum - 1), (g.ynum, g.xnum)
    f = {k: to_device(rng.uniform(-1e7, 1e7, pshape)) for k in ("exx", "pr1", "sxx2", "dsxx")}
    f.update({k: to_device(rng.uniform(-1e7, 1e7, bshape)) for k in ("exy", "sxy2", "dsxy")})
    m = {"x": dx, "y": dy, "marker_xn": xn, "marker_yn": yn,
         "eta": torch.full((n,), 1e21, dtype=torch.float64, device=dx.device),
         "matid": torch.ones(n, dtype=torch.int16, device=dx.device),
         "mat_mu": to_device(np.array([3e10]))}
    out = update_markers_after_stokes(ctx, m, f, 3.15576e12)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s0.record(stream)
    for _ in range(reps):
        update_markers_after_stokes(ctx, m, f, 3.15576e12, out)
    s1.record(stream)
    torch.cuda.synchronize()
    ms = s0.elapsed_time(s1) / reps
    nbytes = 66 * n + 8 * (4 * npr + 3 * nb)
    return {"kernel": "marker_gather_kernel", "markers": n, "ms": ms, "Mmarkers_per_s": n / (ms * 1e-3) / 1e6,
            "algorithmic_bytes": nbytes, "GB_per_s": nbytes / (ms * 1e-3) / 1e9,
            "frac_of_8TBs": nbytes / (ms * 1e-3) / 8e12}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1 "
                          "(multi-GPU sharding is design-only, see DESIGN.md)"}))
        sys.exit(2)
    import numpy as np
    import torch
    if args.calibrate_copy:
        a = torch.ones(64 * 2 ** 20, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": a.numel() * 8, "bytes_written": b.numel() * 8,
                          "dispatches": 3}))
        return
    import earthbox_gpu
    from earthbox_gpu import workloads
    from earthbox_gpu.device import empty, to_device
    from earthbox_gpu.grid import uniform_grid

    nx, ny, nm = args.xnum, args.ynum, args.markers_per_cell
    ctx = earthbox_gpu.GPUContext(0)
    progress("rk4 velocity field: config-2 Stokes solve ...")
    g, dvx1, dvy1, dt, field = config2_velocity(ctx, nx, ny)
    ctx.set_grid(g)
    x, y = workloads.marker_positions(g, nm, nm, 2)
    desp = empty((ny, nx), np.float64)
    ctx.calculate_spin(dvx1, dvy1, desp)
    dx, dy = to_device(x), to_device(y)
    n = x.size

    def timed(mode, vx1=dvx1, vy1=dvy1, sp=desp, step_dt=dt, fresh=True):
        """W untimed steps, then K steps each bracketed by HIP events on the
        library's stream (torch's current stream); wall time over the K.
        fresh: step k advects its own copy of the initial positions (copies
        made before the timed region); otherwise every step advances the same
        arrays."""
        ctx.set_rk4_arithmetic(mode)
        nsteps = args.warmup + args.steps
        if fresh:
            pos = [(dx.clone(), dy.clone()) for _ in range(nsteps)]
        else:
            pos = [(dx.clone(), dy.clone())] * nsteps
        for k in range(args.warmup):
            ctx.advect_markers_and_rotate_stress_tensor(4, step_dt, None, pos[k][0], pos[k][1], None, None,
                                                        vx1, vy1, sp)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k, (s, e) in enumerate(evs):
            px, py = pos[args.warmup + k]
            s.record(stream)
            ctx.advect_markers_and_rotate_stress_tensor(4, step_dt, None, px, py, None, None, vx1, vy1, sp)
            e.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        del pos
        return wall, float(np.mean([s.elapsed_time(e) for s, e in evs]))

    # the exact (bit-identical) kernel, reported beside the headline
    wall_x, kern_x = timed(ctx.RK4_EXACT)
    ex_ms = wall_x * 1e3 / args.steps
    rk4_exact = {"kernel": "rk4_advect_kernel", "arithmetic": "exact (bit-identical to the oracle)",
                 "ms_per_step": ex_ms, "Mmarkers_per_s": n / (ex_ms * 1e-3) / 1e6, "kernel_ms": kern_x,
                 "traffic": RK4_MEASURED_TRAFFIC.get(("rk4_advect_kernel", nx, ny, nm)),
                 "frac_of_8TBs": RK4_BYTES_PER_MARKER * n / (kern_x * 1e-3) / 1e9 / HBM_PEAK_GBS}
    mode = ctx.RK4_TOLERANCE if args.rk4_arithmetic == "tolerance" else ctx.RK4_EXACT
    fast = mode == ctx.RK4_TOLERANCE and ctx.rk4_tolerance_kernel_available()
    wall, kern_ms = timed(mode)
    ms_per_step = wall * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6
    achieved = RK4_BYTES_PER_MARKER * n / (kern_ms * 1e-3) / 1e9
    kname = "rk4_tol_kernel" if fast else "rk4_advect_kernel"
    traffic = RK4_MEASURED_TRAFFIC.get((kname, nx, ny, nm))
    out = {
        "metric": "Stokes time-to-1e-8 rel. residual (learned+mixed vs FP64 MG); Mmarkers/s RK4",
        "value": value,
        "unit": "Mmarkers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"rk4_advect {nx}x{ny} nodes, {nm}x{nm} markers/cell ({n} markers), "
                               f"config-2 Stokes velocity, dt = 0.5 cell / max|v|, each step from the "
                               f"initial positions",
                   "markers": n, "xnum": nx, "ynum": ny, "velocity_field": field,
                   "arithmetic": ("tolerance: FMA, nominal uniform coordinates; positions within 1e-10 "
                                  "of the domain size of the exact kernel" if fast else "exact")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "kernel_ms": kern_ms,
                     "bytes_per_unit": RK4_BYTES_PER_MARKER},
        "rk4_exact": rk4_exact,
    }
    progress(f"rk4 headline: {kern_ms:.3f} ms per launch ({kname})")
    # the same markers under the synthetic random field (smoothed white noise,
    # |v| <= 1e-9 m/s, dt = 0.5 cell / 1e-9), K steps on the same arrays, as a
    # harder case for the LDS tiles: markers of a wavefront drift apart
    rvx, rvy = workloads.random_velocity_fields(g, 12, vmax=1e-9)
    drvx, drvy = to_device(rvx), to_device(rvy)
    rsp = empty((ny, nx), np.float64)
    ctx.calculate_spin(drvx, drvy, rsp)
    rdt = 0.5 * g.xstpavg / 1e-9
    wall_r, kern_r = timed(mode, drvx, drvy, rsp, rdt, fresh=False)
    out["rk4_random_field"] = {"kernel": kname, "kernel_ms": kern_r, "ms_per_step": wall_r * 1e3 / args.steps,
                               "frac_of_8TBs": RK4_BYTES_PER_MARKER * n / (kern_r * 1e-3) / 1e9 / HBM_PEAK_GBS,
                               "field": "smoothed white noise, |v| <= 1e-9 m/s, dt = 0.5 cell / 1e-9 m/s, "
                                        "K steps advancing the same arrays"}
    ctx.set_grid(g)
    legs = [
        ("marker_gathers", not args.no_marker_gathers, lambda: bench_marker_gathers(ctx, g, dx, dy)),
        ("stokes", not args.no_stokes,
         lambda: bench_stokes(ctx, nx, ny, args.stokes_max_iter, args.stokes_amplitude)),
        ("config1", not args.no_config1, lambda: bench_config1(ctx)),
        ("rk4_config5", not args.no_config5, lambda: bench_rk4_config5(ctx)),
        ("config5", not args.no_config5, lambda: bench_config5(ctx, args.config5_steps)),
        ("heat", not args.no_heat, lambda: bench_heat(ctx)),
        ("scatter", not args.no_scatter, lambda: bench_scatter(ctx, g, x, y)),
        ("stokes3d", not args.no_stokes3d, lambda: bench_stokes3d(ctx, args.stokes3d_n, args.stokes3d_max_iter)),
        ("sinker", not args.no_sinker, lambda: bench_sinker(ctx)),
        ("picard", not args.no_picard, lambda: bench_picard(ctx)),
        ("grid_kernels", not args.no_grid_kernels, lambda: bench_grid_kernels(ctx)),
        ("cnn", not args.no_cnn, lambda: bench_cnn(ctx)),
    ]
    for name, on, fn in legs:
        if not on:
            continue
        t0 = time.perf_counter()
        progress(f"{name} ...")
        out[name] = fn()
        progress(f"{name} done in {time.perf_counter() - t0:.1f} s")
    if not args.no_cpu_baseline:
        progress("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(g, x, y, to_host_np(dvx1), to_host_np(dvy1), to_host_np(desp), dt,
                                           2_000_000)
        if not args.no_sinker:
            out["cpu_baseline"]["mg3d_97_reference_driver"] = cpu_sinker_driver()
    print(json.dumps(out), flush=True)
    ctx.close()


def to_host_np(t):
    from earthbox_gpu.device import to_host
    return to_host(t)


if __name__ == "__main__":
    main()
