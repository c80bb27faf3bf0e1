"""The specified configurations at their full size, through size-independent
properties (the oracle at these sizes only where it finishes in seconds):

- 1001 x 501 layered random system (1.0-decade perturbation): FP32 and FP64
  V-cycles reach ||R - L S|| / ||R|| <= 1e-8, recomputed on the host with the
  oracle's COO build and residual in the reference's triplet order; the two
  solutions agree to 1e-6;
- 18M markers (6 x 6 per cell) scattered to the 1001 x 501 grid: bit-exact
  against the oracle's serial loops over all markers;
- 257^3 unit cube with 8 spheres of contrast 1e2-1e6: GMRES(200) with the
  FP32 V-cycle reaches 1e-8, recomputed from the operator apply;
- the learned correction's network on a 513 x 257 patch: forward and
  gradient against the float64 oracle (BF16 emulated);
- the rifting model on 2001 x 801 nodes with 40M markers: one Picard
  iteration after the marker rheology and scatter; the solve's residual
  recomputed by the COO-order residual kernel is <= 1e-8, and the nodal
  yield count equals the oracle's on the same fields;
- the rifting model's heat step on 2001 x 801 nodes: the heat-field scatter
  of its 40M markers bit-exact against the oracle's serial loops, the heat
  solve's ||R - L T|| / ||R|| <= 1e-12 recomputed on the host from the
  oracle's COO rows, and the adaptive heat loop's sub-steps summing to the
  time step.
"""
import numpy as np
import pytest
import torch

import oracle
from earthbox_gpu import heat, learned, rheology, timeloop, workloads
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.grid import uniform_grid
from earthbox_gpu.picard import picard_loop_nodes
from earthbox_gpu.plastic import interpolate_basic_viscoplastic_viscosity_to_pressure_grid
from earthbox_gpu.scatter import interpolate_marker_stokes_parameters_to_grid
from earthbox_gpu.stokes import StokesSolverGPU
from earthbox_gpu.stokes3d import Stokes3DSolverGPU

pytestmark = pytest.mark.gpu


def test_config2_solve_residual_recomputed(gpu_ctx):
    print("1001 x 501 layered random system, FP32 and FP64 V-cycles", flush=True)
    pr = workloads.layered_random(1001, 501, amplitude=1.0)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    sol = StokesSolverGPU(gpu_ctx)
    sol.set_system(pr)
    xs = {}
    for prec in (1, 0):
        S, st = sol.solve_system(precision=prec, rtol=1e-8, max_iterations=2000)
        print(prec, st, flush=True)
        assert st["converged"]
        Sg = to_host(S)
        res = oracle.coo_residual(pr.N, Li, Lj, Lv, Sg, R)
        rel = np.linalg.norm(res) / np.linalg.norm(R)
        assert rel <= 1.0001e-8, rel
        xs[prec] = Sg
    d = np.linalg.norm(xs[1] - xs[0]) / np.linalg.norm(xs[0])
    print(f"FP32 vs FP64 V-cycle solutions differ by {d:.2e}", flush=True)
    assert d < 1e-6


def test_config2_scatter_18M_bit_exact(gpu_ctx):
    print("18M-marker scatter on 1001 x 501 nodes", flush=True)
    from scatter_cases import previous_fields
    g = uniform_grid(1001, 501, 1.5e6, 6.0e5)
    x, y = workloads.marker_positions(g, 6, 6, 2)
    n = x.size
    rng = np.random.default_rng(5)
    m = {"x": x, "y": y, "rho": 3000 + 300 * rng.random(n), "eta": 10 ** rng.uniform(19, 23, n),
         "matid": rng.integers(1, 4, n).astype(np.int16), "mat_mu": np.array([3e10, 6e10, 1e10])}
    prev = previous_fields(g, 4)
    keys = ("rho1", "rho1_vy", "etas1", "etan1", "mus1", "mun1")
    exp = oracle.scatter_stokes(g, m, prev)
    gpu_ctx.set_grid(g)
    dm = {k: to_device(v) for k, v in m.items()}
    df = {k: to_device(prev[k]) for k in keys}
    for k in ("wtnodes", "wtetas", "wtetan", "wtnodes_vy"):
        df[k] = to_device(np.zeros_like(exp[k], order="F"))
    interpolate_marker_stokes_parameters_to_grid(gpu_ctx, dm, df)
    gpu_ctx.synchronize()
    for k in keys + ("wtnodes", "wtetas", "wtetan", "wtnodes_vy"):
        assert np.array_equal(to_host(df[k]), exp[k]), k


def test_config3_spheres_257(gpu_ctx):
    print("257^3 spheres, contrast 1e2-1e6, GMRES(200) + FP32 V-cycle", flush=True)
    w = workloads.spheres3d(n=257, seed=3, log_contrast=(2.0, 6.0), device=gpu_ctx.device)
    sol = Stokes3DSolverGPU(gpu_ctx)
    sol.set_system(*w["dims"], w["etan"], w["etaxy"], w["etaxz"], w["etayz"])
    rhs = sol.pack({"vx": w["RX"], "vy": w["RY"], "vz": w["RZ"], "pr": w["RC"]})
    del w
    x, st = sol.solve(rhs, precision=1, rtol=1e-8, max_iterations=800, restart=200, flexible=0)
    print(st, flush=True)
    assert st["converged"], st
    ax = sol.vector()
    sol.apply(x.clone(), ax)
    parts = sol.split(rhs)
    for k in ("vx", "vy", "vz"):
        p = parts[k]
        p[0].zero_(); p[-1].zero_(); p[:, 0].zero_(); p[:, -1].zero_(); p[:, :, 0].zero_(); p[:, :, -1].zero_()
    parts["pr"][0, 0, 0] = 0.0
    rel = (torch.linalg.vector_norm(rhs - ax) / torch.linalg.vector_norm(rhs)).item()
    print(f"recomputed relative residual {rel:.3e}", flush=True)
    assert rel <= 1.05e-8
    del sol, rhs, x, ax
    torch.cuda.empty_cache()


def test_config4_network_513x257(gpu_ctx):
    import oracle.cnn as ocnn
    from test_oracle_cnn import _params
    print("network forward and gradient on a 513 x 257 patch", flush=True)
    net = learned.LearnedCorrection(gpu_ctx, seed=3)
    p = _params(7).astype(np.float32)
    net.set_params(p)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((1, 513, 257, 7)).astype(np.float32)
    t = rng.standard_normal((1, 513, 257, 3)).astype(np.float32)
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    out = net.forward(xd).cpu().numpy()
    ref = ocnn.forward(p, x)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(out, ref) < 2e-2, rel(out, ref)
    loss, g = net.loss_and_grad(xd, td, host_grad=True)
    rloss, rg = ocnn.loss_and_grad(p, x, t)
    print(f"loss {loss} oracle {rloss} grad rel {rel(g, rg):.3e}", flush=True)
    assert abs(loss - rloss) <= 1e-2 * abs(rloss)
    assert rel(g, rg) < 5e-2


def test_config5_picard_iteration_2001x801(gpu_ctx):
    print("rift model 2001 x 801, 40M markers: rheology, scatter, one Picard iteration", flush=True)
    tmpl, m, mats, par = workloads.rift2d_markers()
    model = timeloop.upload(gpu_ctx, tmpl, m, mats, par)
    del m
    dm, gf = model.markers, model.grid
    g = tmpl.grid
    gpu_ctx.get_marker_inside_flags(dm["x"], dm["y"], out=dm["inside_flags"])
    rheology.update_marker_rheology(gpu_ctx, dm, model.mats, dict(par, ntimestep=1), dm)
    interpolate_marker_stokes_parameters_to_grid(gpu_ctx, {**dm, "dilatation": None, "meltfrac": None},
                                                 {**gf, "marker_xn": dm["marker_xn"], "marker_yn": dm["marker_yn"]})
    interpolate_basic_viscoplastic_viscosity_to_pressure_grid(gpu_ctx, gf["etas1"], gf["plastics"], gf["plasticn"],
                                                              gf["plastic_yield"], gf["etan1"])
    before = {k: to_host(gf[k]) for k in oracle.PLASTIC_FIELDS if k in gf}
    ye_before = to_host(gf["yield_error"])
    ppar = dict(timestep=par["timestep"], timesum=0.0, ntimestep=1, viscosity_min=par["viscosity_min"],
                viscosity_max=par["viscosity_max"], iuse_fluid_pressure_for_yield=0)
    solver = StokesSolverGPU(gpu_ctx)
    sol, log = picard_loop_nodes(gpu_ctx, tmpl, gf, ppar, 1, 1e-3, solver=solver, solve_opts=dict(rtol=1e-8))
    print(log, flush=True)
    res = solver.stokes_calc_nonlinear_system_residual(sol["S"])
    rel = (torch.linalg.vector_norm(res) / torch.linalg.vector_norm(solver.rhs())).item()
    print(f"recomputed relative residual {rel:.3e}", flush=True)
    assert rel <= 1.0001e-8
    f = dict(before, sxy2=to_host(sol["sxy2"]), sxx2=to_host(sol["sxx2"]), pr1=to_host(sol["pr1"]))
    *_, n, err = oracle.update_nodes_for_plastic_yielding(g, f, ppar, ye_before)
    print(f"yielding nodes: gpu {log[0]['nnode_yield']}, oracle {n}", flush=True)
    assert n > 0 and log[0]["nnode_yield"] == n
    assert log[0]["global_yield_error"] == pytest.approx(err, rel=1e-9)
    del model, solver
    torch.cuda.empty_cache()


def test_config5_heat_step_2001x801(gpu_ctx):
    print("rift model 2001 x 801, 40M markers: heat scatter, heat solve, heat loop", flush=True)
    tmpl, m, mats, par = workloads.rift2d_markers()
    therm, tmats, bc, hpar = workloads.rift2d_thermal(tmpl, m)
    g = tmpl.grid
    ny, nx = g.ynum, g.xnum
    flags = oracle.inside_flags(m["x"], m["y"], g.limits)
    hm = {"x": m["x"], "y": m["y"], "flags": flags, "TK": m["TK"], "matid": m["matid"], **therm}
    zero = np.zeros((ny, nx), order="F")
    exp = oracle.scatter_heat(g, hm, {"tk2": zero, "kt1": zero, "rhocp1": zero, "hr1": zero, "ha1": zero},
                              tmats["mat_hr"], None, bc)
    print("oracle scatter done", flush=True)
    gpu_ctx.set_grid(g)
    dm = {k: to_device(v) for k, v in hm.items() if k != "flags"}
    dm["inside_flags"] = to_device(flags)
    del m, hm
    fields = {k: to_device(zero) for k in heat.HEAT_FIELDS}
    dbc = {k: to_device(v) for k, v in bc.items()}
    heat.interpolate_marker_heat_parameters_to_grid(gpu_ctx, dm, {"mat_hr": to_device(tmats["mat_hr"])}, fields, dbc)
    gpu_ctx.synchronize()
    for k in heat.HEAT_FIELDS:
        assert np.array_equal(to_host(fields[k]), exp[k]), k
    RT1 = heat.calculate_heat_rhs_terms_grid(gpu_ctx, fields)   # radiogenic heat only
    dt = par["timestep"]
    torch.cuda.synchronize()
    tk2, st = heat.discretize_and_solve_conductive_heat_equation(gpu_ctx, RT1, fields["tk1"], fields["rhocp1"],
                                                                 fields["kt1"], dbc, dt)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    tk2, st = heat.discretize_and_solve_conductive_heat_equation(gpu_ctx, RT1, fields["tk1"], fields["rhocp1"],
                                                                 fields["kt1"], dbc, dt, tk2)
    t1.record()
    torch.cuda.synchronize()
    print(f"heat solve: {st}, {t0.elapsed_time(t1):.2f} ms", flush=True)
    assert st["converged"] and st["iterations"] <= 40
    N, Li, Lj, Lv, R = oracle.heat_build(g, bc, to_host(RT1), exp["tk1"], exp["rhocp1"], exp["kt1"], dt)
    r = R - oracle.coo_spmv(N, Li, Lj, Lv, to_host(tk2).ravel(order="F"))
    rel = np.linalg.norm(r) / np.linalg.norm(R)
    print(f"host-recomputed relative residual {rel:.3e}", flush=True)
    assert rel <= 2e-12
    out, lst = heat.run_conductive_heat_solver_loop(gpu_ctx, RT1, fields["tk1"], fields["rhocp1"], fields["kt1"],
                                                    dbc, 10 * dt, hpar["max_temp_change"])
    torch.cuda.synchronize()
    print(f"heat loop over 10 time steps: {lst}", flush=True)
    assert lst["converged"] and lst["timestep_sum"] == pytest.approx(10 * dt, rel=1e-12)
    chg = np.abs(to_host(out["dtk1"])).max()
    assert 0 < chg and np.isfinite(chg)
    del dm, fields, out
    torch.cuda.empty_cache()
