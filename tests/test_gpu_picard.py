"""The GPU Picard loop over nodes against the same loop composed from the
oracle's restatements with a sparse direct solve, on a 41 x 41 falling block
with a cohesion that the stresses around the block exceed.  The loop, the
yield counts and the stopping step agree; viscosities and velocities agree to
the iterative solve's tolerance at every node whose yield flag matches.  The
change norm compares with the previous step's velocity (zeros on the first
step), and the displacement-limited time step is recomputed after every solve
when iupdate_timestep is 1."""
import dataclasses

import numpy as np
import pytest

import oracle
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.grid import uniform_grid
from earthbox_gpu.picard import STATE, check_stop_global_loop_nodes, picard_loop_nodes
from earthbox_gpu.stokes_problem import StokesProblem, free_slip_bcs

pytestmark = pytest.mark.gpu

PAR = dict(timestep=1e12, timesum=1e13, ntimestep=1, viscosity_min=1e18, viscosity_max=1e25)


def _setup(n=41):
    g = uniform_grid(n, n, 1e6, 1e6)
    def inside(X, Y):
        return (np.abs(X - 5e5) <= 1e5) & (np.abs(Y - 5e5) <= 1e5)
    Xb, Yb = np.meshgrid(g.gridx_b, g.gridy_b)
    Xp, Yp = np.meshgrid(g.gridx_pr, g.gridy_pr)
    Xv, Yv = np.meshgrid(g.gridx_vy, g.gridy_b)
    F = np.asfortranarray
    ny, nx = g.ynum, g.xnum
    s = {"etas1": F(np.where(inside(Xb, Yb), 1e24, 1e21)), "rho1": F(np.where(inside(Xb, Yb), 3300.0, 3200.0)),
         "rho1_vy": F(np.where(inside(Xv, Yv), 3300.0, 3200.0)), "etan1": F(np.where(inside(Xp, Yp), 1e24, 1e21)),
         "mus1": F(np.full((ny, nx), 1e11)), "mun1": F(np.full((ny - 1, nx - 1), 1e11)),
         "sxy1": F(np.zeros((ny, nx))), "sxx1": F(np.zeros((ny - 1, nx - 1))),
         "plastics": F(np.zeros((ny, nx))), "plasticn": F(np.zeros((ny - 1, nx - 1))),
         "plastic_yield": F(np.zeros((ny, nx))), "cohesion_grid": F(np.full((ny, nx), 4e7)),
         "fric_degrees_grid": F(np.zeros((ny, nx))), "dilatation_grid": F(np.zeros((ny - 1, nx - 1))),
         "eii_plastic_pressure": F(np.zeros((ny - 1, nx - 1))), "yield_error": F(np.zeros((ny, nx)))}
    s["eta_flow"] = s["etas1"].copy(order="F")
    z = lambda shp: np.zeros(shp, order="F")
    template = StokesProblem(grid=g, etas0=s["etas1"], etan0=s["etan1"], rho1_vy=s["rho1_vy"], RX1=z((ny + 1, nx)),
                             RY1=z((ny, nx + 1)), RC1=z((ny - 1, nx - 1)), bcs=free_slip_bcs(g), gravity_y=10.0)
    return g, s, template


def _cpu_loop(g, s, template, nglobal, tol, old=None, iupdate_timestep=0, displ_max=0.5):
    """The loop composed from the oracle; old: the previous step's velocity
    as one concatenated vector (None: zeros, the first time step)."""
    s = {k: v.copy(order="F") for k, v in s.items()}
    dt = PAR["timestep"]
    par = dict(PAR)
    log = []
    for it in range(1, nglobal + 1):
        if it > 1:
            s["etan1"] = oracle.harmonic_normal_viscosity(g, s["etas1"], s["plastics"], s["plasticn"],
                                                          s["plastic_yield"], s["etan1"])
        ve = oracle.stokes_viscoelastic_rhs(g, s, dt, 0.0, template.gravity_y, 0)
        prob = dataclasses.replace(template, etas0=ve["etas0"], etan0=ve["etan0"], RX1=ve["RX1"], RY1=ve["RY1"],
                                   RC1=ve["RC1"], rho1_vy=s["rho1_vy"], timestep=dt,
                                   pscale=2.0 * ve["etan0"][0, 0] / (g.xstpavg + g.ystpavg))
        Li, Lj, Lv, R = oracle.stokes_build(prob)
        S = oracle.direct_solve(prob.N, Li, Lj, Lv, R)
        vx1, vy1, pr1 = oracle.stokes_unpack(prob, S)
        if iupdate_timestep:
            dt, _ = oracle.calculate_new_timestep(dt, displ_max, g.xstpavg, g.ystpavg, vx1, vy1)
            par["timestep"] = dt
        ss = oracle.grid_strain_rate_and_stress(g, vx1, vy1, s, dt)
        v = np.concatenate([vx1.ravel(order="F"), vy1.ravel(order="F")])
        dv = oracle.get_L2_norms(v, old if old is not None else np.zeros_like(v))[1]
        old = v
        f = {k: s[k] for k in oracle.PLASTIC_FIELDS if k in s}
        f.update(sxy2=ss["sxy2"], sxx2=ss["sxx2"], pr1=pr1)
        ev, psr, pyn, ye, n, err = oracle.update_nodes_for_plastic_yielding(g, f, par, s["yield_error"])
        s["yield_error"] = ye
        if n > 0:
            s["etas1"], s["plastic_yield"] = ev, pyn
            s["eii_plastic_pressure"] = oracle.plastic_strain_rate_at_pressure_nodes(g, psr)
        log.append({"iglobal": it, "nnode_yield": n, "global_yield_error": err, "dvxy_rel_L2": dv,
                    "timestep": dt})
        if check_stop_global_loop_nodes(n, it, nglobal, err, dv, tol):
            break
    return s, {"vx1": vx1, "vy1": vy1, "pr1": pr1}, log


def test_picard_loop_nodes_matches_oracle_loop(gpu_ctx):
    print("picard loop over nodes, 41 x 41 falling block", flush=True)
    g, s, template = _setup()
    nglobal, tol = 6, 1e-2
    cs, csol, clog = _cpu_loop(g, s, template, nglobal, tol)
    gpu_ctx.set_grid(g)
    ds = {k: to_device(v) for k, v in s.items()}
    gsol, glog = picard_loop_nodes(gpu_ctx, template, ds, PAR, nglobal, tol, solve_opts=dict(rtol=1e-11))
    print("cpu", [(r["nnode_yield"], f"{r['global_yield_error']:.4g}") for r in clog], flush=True)
    print("gpu", [(r["nnode_yield"], f"{r['global_yield_error']:.4g}") for r in glog], flush=True)
    assert clog[0]["nnode_yield"] > 0, "the test problem must yield"
    assert len(glog) == len(clog)
    for a, b in zip(glog, clog):
        assert abs(a["nnode_yield"] - b["nnode_yield"]) <= max(2, b["nnode_yield"] // 100)
        assert a["global_yield_error"] == pytest.approx(b["global_yield_error"], rel=1e-3)
    same = to_host(ds["plastic_yield"]) == cs["plastic_yield"]
    assert same.mean() > 0.99
    np.testing.assert_allclose(to_host(ds["etas1"])[same], cs["etas1"][same], rtol=1e-5)
    for k in ("vx1", "vy1"):
        ref = csol[k]
        assert np.linalg.norm(to_host(gsol[k]) - ref) <= 1e-5 * np.linalg.norm(ref)


def test_picard_stops_at_first_iteration_when_prior_velocity_converged(gpu_ctx):
    """The change norm of the first iteration compares with the previous time
    step's velocity (backup_stokes_arrays!), so a loop that starts from a
    converged state and that velocity stops after one solve, as the
    reference does."""
    print("picard loop restarted from a converged state", flush=True)
    g, s, template = _setup()
    nglobal, tol = 20, 1e-2
    cs, csol, clog = _cpu_loop(g, s, template, nglobal, tol)
    assert clog[-1]["dvxy_rel_L2"] < tol, "the first loop must stop on the change norm"
    vold = np.concatenate([csol["vx1"].ravel(order="F"), csol["vy1"].ravel(order="F")])
    cs2, _, clog2 = _cpu_loop(g, cs, template, nglobal, tol, old=vold)
    gpu_ctx.set_grid(g)
    ds = {k: to_device(v) for k, v in s.items()}
    gsol, glog = picard_loop_nodes(gpu_ctx, template, ds, PAR, nglobal, tol, solve_opts=dict(rtol=1e-11))
    gsol2, glog2 = picard_loop_nodes(gpu_ctx, template, ds, PAR, nglobal, tol, solve_opts=dict(rtol=1e-11),
                                     velocity_old=(gsol["vx1"], gsol["vy1"]))
    print("cpu", [f"{r['dvxy_rel_L2']:.3g}" for r in clog2], "gpu", [f"{r['dvxy_rel_L2']:.3g}" for r in glog2],
          flush=True)
    assert glog[0]["dvxy_rel_L2"] == pytest.approx(1.0) and clog[0]["dvxy_rel_L2"] == pytest.approx(1.0)
    assert len(clog2) == 1 and len(glog2) == 1
    assert glog2[0]["dvxy_rel_L2"] < tol
    assert glog2[0]["dvxy_rel_L2"] == pytest.approx(clog2[0]["dvxy_rel_L2"], rel=0.1, abs=1e-6)


def test_picard_updates_timestep_after_each_solve(gpu_ctx):
    """iupdate_timestep = 1: the displacement limit shortens the step after
    each solve, and the shortened step drives the stress forecast, the
    yielding and the next right-hand side, in both loops alike."""
    print("picard loop with the displacement-limited time step", flush=True)
    g, s, template = _setup()
    nglobal, tol, displ = 4, 1e-3, 0.05
    cs, csol, clog = _cpu_loop(g, s, template, nglobal, tol, iupdate_timestep=1, displ_max=displ)
    gpu_ctx.set_grid(g)
    ds = {k: to_device(v) for k, v in s.items()}
    gsol, glog = picard_loop_nodes(gpu_ctx, template, ds, PAR, nglobal, tol, solve_opts=dict(rtol=1e-11),
                                   iupdate_timestep=1, marker_cell_displ_max=displ)
    print("cpu dt", [r["timestep"] for r in clog], "gpu dt", [r["timestep"] for r in glog], flush=True)
    assert clog[0]["timestep"] < PAR["timestep"], "the limit must bind in this test"
    assert len(glog) == len(clog)
    for a, b in zip(glog, clog):
        assert a["timestep"] == pytest.approx(b["timestep"], rel=1e-6)
        assert abs(a["nnode_yield"] - b["nnode_yield"]) <= max(2, b["nnode_yield"] // 100)
        assert a["global_yield_error"] == pytest.approx(b["global_yield_error"], rel=1e-3)
    assert gsol["timestep"] == glog[-1]["timestep"]


def test_picard_warm_start_matches_oracle_with_fewer_iterations(gpu_ctx):
    """Each solve started from the previous one's solution (pressures
    rescaled to the new pscale): the same loop, yield counts and stopping
    step as the oracle-composed loop, the velocity within the solve's
    tolerance, and fewer Krylov iterations than cold starts."""
    print("picard loop over nodes with warm-started solves", flush=True)
    g, s, template = _setup()
    nglobal, tol = 6, 1e-2
    cs, csol, clog = _cpu_loop(g, s, template, nglobal, tol)
    gpu_ctx.set_grid(g)
    runs = {}
    for warm in (False, True):
        ds = {k: to_device(v) for k, v in s.items()}
        runs[warm] = picard_loop_nodes(gpu_ctx, template, ds, PAR, nglobal, tol, solve_opts=dict(rtol=1e-11),
                                       warm_start=warm)
    gsol, glog = runs[True]
    cold = sum(r["iterations"] for r in runs[False][1])
    warm_it = sum(r["iterations"] for r in glog)
    print(f"krylov iterations cold {cold} warm {warm_it}", flush=True)
    assert len(glog) == len(clog)
    for a, b in zip(glog, clog):
        assert abs(a["nnode_yield"] - b["nnode_yield"]) <= max(2, b["nnode_yield"] // 100)
        assert a["global_yield_error"] == pytest.approx(b["global_yield_error"], rel=1e-3)
    for k in ("vx1", "vy1"):
        ref = csol[k]
        assert np.linalg.norm(to_host(gsol[k]) - ref) <= 1e-5 * np.linalg.norm(ref)
    assert warm_it < cold
