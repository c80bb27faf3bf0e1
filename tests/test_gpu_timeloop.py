"""The device time step of the rifting model (marker rheology, scatter,
Picard loop over nodes, grid -> marker viscosity, time-step limit, post-solve
marker updates, RK4 with stress rotation, marker strain) against the same
steps composed from the oracle with a direct solve, on a small grid of the
rifting workload: two steps; yield counts within 1 %, the time step to 1e-6,
marker positions within 1e-10 of the domain size, marker viscosities and
stresses to 1e-6 of their range.  With the heat path on, marker and grid
temperatures agree to 1e-6 K (the heat solve stands in for a direct solve
at relative residual 1e-12)."""
import numpy as np
import pytest
import torch

from earthbox_gpu import timeloop, workloads
from earthbox_gpu.device import to_host
from timeloop_oracle import initial_grid, time_step

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("with_heat", [False, True])
def test_rift_time_steps_match_oracle(gpu_ctx, with_heat):
    print(f"rift model, two time steps on 101 x 41 nodes, 4 x 4 markers per cell, heat path {with_heat}",
          flush=True)
    tmpl, m, mats, par = workloads.rift2d_markers(101, 41, nm=4, xsize=1.0e5, ysize=4.0e4, seed=7)
    par = dict(par, nglobal=6, tolerance_picard=1e-3)
    thermal = workloads.rift2d_thermal(tmpl, m, max_temp_change=20.0) if with_heat else None
    model = timeloop.upload(gpu_ctx, tmpl, m, mats, par, thermal)
    mo = {k: v.copy() for k, v in m.items()}
    if with_heat:
        mo.update({k: v.copy() for k, v in thermal[0].items()})
    grid = initial_grid(tmpl.grid)
    dt, ts, vold = par["timestep"], 0.0, None
    L = max(tmpl.grid.xsize, tmpl.grid.ysize)
    for step in (1, 2):
        glog = timeloop.time_step(gpu_ctx, model, solve_opts=dict(rtol=1e-11))
        dt_new, vx1, vy1, clog = time_step(tmpl, mo, mats, par, grid, dt, ts, step, vold,
                                           thermal[1:] if with_heat else None)
        ts += dt_new
        dt = dt_new
        vold = np.concatenate([vx1.ravel(order="F"), vy1.ravel(order="F")])
        print(f"step {step}: gpu dt {glog['timestep']:.6e} picard {glog['picard_iterations']} yield "
              f"{glog['nnode_yield']}; oracle dt {dt_new:.6e} picard {len(clog)} yield "
              f"{[r['nnode_yield'] for r in clog]}", flush=True)
        assert glog["max_relres"] <= 1e-10
        assert glog["picard_iterations"] == len(clog)
        for a, b in zip(glog["nnode_yield"], [r["nnode_yield"] for r in clog]):
            assert abs(a - b) <= max(2, b // 100)
        assert glog["timestep"] == pytest.approx(dt_new, rel=1e-6)
        gm = {k: to_host(v) for k, v in model.markers.items() if k in mo}
        err = max(np.abs(gm["x"] - mo["x"]).max(), np.abs(gm["y"] - mo["y"]).max())
        print(f"  positions: max |d| = {err:.3e} m ({err / L:.2e} of the domain)", flush=True)
        assert err <= 1e-10 * L
        if with_heat:
            print(f"  heat: {glog['heat']}", flush=True)
            dtk = np.abs(gm["TK"] - mo["TK"]).max()
            print(f"  marker temperature: max |d| = {dtk:.3e} K", flush=True)
            assert dtk <= 1e-6 and glog["heat"]["converged"]
            assert np.abs(to_host(model.grid["tk2"]) - grid["tk2"]).max() <= 1e-6
        for k in ("eta_flow", "eta", "sxx", "sxy", "GII", "strain_plastic"):
            scale = np.abs(mo[k]).max() + 1e-300
            d = np.abs(gm[k] - mo[k])
            print(f"  {k}: max |d| / range {d.max() / scale:.2e}, share above 1e-6: {(d > 1e-6 * scale).mean():.2e}",
                  flush=True)
            assert (d > 1e-6 * scale).mean() <= 1e-3, k
    assert model.ntimestep == 2 and model.timesum == pytest.approx(ts, rel=1e-6)


def test_rift_time_steps_marker_plasticity_match_oracle(gpu_ctx):
    """itype_global 0: the Picard loop over markers (marker yielding,
    viscosity-only scatter, per-solve marker gathers) inside the time step;
    failed-marker counts within 1 %, the rest as the node-based test."""
    print("rift model, two time steps on 101 x 41 nodes, Picard loop over markers", flush=True)
    tmpl, m, mats, par = workloads.rift2d_markers(101, 41, nm=4, xsize=1.0e5, ysize=4.0e4, seed=7)
    par = dict(par, nglobal=6, tolerance_picard=1e-3, itype_global=0, itype_plasticity=0)
    model = timeloop.upload(gpu_ctx, tmpl, m, mats, par)
    mo = {k: v.copy() for k, v in m.items()}
    grid = initial_grid(tmpl.grid)
    dt, ts, vold = par["timestep"], 0.0, None
    L = max(tmpl.grid.xsize, tmpl.grid.ysize)
    failed = 0
    for step in (1, 2):
        glog = timeloop.time_step(gpu_ctx, model, solve_opts=dict(rtol=1e-11))
        dt_new, vx1, vy1, clog = time_step(tmpl, mo, mats, par, grid, dt, ts, step, vold)
        ts += dt_new
        dt = dt_new
        vold = np.concatenate([vx1.ravel(order="F"), vy1.ravel(order="F")])
        cf = [r["nmarker_failed"] for r in clog]
        print(f"step {step}: gpu dt {glog['timestep']:.6e} picard {glog['picard_iterations']} failed "
              f"{glog['nmarker_failed']}; oracle dt {dt_new:.6e} picard {len(clog)} failed {cf}", flush=True)
        assert glog["max_relres"] <= 1e-10
        assert glog["picard_iterations"] == len(clog)
        for a, b in zip(glog["nmarker_failed"], cf):
            assert (a is None) == (b is None)
            if b is not None:
                assert abs(a - b) <= max(2, b // 100)
                failed = max(failed, b)
        assert glog["timestep"] == pytest.approx(dt_new, rel=1e-6)
        gm = {k: to_host(v) for k, v in model.markers.items() if k in mo}
        err = max(np.abs(gm["x"] - mo["x"]).max(), np.abs(gm["y"] - mo["y"]).max())
        print(f"  positions: max |d| = {err:.3e} m ({err / L:.2e} of the domain)", flush=True)
        assert err <= 1e-10 * L
        for k in ("eta", "sxx", "sxy", "GII", "strain_plastic"):
            scale = np.abs(mo[k]).max() + 1e-300
            d = np.abs(gm[k] - mo[k])
            print(f"  {k}: max |d| / range {d.max() / scale:.2e}, share above 1e-6: {(d > 1e-6 * scale).mean():.2e}",
                  flush=True)
            assert (d > 1e-6 * scale).mean() <= 2e-3, k
    assert failed > 0, "the rift markers must yield"
