"""The marker plastic-yielding step and the Picard loop over markers on the GPU
against the oracle.

The yielding step (MarkerPlasticity.update_plastic_yielding!) on random
markers: the device runs the same arithmetic except asin/cos of the yield
stress and the square roots, so values agree to 1e-14 relative (the elastic
forecast's viscosity to 1e-14 times sy/(F - sy), the amplification of its
difference of forecast F and yield stress sy) and the yield flags agree
except for markers whose forecast sits within 1e-12 of the yield stress.

The loop (PlasticityLoopMarkers.picard_loop) on a 41 x 41 falling block with
16 markers per cell and a cohesion that the stresses around the block
exceed, against the same loop composed from the oracle's restatements with a
sparse direct solve: the loop length and the change norms agree, the
failure flags of 99 % of the markers, their viscosities to the iterative
solve's tolerance where the flags match, and the velocities.
"""
import numpy as np
import pytest

import oracle
from earthbox_gpu import rheology, workloads
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.grid import uniform_grid
from earthbox_gpu.picard import picard_loop_markers
from earthbox_gpu.stokes_problem import StokesProblem, free_slip_bcs
from marker_plastic_cases import AIR, PARAMS, PLATE, WALL, markers as random_markers
from timeloop_oracle import picard_markers

pytestmark = pytest.mark.gpu

MARKER_IN = ("y", "cohesion", "fric", "exx", "exy", "sr_ratio", "pr")


def _forecast(m, par, kind):
    mu, dt, r = m["mat_mu"][m["matid"] - 1], par["timestep"], m["sr_ratio"]
    if kind == 0:
        z = m["eta"] / (m["eta"] + dt * mu)
        return np.hypot(m["sxx"] * z + 2 * m["eta"] * m["exx"] * r * (1 - z),
                        m["sxy"] * z + 2 * m["eta"] * m["exy"] * r * (1 - z))
    return np.hypot(m["sxx"] + 2 * mu * dt * m["exx"] * r, m["sxy"] + 2 * mu * dt * m["exy"] * r)


def _sy(m, par):
    c, f, p = m["cohesion"], m["fric"], m["pr"]
    if par["iuse_fluid_pressure_for_yield"]:
        pf = 1000.0 * 9.81 * np.maximum(0.0, m["y"] - par["y_sealevel"])
        p = np.where(p < pf, 0.0, p - pf)
    return np.clip(np.maximum(c * np.sqrt(1 - f * f) + f * p, 0.0), par["yield_stress_min"],
                   par["yield_stress_max"])


@pytest.mark.parametrize("kind,variant", [(0, "plain"), (1, "plain"), (0, "fluid"), (0, "wall"), (1, "limits")])
def test_marker_plastic_yielding_matches_oracle(gpu_ctx, kind, variant):
    print(f"marker plastic yielding forecast={kind} {variant}", flush=True)
    m = random_markers(200_000, 21 + kind)
    par = dict(PARAMS)
    dom = (-1, -1, -1)
    if variant == "fluid":
        par["iuse_fluid_pressure_for_yield"] = 1
    elif variant == "wall":
        par.update(no_yielding_in_mobile_wall=1, matid_sticky_air=AIR, matid_mobile_wall=WALL,
                   matid_plate_extension=PLATE)
        dom = (AIR, WALL, PLATE)
    elif variant == "limits":
        par.update(viscosity_min=1e20, viscosity_max=1e23, yield_stress_min=5e6, yield_stress_max=5e7)
    exp = oracle.update_marker_plastic_yielding(m, dict(par, forecast=kind))
    dm = {k: to_device(m[k]) for k in MARKER_IN + ("matid", "sxx", "sxy", "eta", "pfailure")}
    dm["inside_flags"] = to_device(m["flags"])
    rheology.update_marker_plastic_yielding(gpu_ctx, dm, to_device(m["mat_mu"]), par, kind, *dom)
    gpu_ctx.synchronize()
    got = [to_host(dm[k]) for k in ("sxx", "sxy", "eta", "pfailure")]
    F, sy = _forecast(m, par, kind), _sy(m, par)
    edge = np.abs(F - sy) <= 1e-12 * sy
    ok = ~edge
    assert (exp[3] != m["pfailure"]).sum() > 1000, "the case must yield"
    np.testing.assert_array_equal(got[3][ok], exp[3][ok])
    for a, b in zip(got[:2], exp[:2]):
        np.testing.assert_allclose(a[ok], b[ok], rtol=1e-14, atol=0)
    # the elastic correction mu*dt*sy/(F - sy) amplifies an ulp of sy or F by sy/(F - sy)
    amp = np.maximum(1.0, sy / np.abs(F - sy)) if kind == 1 else 1.0
    assert np.all((np.abs(got[2] - exp[2]) <= 1e-14 * amp * np.abs(exp[2]))[ok])


def _setup(n=41, per_cell=4, seed=5):
    g = uniform_grid(n, n, 1e6, 1e6)
    x, y = workloads.marker_positions(g, per_cell, per_cell, seed)
    nm = x.size
    rng = np.random.default_rng(seed)
    blk = (np.abs(x - 5e5) <= 1e5) & (np.abs(y - 5e5) <= 1e5)
    flags = oracle.inside_flags(x, y, g.limits)
    m = {"x": x, "y": y, "inside_flags": flags, "matid": np.ones(nm, np.int16), "mat_mu": np.array([1e11]),
         "eta": np.where(blk, 1e24, 1e21), "rho": np.where(blk, 3300.0, 3200.0),
         "cohesion": np.full(nm, 4e7), "fric": np.zeros(nm),
         "sxx": np.zeros(nm), "sxy": np.zeros(nm), "exx": np.zeros(nm), "exy": np.zeros(nm),
         "sr_ratio": np.ones(nm), "pr": np.zeros(nm),
         "pfailure": (rng.random(nm) < 0.05).astype(np.float32)}
    sc = oracle.scatter_stokes(g, m, {})
    ny, nx = g.ynum, g.xnum
    s = {k: sc[k] for k in ("etas1", "etan1", "mus1", "mun1", "sxy1", "sxx1", "rho1", "rho1_vy", "plastics",
                            "plasticn")}
    s["plastic_yield"] = np.zeros((ny, nx), order="F")
    m["marker_xn"], m["marker_yn"] = sc["marker_xn"], sc["marker_yn"]
    z = lambda shp: np.zeros(shp, order="F")
    template = StokesProblem(grid=g, etas0=s["etas1"], etan0=s["etan1"], rho1_vy=s["rho1_vy"], RX1=z((ny + 1, nx)),
                             RY1=z((ny, nx + 1)), RC1=z((ny - 1, nx - 1)), bcs=free_slip_bcs(g), gravity_y=10.0)
    par = dict(PARAMS, ntimestep=1, timestep=1e12)
    return g, s, m, template, par


def _cpu_loop(g, s, m, template, par, nglobal, tol, kind=0, iupdate_timestep=0, displ_max=0.5):
    s = {k: v.copy(order="F") for k, v in s.items()}
    m = {k: v.copy() for k, v in m.items()}
    vx1, vy1, pr1, _, log = picard_markers(g, template, s, m, dict(par, forecast=kind), nglobal, tol, None,
                                           iupdate_timestep, displ_max)
    return s, m, {"vx1": vx1, "vy1": vy1, "pr1": pr1}, log


DEV_KEYS = ("x", "y", "inside_flags", "marker_xn", "marker_yn", "matid", "eta", "sxx", "sxy", "exx", "exy",
            "sr_ratio", "pr", "cohesion", "fric", "pfailure")


@pytest.mark.parametrize("kind,iupdate", [(0, 0), (1, 1)])
def test_picard_loop_markers_matches_oracle_loop(gpu_ctx, kind, iupdate):
    print(f"picard loop over markers, 41 x 41 falling block, forecast={kind} iupdate_timestep={iupdate}",
          flush=True)
    g, s, m, template, par = _setup()
    nglobal, tol, displ = 6, 1e-3, 0.05
    cs, cm, csol, clog = _cpu_loop(g, s, m, template, par, nglobal, tol, kind, iupdate, displ)
    gpu_ctx.set_grid(g)
    ds = {k: to_device(v) for k, v in s.items()}
    dm = {k: to_device(m[k]) for k in DEV_KEYS}
    gsol, glog = picard_loop_markers(gpu_ctx, template, ds, dm, to_device(m["mat_mu"]), par, nglobal, tol,
                                     forecast=kind, solve_opts=dict(rtol=1e-11), iupdate_timestep=iupdate,
                                     marker_cell_displ_max=displ)
    print("cpu", [(r["nmarker_failed"], f"{r['dvxy_rel_L2']:.4g}", r["timestep"]) for r in clog], flush=True)
    print("gpu", [(r["nmarker_failed"], f"{r['dvxy_rel_L2']:.4g}", r["timestep"]) for r in glog], flush=True)
    assert len(clog) >= 3 and clog[1]["nmarker_failed"] > (m["pfailure"] == 1).sum() + 100, \
        "the test problem must yield"
    assert len(glog) == len(clog)
    for a, b in zip(glog, clog):
        assert a["timestep"] == pytest.approx(b["timestep"], rel=1e-6)
        assert a["dvxy_rel_L2"] == pytest.approx(b["dvxy_rel_L2"], rel=1e-3, abs=1e-8)
        if b["nmarker_failed"] is not None:
            assert abs(a["nmarker_failed"] - b["nmarker_failed"]) <= max(3, b["nmarker_failed"] // 200)
    same = to_host(dm["pfailure"]) == cm["pfailure"]
    assert same.mean() > 0.99
    np.testing.assert_allclose(to_host(dm["eta"])[same], cm["eta"][same], rtol=1e-5)
    for k in ("vx1", "vy1"):
        ref = csol[k]
        assert np.linalg.norm(to_host(gsol[k]) - ref) <= 1e-5 * np.linalg.norm(ref)
    close = np.isclose(to_host(ds["etas1"]), cs["etas1"], rtol=1e-5, atol=0)
    assert close.mean() > 0.98
