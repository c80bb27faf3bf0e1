"""Model time steps on device arrays, in the order of the reference's time
loop (time_loop/TimeLoop.jl:163-273), restricted to the hot path: marker
rheology, marker -> grid scatter, the Picard loop over nodes (itype_global 1;
over markers with 0) with the Stokes solve, grid -> marker viscoplastic
viscosity (node-based plasticity only), the displacement-limited time step,
the post-solve grid and marker updates, RK4 advection with stress rotation and
the marker strain update.  Everything stays on the device between steps; only
scalars (solver statistics, yield counts, the time step) come back to the host.

With thermal inputs (upload's thermal argument) the step also runs the heat
path: the heat-field scatter with the thermal boundary conditions before the
Stokes scatter (TimeLoop.jl:219-221), HeatSolver.update_marker_temp_for_grid_and_subgrid_changes!
after the marker stress update (TimeLoop.jl:250-257) and the sticky
temperature correction after advection (TimeLoop.jl:268-269).  Without them
marker temperatures stay as given.

Not in the loop (outside the accelerated path): the rock-property updates
(conductivity, rho*Cp and the adiabatic term stay at their uploaded values),
melting, surface processes, output and the transient boundary conditions.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np
import torch

from . import heat, rheology
from .device import empty, to_device
from .marker_gather import OUTPUTS as GATHER_OUTPUTS, STRESS as GATHER_STRESS, update_markers_after_stokes
from .picard import picard_loop_markers, picard_loop_nodes
from .plastic import interpolate_basic_viscoplastic_viscosity_to_pressure_grid
from .scatter import interpolate_marker_stokes_parameters_to_grid, new_grid_fields
from .stokes import StokesSolverGPU, calculate_new_timestep
from .stokes_problem import StokesProblem
from .viscoelastic import update_grid_strain_rate_and_stress

MARKER_FLOAT64 = ("x", "y", "TK", "rho", "pr", "sxx", "sxy", "exx", "exy", "sr_ratio", "GII", "strain_plastic",
                  "strain_rate_plastic", "fric_ini", "cohesion", "fric", "eta_flow", "preexp", "eta")


@dataclasses.dataclass
class DeviceModel:
    """The model state a time step reads and writes, on the device."""
    template: StokesProblem          # grid, velocity BCs, gravity
    markers: dict                    # marker arrays (MARKER_FLOAT64, matid int16, pfailure float32, ...)
    mats: dict                       # material tables
    params: dict                     # scalars (rheology, plasticity, Picard)
    grid: dict                       # grid fields (scatter outputs, Picard state)
    timestep: float
    timesum: float = 0.0
    ntimestep: int = 0
    velocity_old: tuple | None = None
    solver: StokesSolverGPU | None = None
    thermal_bc: dict | None = None   # device thermal boundary conditions; None: no heat path
    last_solution: tuple | None = None   # (S, pscale) of the last Stokes solve (warm starts)


def upload(ctx, template: StokesProblem, markers: dict, mats: dict, params: dict,
           thermal: tuple | None = None) -> DeviceModel:
    """Copy a model (numpy arrays, as workloads.rift2d_markers returns it) to
    the device and allocate the grid fields and marker work arrays.  thermal:
    (marker arrays kt, rhocp, ha; mats with mat_hr; thermal boundary
    conditions; heat scalars), as workloads.rift2d_thermal returns them, turns
    the heat path on."""
    g = template.grid
    ctx.set_grid(g)
    n = markers["x"].size
    if thermal is not None:
        therm, tmats, tbc, hpar = thermal
        markers = {**markers, **therm}
        mats = {**mats, **tmats}
        params = {**params, **hpar}
    dm = {k: to_device(v) for k, v in markers.items()}
    dev = dm["x"].device
    dm["inside_flags"] = torch.empty(n, dtype=torch.int8, device=dev)
    dm["marker_xn"] = torch.zeros(n, dtype=torch.int32, device=dev)
    dm["marker_yn"] = torch.zeros(n, dtype=torch.int32, device=dev)
    dm["mat_mu"] = to_device(mats["mat_mu"])
    grid = new_grid_fields(g.xnum, g.ynum, 0.0)
    ny, nx = g.ynum, g.xnum
    for k, shp in (("plastic_yield", (ny, nx)), ("yield_error", (ny, nx)),
                   ("eii_plastic_pressure", (ny - 1, nx - 1)), ("esp", (ny, nx))):
        grid[k] = empty(shp, np.float64)
        grid[k].zero_()
    bc = None
    if thermal is not None:
        # tk2 is the backup of tk1 (the previous heat solution); zero before
        # the first step, where every node receives marker weight
        for k in ("tk1", "tk2", "kt1", "rhocp1", "hr1", "ha1"):
            grid[k] = empty((ny, nx), np.float64)
            grid[k].zero_()
        bc = {k: to_device(v) for k, v in tbc.items()}
    return DeviceModel(template=template, markers=dm, mats={k: to_device(v) for k, v in mats.items()},
                       params=dict(params), grid=grid, timestep=float(params["timestep"]), thermal_bc=bc)


def time_step(ctx, model: DeviceModel, solve_opts: dict | None = None, rk4_arithmetic: int | None = None) -> dict:
    """One time step (TimeLoop.jl:185-190 for the hot path).  Returns a log
    with the Picard iterations, the Krylov iterations and solve times (set-up
    included) of every solve, yield counts and the new time step."""
    p, m, gf = model.params, model.markers, model.grid
    g = model.template.grid
    t_step = time.perf_counter()
    model.solver = model.solver or StokesSolverGPU(ctx)
    # pre-solver steps (execute_pre_solver_steps!, TimeLoop.jl:206-222)
    ctx.get_marker_inside_flags(m["x"], m["y"], out=m["inside_flags"])
    step = model.ntimestep + 1   # the reference's time-step counter of this step
    rpar = dict(p, timestep=model.timestep, ntimestep=step)
    heat_on = model.thermal_bc is not None
    if heat_on:   # Interpolation.interpolate_marker_heat_parameters_to_grid! (TimeLoop.jl:220)
        gf["tk1"].copy_(gf["tk2"])
        heat.interpolate_marker_heat_parameters_to_grid(
            ctx, m, model.mats, {k: gf[k] for k in heat.HEAT_FIELDS}, model.thermal_bc)
    rheology.update_marker_rheology(ctx, m, model.mats, rpar, m)
    interpolate_marker_stokes_parameters_to_grid(ctx, {**m, "dilatation": None, "meltfrac": None},
                                                 {**gf, "marker_xn": m["marker_xn"], "marker_yn": m["marker_yn"]})
    interpolate_basic_viscoplastic_viscosity_to_pressure_grid(ctx, gf["etas1"], gf["plastics"], gf["plasticn"],
                                                              gf["plastic_yield"], gf["etan1"])
    # Stokes-continuity solver steps (execute_stokes_continuity_solver_steps!, TimeLoop.jl:230-243)
    plastic_params = dict(timestep=model.timestep, timesum=model.timesum, ntimestep=step,
                          viscosity_min=p["viscosity_min"], viscosity_max=p["viscosity_max"],
                          iuse_fluid_pressure_for_yield=p.get("iuse_fluid_pressure_for_yield", 0))
    node_based = int(p.get("itype_global", 1)) == 1
    warm = bool(int(p.get("warm_start", 1)))   # solves start from the previous solution
    if node_based:
        sol, plog = picard_loop_nodes(ctx, model.template, gf, plastic_params, int(p["nglobal"]),
                                      float(p["tolerance_picard"]), solve_opts=solve_opts, solver=model.solver,
                                      velocity_old=model.velocity_old, warm_start=warm,
                                      initial_guess=model.last_solution)
    else:   # PlasticityLoopMarkers.picard_loop (itype_global 0)
        mpar = {k: p.get(k, d) for k, d in (("y_sealevel", 0.0), ("yield_stress_min", 0.0),
                                             ("yield_stress_max", 1e32), ("no_yielding_in_mobile_wall", 0))}
        mpar.update(plastic_params)
        sol, plog = picard_loop_markers(ctx, model.template, gf, m, m["mat_mu"], mpar, int(p["nglobal"]),
                                        float(p["tolerance_picard"]), forecast=int(p.get("itype_plasticity", 0)),
                                        matid_domains=tuple(int(p.get(k, -1)) for k in (
                                            "matid_sticky_air", "matid_mobile_wall", "matid_plate_extension")),
                                        solve_opts=solve_opts, solver=model.solver, velocity_old=model.velocity_old,
                                        warm_start=warm, initial_guess=model.last_solution)
    vx1, vy1, pr1 = sol["vx1"], sol["vy1"], sol["pr1"]
    if node_based:   # interpolate_grid_viscoplastic_viscosity_to_markers! (node-based plasticity only)
        rheology.calc_marker_viscoplastic_viscosity(ctx, m, gf["plastic_yield"], gf["etas1"])
    dt, _ = calculate_new_timestep(ctx, vx1, vy1, model.timestep, p["marker_cell_displ_max"])
    # execute_post_stokes_solver_steps! with the new step, then the marker stress update
    ss = update_grid_strain_rate_and_stress(ctx, {"vx1": vx1, "vy1": vy1, **gf}, dt)
    fields = {"exx": ss["exx"], "exy": ss["exy"], "pr1": pr1, "sxx2": ss["sxx2"], "sxy2": ss["sxy2"],
              "dsxx": ss["dsxx"], "dsxy": ss["dsxy"]}
    gm = {"x": m["x"], "y": m["y"], "marker_xn": m["marker_xn"], "marker_yn": m["marker_yn"],
          "inside_flags": m["inside_flags"], "eta": m["eta"], "matid": m["matid"], "mat_mu": m["mat_mu"]}
    out = {"marker_exx": m["exx"], "marker_exy": m["exy"], "marker_pr": m["pr"], "marker_sr_ratio": m["sr_ratio"],
           "marker_sxx": m["sxx"], "marker_sxy": m["sxy"]}
    update_markers_after_stokes(ctx, gm, fields, dt, out, GATHER_OUTPUTS + GATHER_STRESS)
    hlog = {}
    if heat_on:   # execute_heat_solver_steps! (TimeLoop.jl:250-257)
        hf = {k: gf[k] for k in heat.HEAT_FIELDS + ("rho1", "etas1", "etan1")}
        hf.update(vx1=vx1, vy1=vy1, sxy2=ss["sxy2"], sxx2=ss["sxx2"])
        t_heat = time.perf_counter()
        hr = heat.update_marker_temp_for_grid_and_subgrid_changes(
            ctx, {**gm, "TK": m["TK"]}, hf, model.thermal_bc, dt, float(p["max_temp_change"]),
            float(p.get("subgrid_diff_coef_temp", 0.0)), int(p.get("iuse_heat", 1)), 0.0,
            float(model.template.gravity_y), int(p.get("iuse_adiabatic_heating", 0)),
            int(p.get("iuse_shear_heating", 0)))
        if hr:
            gf["tk2"].copy_(hr["tk2"])
            hlog = dict(hr["stats"], seconds=time.perf_counter() - t_heat)
    # advection steps (execute_advection_steps!, TimeLoop.jl:259-262)
    ctx.calculate_spin(vx1, vy1, gf["esp"])
    if rk4_arithmetic is not None:
        ctx.set_rk4_arithmetic(rk4_arithmetic)
    ctx.advect_markers_and_rotate_stress_tensor(4, dt, m["inside_flags"], m["x"], m["y"], m["marker_xn"],
                                                m["marker_yn"], vx1, vy1, gf["esp"], None, None, None,
                                                m["sxx"], m["sxy"])
    rheology.update_marker_strain(ctx, m, dt, p.get("plastic_healing_rate", 0.0))
    if heat_on and int(p.get("iuse_sticky_correction", 0)) == 1:
        heat.apply_sticky_temperature_correction(ctx, gm, m["TK"], int(p["matid_sticky_air"]),
                                                 int(p["matid_sticky_water"]), float(p["temperature_top"]))
    # advance_model_time!
    model.timesum += dt
    model.timestep = dt
    model.ntimestep = step
    model.velocity_old = (vx1, vy1)
    model.last_solution = (sol["S"], sol["pscale"])
    torch.cuda.synchronize()
    return {"ntimestep": model.ntimestep, "timestep": dt, "seconds": time.perf_counter() - t_step,
            "picard_iterations": len(plog), "krylov_iterations": [r["iterations"] for r in plog],
            "solve_seconds_with_setup": [r["solve_seconds_with_setup"] for r in plog],
            "max_relres": max(r["relres"] for r in plog),
            "nnode_yield": [r.get("nnode_yield") for r in plog],
            "nmarker_failed": [r.get("nmarker_failed") for r in plog],
            "learned_accepted": sum(r["learned_accepted"] for r in plog),
            "final_dvxy_rel_L2": plog[-1]["dvxy_rel_L2"], "heat": hlog}
