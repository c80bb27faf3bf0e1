"""The Picard (global plasticity) loop over nodes, on device arrays.

picard_loop_nodes mirrors PlasticityLoopNodes.picard_loop
(global_plasticity_loop/plasticity_loop/plasticity_loop_nodes/PlasticityLoopNodes.jl:36-92):
each iteration interpolates the viscoplastic viscosity to the pressure grid
(after the first), forms the viscoelastic terms and right-hand sides, solves
the Stokes system, processes the solution on the grid (strain rates, forecast
stresses, change norms), updates the nodes for plastic yielding, copies the
new viscosity, yield flags and plastic strain rate back when nodes yield, and
stops by LoopTermination.check_stop_global_loop_nodes
(global_plasticity_loop/utils/LoopTermination.jl:16-31).

picard_loop_markers mirrors PlasticityLoopMarkers.picard_loop
(plasticity_loop/plasticity_loop_markers/PlasticityLoopMarkers.jl:36-75):
after the first iteration the markers yield
(MarkerPlasticity.update_plastic_yielding!), their viscosity alone is
scattered back to etas1 and etan1 with the previous values as backups
(BackupTransportArrays.backup_and_clear_viscosity_transport_arrays!,
MarkersToStokesArrays.interpolate_viscosity_to_transport_arrays!) and the
harmonic normal viscosity follows; each solve is processed on the grid and
gathered to the markers (strain rate, strain-rate ratio, pressure:
StokesContinuitySolver.execute_post_stokes_solver_steps!), and the loop stops
by LoopTermination.check_stop_global_loop_markers (LoopTermination.jl:3-14).
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np
import torch

from .device import empty
from .marker_gather import OUTPUTS as GATHER_OUTPUTS, update_markers_after_stokes
from .rheology import VISCOELASTIC_FORECAST, update_marker_plastic_yielding
from .scatter import interpolate_marker_stokes_parameters_to_grid
from .plastic import (interpolate_basic_viscoplastic_viscosity_to_pressure_grid,
                      plastic_strain_rate_at_pressure_nodes, update_nodes_for_plastic_yielding)
from .stokes import StokesSolverGPU, bc_mirror_factors, calculate_new_timestep, change_norms
from .stokes_problem import StokesProblem
from .viscoelastic import calc_viscoelastic_terms_and_rhs, update_grid_strain_rate_and_stress

STATE = ("etas1", "eta_flow", "mus1", "sxy1", "rho1", "rho1_vy", "etan1", "mun1", "sxx1", "plastics",
         "plasticn", "plastic_yield", "cohesion_grid", "fric_degrees_grid", "dilatation_grid",
         "eii_plastic_pressure", "yield_error")


def warm_start_vector(S_prev, pscale_prev: float, pscale: float):
    """The previous solve's solution as the next solve's initial guess: its
    scaled pressures (every third unknown, pr / pscale) rescaled to the new
    pscale.  The reference solves directly, so the initial guess changes the
    result only within the iterative solve's tolerance."""
    S0 = S_prev.clone()
    if pscale != pscale_prev:
        S0[2::3] *= pscale_prev / pscale
    return S0


def check_stop_global_loop_nodes(nnode_yield, iglobal, nglobal, global_yield_error, convergence_criterion,
                                 tolerance) -> bool:
    """LoopTermination.check_stop_global_loop_nodes (LoopTermination.jl:16-31)."""
    return (nnode_yield == 0 or iglobal == nglobal or global_yield_error < tolerance
            or convergence_criterion < tolerance)


def picard_loop_nodes(ctx, template: StokesProblem, state: dict, plastic_params: dict, nglobal: int,
                      tolerance: float, gravity_x: float = 0.0, iuse_harmonic_avg_normal_viscosity: int = 1,
                      no_yielding_in_mobile_wall: bool = False, no_yielding_in_plate_extension: bool = False,
                      solve_opts: dict | None = None, solver: StokesSolverGPU | None = None,
                      velocity_old: tuple | None = None, iupdate_timestep: int = 0,
                      marker_cell_displ_max: float = 0.5, warm_start: bool = False,
                      initial_guess: tuple | None = None):
    """template: the system's grid, velocity boundary conditions, pressure BC,
    gravity and interface-stabilization flag (its fields are replaced each
    iteration); state: device arrays keyed like STATE, updated in place as the
    reference updates ModelData; plastic_params: the scalars of
    update_nodes_for_plastic_yielding (timestep included).
    velocity_old: (vx1, vy1) of the previous time step, which
    backup_stokes_arrays! leaves for calculate_stokes_norms!
    (StokesContinuitySolver.jl:228-236); None on the first time step (zeros,
    so the first change norm is 1).  The change norm is computed on every
    iteration against the previous solution.
    iupdate_timestep = 1: after every solve the time step is limited by the
    displacement limit (TimeStep.adjust_time_step_using_displacement_limit!,
    StokesContinuitySolver.jl:223-226, Update.jl:43-61) before the stress
    forecast and the yielding, and the new step carries into the next
    iteration's right-hand side.
    warm_start: False (every solve from zero, default) or True: each solve
    starts from the previous one's solution, the first from initial_guess
    ((S, pscale) of an earlier solve, e.g. the previous time step's) when
    given, see warm_start_vector.
    Returns (solution dict with vx1, vy1, pr1, the last solution vector S and
    its pscale, exx, exy, eii, sxy2, sxx2, dsxy, dsxx and the final timestep;
    per-iteration log)."""
    g = template.grid
    solver = solver or StokesSolverGPU(ctx)
    opts = dict(precision=1, rtol=1e-10, max_iterations=4000, bc_d=bc_mirror_factors(template))
    opts.update(solve_opts or {})
    dt = float(plastic_params["timestep"])
    plastic_params = dict(plastic_params)
    ny, nx = g.ynum, g.xnum
    eii_basic = empty((ny, nx), np.float64)
    vx_old, vy_old = velocity_old if velocity_old is not None else (None, None)
    log = []
    sol = {}
    for iglobal in range(1, nglobal + 1):
        if iglobal > 1:
            interpolate_basic_viscoplastic_viscosity_to_pressure_grid(
                ctx, state["etas1"], state["plastics"], state["plasticn"], state["plastic_yield"], state["etan1"],
                iuse_harmonic_avg_normal_viscosity)
        # StokesContinuitySolver.solve_viscoelastic_stokes_continuity_equations!
        ve = calc_viscoelastic_terms_and_rhs(ctx, state, dt, gravity_x, template.gravity_y,
                                             template.iuse_interface_stabilization)
        etan11 = float(ve["etan0"][0, 0].item())
        prob = dataclasses.replace(template, pscale=2.0 * etan11 / (g.xstpavg + g.ystpavg), timestep=dt)
        t0 = time.perf_counter()
        solver.set_system(prob, fields={"etas0": ve["etas0"], "etan0": ve["etan0"], "rho1_vy": state["rho1_vy"],
                                        "RX1": ve["RX1"], "RY1": ve["RY1"], "RC1": ve["RC1"]})
        if warm_start and initial_guess is not None:
            S, st = solver.solve_system(warm_start_vector(initial_guess[0], initial_guess[1], prob.pscale),
                                        **{**opts, "use_initial_guess": 1})
        else:
            S, st = solver.solve_system(**opts)   # returns after the device finished (stats read back)
        initial_guess = (S, prob.pscale)
        solve_seconds = time.perf_counter() - t0
        vx1, vy1, pr1 = solver.process_stokes_solution(S)
        # stokes_solution_processing! on the grid: the time step limit (the
        # strain rates do not depend on it), then the stress forecast with it
        if iupdate_timestep == 1:
            dt, _ = calculate_new_timestep(ctx, vx1, vy1, dt, marker_cell_displ_max)
            plastic_params["timestep"] = dt
        ss = update_grid_strain_rate_and_stress(ctx, {"vx1": vx1, "vy1": vy1, **state}, dt)
        dvxy_rel_L2 = change_norms(ctx, [vx1, vy1], [vx_old, vy_old])[1]
        vx_old, vy_old = vx1.clone(), vy1.clone()
        # PlasticFailureNodes.update_nodes_for_plastic_yielding
        fields = {k: state[k] for k in ("plastic_yield", "etas1", "eta_flow", "mus1", "cohesion_grid",
                                        "fric_degrees_grid")}
        fields.update(sxy2=ss["sxy2"], sxx2=ss["sxx2"], pr1=pr1)
        out, nnode_yield, global_yield_error = update_nodes_for_plastic_yielding(
            ctx, fields, state["yield_error"], plastic_params, no_yielding_in_mobile_wall,
            no_yielding_in_plate_extension)
        if nnode_yield > 0:
            state["etas1"].copy_(out["etavp_new"])
            state["plastic_yield"].copy_(out["plastic_yield_new"])
            eii_basic.copy_(out["plastic_strain_rate"])
            plastic_strain_rate_at_pressure_nodes(ctx, eii_basic, state["eii_plastic_pressure"])
        log.append({"iglobal": iglobal, "iterations": st["iterations"], "relres": st["relres"],
                    "nnode_yield": nnode_yield, "global_yield_error": global_yield_error,
                    "dvxy_rel_L2": dvxy_rel_L2, "timestep": dt, "solve_seconds_with_setup": solve_seconds,
                    "learned_accepted": st["learned_accepted"]})
        sol = {"vx1": vx1, "vy1": vy1, "pr1": pr1, "S": S, "pscale": prob.pscale, "timestep": dt, **ss}
        if check_stop_global_loop_nodes(nnode_yield, iglobal, nglobal, global_yield_error, dvxy_rel_L2,
                                        tolerance):
            break
    torch.cuda.synchronize()
    return sol, log


def check_stop_global_loop_markers(iglobal, nglobal, convergence_criterion, tolerance) -> bool:
    """LoopTermination.check_stop_global_loop_markers (LoopTermination.jl:3-14)."""
    return iglobal == nglobal or convergence_criterion < tolerance


MARKERS = ("x", "y", "inside_flags", "marker_xn", "marker_yn", "matid", "eta", "sxx", "sxy", "exx", "exy",
           "sr_ratio", "pr", "cohesion", "fric", "pfailure")


def picard_loop_markers(ctx, template: StokesProblem, state: dict, markers: dict, mat_mu, plastic_params: dict,
                        nglobal: int, tolerance: float, gravity_x: float = 0.0,
                        iuse_harmonic_avg_normal_viscosity: int = 1, forecast: int = VISCOELASTIC_FORECAST,
                        matid_domains: tuple = (-1, -1, -1), solve_opts: dict | None = None,
                        solver: StokesSolverGPU | None = None, velocity_old: tuple | None = None,
                        iupdate_timestep: int = 0, marker_cell_displ_max: float = 0.5,
                        warm_start: bool = False, initial_guess: tuple | None = None):
    """template, state, solve_opts, solver, velocity_old, iupdate_timestep,
    warm_start, initial_guess: as picard_loop_nodes (state needs etas1, etan1, mus1, mun1, sxy1, sxx1,
    rho1, rho1_vy and, for the harmonic normal viscosity, plastics, plasticn,
    plastic_yield).  markers: device vectors keyed like MARKERS (marker_xn,
    marker_yn as the scatter before the loop stored them); eta, sxx, sxy and
    pfailure change where markers yield, and exx, exy, sr_ratio, pr hold the
    last solve's gathers on return (on entry: the previous time step's, which
    the first yielding reads).  mat_mu: device shear moduli by material.
    plastic_params: the scalars of rheology.update_marker_plastic_yielding
    (timestep included; ntimestep > 0 lets markers yield).  matid_domains:
    (sticky air, mobile wall, plate extension) ids, used when
    plastic_params["no_yielding_in_mobile_wall"] is 1.
    Returns (solution dict as picard_loop_nodes; per-iteration log with the
    number of markers whose pfailure is 1 after the yielding step, None on
    the first iteration)."""
    g = template.grid
    solver = solver or StokesSolverGPU(ctx)
    opts = dict(precision=1, rtol=1e-10, max_iterations=4000, bc_d=bc_mirror_factors(template))
    opts.update(solve_opts or {})
    dt = float(plastic_params["timestep"])
    plastic_params = dict(plastic_params)
    n = int(markers["x"].numel())
    vx_old, vy_old = velocity_old if velocity_old is not None else (None, None)
    scatter_markers = {k: markers[k] for k in ("x", "y", "inside_flags", "eta")}
    gather_markers = {k: markers[k] for k in ("x", "y", "marker_xn", "marker_yn", "inside_flags", "eta", "matid")}
    gather_markers["mat_mu"] = mat_mu
    gather_out = {"marker_exx": markers["exx"], "marker_exy": markers["exy"], "marker_pr": markers["pr"],
                  "marker_sr_ratio": markers["sr_ratio"]}
    log = []
    sol = {}
    for iglobal in range(1, nglobal + 1):
        nfailed = None
        if iglobal > 1:
            plastic_params["timestep"] = dt
            update_marker_plastic_yielding(ctx, markers, mat_mu, plastic_params, forecast, *matid_domains)
            nfailed = int((markers["pfailure"] == 1).sum().item()) if n else 0
            # etas1/etan1 enter as the backups kept where a node receives no weight
            interpolate_marker_stokes_parameters_to_grid(ctx, scatter_markers,
                                                         {"etas1": state["etas1"], "etan1": state["etan1"]})
            interpolate_basic_viscoplastic_viscosity_to_pressure_grid(
                ctx, state["etas1"], state["plastics"], state["plasticn"], state["plastic_yield"], state["etan1"],
                iuse_harmonic_avg_normal_viscosity)
        ve = calc_viscoelastic_terms_and_rhs(ctx, state, dt, gravity_x, template.gravity_y,
                                             template.iuse_interface_stabilization)
        etan11 = float(ve["etan0"][0, 0].item())
        prob = dataclasses.replace(template, pscale=2.0 * etan11 / (g.xstpavg + g.ystpavg), timestep=dt)
        t0 = time.perf_counter()
        solver.set_system(prob, fields={"etas0": ve["etas0"], "etan0": ve["etan0"], "rho1_vy": state["rho1_vy"],
                                        "RX1": ve["RX1"], "RY1": ve["RY1"], "RC1": ve["RC1"]})
        if warm_start and initial_guess is not None:
            S, st = solver.solve_system(warm_start_vector(initial_guess[0], initial_guess[1], prob.pscale),
                                        **{**opts, "use_initial_guess": 1})
        else:
            S, st = solver.solve_system(**opts)
        initial_guess = (S, prob.pscale)
        solve_seconds = time.perf_counter() - t0
        vx1, vy1, pr1 = solver.process_stokes_solution(S)
        # stokes_solution_processing!: time step limit, grid stresses, marker gathers, norms
        if iupdate_timestep == 1:
            dt, _ = calculate_new_timestep(ctx, vx1, vy1, dt, marker_cell_displ_max)
        ss = update_grid_strain_rate_and_stress(ctx, {"vx1": vx1, "vy1": vy1, **state}, dt)
        fields = {"exx": ss["exx"], "exy": ss["exy"], "pr1": pr1, "sxx2": ss["sxx2"], "sxy2": ss["sxy2"],
                  "dsxx": ss["dsxx"], "dsxy": ss["dsxy"]}
        update_markers_after_stokes(ctx, gather_markers, fields, dt, gather_out, GATHER_OUTPUTS)
        dvxy_rel_L2 = change_norms(ctx, [vx1, vy1], [vx_old, vy_old])[1]
        vx_old, vy_old = vx1.clone(), vy1.clone()
        log.append({"iglobal": iglobal, "iterations": st["iterations"], "relres": st["relres"],
                    "nmarker_failed": nfailed, "dvxy_rel_L2": dvxy_rel_L2, "timestep": dt,
                    "solve_seconds_with_setup": solve_seconds, "learned_accepted": st["learned_accepted"]})
        sol = {"vx1": vx1, "vy1": vy1, "pr1": pr1, "S": S, "pscale": prob.pscale, "timestep": dt, **ss}
        if check_stop_global_loop_markers(iglobal, nglobal, dvxy_rel_L2, tolerance):
            break
    torch.cuda.synchronize()
    return sol, log
