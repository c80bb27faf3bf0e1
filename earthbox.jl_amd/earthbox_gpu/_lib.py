"""ctypes binding of ``libearthbox_gpu.so`` (the C ABI in include/earthbox_gpu.h).

The library is loaded from this directory.  There is no fallback: if the
shared library is missing or a symbol is absent, importing raises.
"""
from __future__ import annotations

import ctypes as C
import pathlib

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libearthbox_gpu.so"

EBG_OK = 0
EBG_NOT_CONVERGED = 1

c_int64 = C.c_int64
c_int = C.c_int
c_double = C.c_double
c_void_p = C.c_void_p
c_size_t = C.c_size_t
c_char_p = C.c_char_p

# name -> (restype, argtypes).  Pointers to device memory are passed as c_void_p.
SIGNATURES: dict[str, tuple] = {
    "ebg_ctx_create": (c_int, [C.POINTER(c_void_p), c_int]),
    "ebg_ctx_destroy": (c_int, [c_void_p]),
    "ebg_ctx_set_stream": (c_int, [c_void_p, c_void_p]),
    "ebg_ctx_synchronize": (c_int, [c_void_p]),
    "ebg_last_error": (c_char_p, []),
    "ebg_device_alloc": (c_int, [C.POINTER(c_void_p), c_size_t]),
    "ebg_device_free": (c_int, [c_void_p]),
    "ebg_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ebg_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ebg_grid2d_set": (c_int, [c_void_p, c_int64, c_int64] + [c_void_p] * 8 + [c_double, c_double]),
    "ebg_marker_inside_flags": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "ebg_marker_cell_indices": (c_int, [c_void_p, c_int64] + [c_void_p] * 5),
    "ebg_rk4_advect": (c_int, [c_void_p, c_int64, c_int, c_double] + [c_void_p] * 13 + [c_int]),
    "ebg_rk4_set_arithmetic": (c_int, [c_void_p, c_int]),
    "ebg_rk4_set_tiling": (c_int, [c_void_p, c_int]),
    "ebg_scatter_set_order": (c_int, [c_void_p, c_int]),
    "ebg_stokes_set_graphs": (c_int, [c_void_p, c_int]),
    "ebg_rk4_tolerance_kernel_available": (c_int, [c_void_p]),
    "ebg_calculate_spin": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_interpolate_marker_stokes_parameters_to_grid": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_stokes_viscoelastic_rhs": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_update_nodes_for_plastic_yielding": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_update_nodes_for_plastic_yielding_noalloc": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_update_marker_rheology": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_update_marker_plastic_yielding": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_calc_marker_viscoplastic_viscosity": (c_int, [c_void_p, c_int64] + [c_void_p] * 10),
    "ebg_update_marker_strain": (c_int, [c_void_p, c_int64, c_void_p, c_double, c_double] + [c_void_p] * 10),
    "ebg_grid_strain_rate_and_stress": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_harmonic_normal_viscosity": (c_int, [c_void_p] * 6),
    "ebg_marker_strain_rate_and_pressure": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_plastic_strain_rate_at_pressure_nodes": (c_int, [c_void_p] * 3),
    "ebg_stokes_set_system": (c_int, [c_void_p] + [c_void_p] * 16 + [c_int, c_double, c_double,
                                                                     c_double, c_double, c_int]),
    "ebg_stokes_apply": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_stokes_rhs": (c_int, [c_void_p, c_void_p]),
    "ebg_stokes_residual": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_stokes_unpack": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_stokes_solve_opts_default": (c_int, [c_void_p]),
    "ebg_compute_stokes_residuals": (c_int, [c_void_p] * 7),
    "ebg_change_norms": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_calculate_new_timestep": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_double, c_void_p,
                                           c_void_p]),
    "ebg_stokes_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_mg2d_create": (c_int, [c_void_p, c_int64, c_int64, c_double, c_double, c_int, c_void_p]),
    "ebg_mg2d_level_shape": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "ebg_mg2d_copy_field": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "ebg_mg2d_restrict_viscosity": (c_int, [c_void_p]),
    "ebg_mg2d_smoother": (c_int, [c_void_p, c_int, c_int, c_double, c_double, c_double, c_void_p]),
    "ebg_stokes_precondition": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_stokes_solve_rhs": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_cnn_create": (c_int, [c_void_p, C.c_uint64]),
    "ebg_cnn_num_params": (c_int, [c_void_p, c_void_p]),
    "ebg_cnn_get_params": (c_int, [c_void_p, c_void_p]),
    "ebg_cnn_set_params": (c_int, [c_void_p, c_void_p]),
    "ebg_cnn_forward": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "ebg_cnn_loss": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "ebg_cnn_loss_and_grad": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                      c_void_p]),
    "ebg_cnn_adamw_step": (c_int, [c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]),
    "ebg_cnn_make_sample": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_mg2d_smoother_tiled": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_double, c_double,
                                        c_double, c_void_p]),
    "ebg_mg2d_vcycles": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_double, c_double,
                                 c_double, c_double, c_double, c_void_p]),
    "ebg_mg3d_create": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_double, c_double, c_double,
                                c_int, c_void_p]),
    "ebg_mg3d_level_shape": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "ebg_mg3d_set_fused_limit": (c_int, [c_void_p, c_int64]),
    "ebg_mg3d_copy_field": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "ebg_mg3d_restrict_viscosity": (c_int, [c_void_p]),
    "ebg_mg3d_set_bc": (c_int, [c_void_p, c_int]),
    "ebg_mg3d_smoother": (c_int, [c_void_p, c_int, c_int, c_double, c_double, c_double]),
    "ebg_mg3d_restrict": (c_int, [c_void_p, c_int]),
    "ebg_mg3d_prolong": (c_int, [c_void_p, c_int, c_double, c_double]),
    "ebg_mg3d_vcycles": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_double, c_double,
                                 c_double, c_double, c_double]),
    "ebg_mg3d_run": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_mg3d_run_multimulti": (c_int, [c_void_p] * 7 + [c_int]),
    "ebg_mg2d_run": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_mg2d_run_multimulti": (c_int, [c_void_p] * 8 + [c_int]),
    "ebg_mg2d_direct_solve": (c_int, [c_void_p, c_int, c_void_p]),
    "ebg_stokes3d_set_system": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_double, c_double,
                                        c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_stokes3d_vector_length": (c_int, [c_void_p, c_void_p]),
    "ebg_stokes3d_apply": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_stokes3d_precondition": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_stokes3d_solve": (c_int, [c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p]),
    "ebg_interpolate_marker_heat_parameters_to_grid": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_apply_thermal_bc": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_heat_rhs": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ebg_heat_system_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_heat_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_heat_residuals": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_run_conductive_heat_solver_loop": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ebg_grid_to_marker_temperature": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "ebg_apply_sticky_temperature_correction": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_double,
                                                        c_void_p]),
    "ebg_subgrid_heat_diffusion_update": (c_int, [c_void_p] * 6),
    "ebg_update_subgrid_stress": (c_int, [c_void_p, c_void_p, c_void_p]),
}


class ViscoelasticIn(C.Structure):
    _fields_ = [("etas1", c_void_p), ("mus1", c_void_p), ("sxy1", c_void_p), ("rho1", c_void_p),
                ("etan1", c_void_p), ("mun1", c_void_p), ("sxx1", c_void_p), ("rho1_vy", c_void_p),
                ("dilatation_grid", c_void_p), ("eii_plastic_pressure", c_void_p),
                ("timestep", c_double), ("gravity_x", c_double), ("gravity_y", c_double),
                ("iuse_interface_stabilization", c_int)]


class ViscoelasticOut(C.Structure):
    _fields_ = [("etas0", c_void_p), ("sxy0", c_void_p), ("etan0", c_void_p), ("sxx0", c_void_p),
                ("RX1", c_void_p), ("RY1", c_void_p), ("RC1", c_void_p)]


class StrainStressIn(C.Structure):
    _fields_ = [("vx1", c_void_p), ("vy1", c_void_p), ("etas1", c_void_p), ("mus1", c_void_p),
                ("sxy1", c_void_p), ("etan1", c_void_p), ("mun1", c_void_p), ("sxx1", c_void_p),
                ("timestep", c_double)]


class StrainStressOut(C.Structure):
    _fields_ = [("exx", c_void_p), ("exy", c_void_p), ("eii", c_void_p), ("sxy2", c_void_p),
                ("sxx2", c_void_p), ("dsxy", c_void_p), ("dsxx", c_void_p)]


class MarkerGatherIn(C.Structure):
    _fields_ = ([("marknum", c_int64)] +
                [(k, c_void_p) for k in ("x", "y", "marker_xn", "marker_yn", "inside_flags", "exx", "exy",
                                         "pr1", "sxx2", "sxy2", "dsxx", "dsxy", "eta", "matid", "mat_mu")] +
                [("nmat", c_int64), ("timestep", c_double)])


class MarkerGatherOut(C.Structure):
    _fields_ = [("marker_exx", c_void_p), ("marker_exy", c_void_p), ("marker_pr", c_void_p),
                ("marker_sr_ratio", c_void_p), ("marker_sxx", c_void_p), ("marker_sxy", c_void_p)]


class PlasticNodesIn(C.Structure):
    _fields_ = [("plastic_yield", c_void_p), ("etas1", c_void_p), ("eta_flow", c_void_p), ("mus1", c_void_p),
                ("cohesion_grid", c_void_p), ("fric_degrees_grid", c_void_p), ("sxy2", c_void_p),
                ("sxx2", c_void_p), ("pr1", c_void_p), ("timestep", c_double), ("timesum", c_double),
                ("ntimestep", c_int64), ("viscosity_min", c_double), ("viscosity_max", c_double),
                ("iuse_fluid_pressure_for_yield", c_int), ("y_sealevel", c_double),
                ("no_yielding_in_mobile_wall", c_int), ("no_yielding_in_plate_extension", c_int),
                ("velocity_internal_x", c_double), ("x_left_mobile_wall", c_double),
                ("x_right_mobile_wall", c_double), ("y_top_mobile_wall", c_double),
                ("y_bottom_mobile_wall", c_double), ("plate_extension_thickness", c_double),
                ("plate_extension_width", c_double)]


class PlasticNodesOut(C.Structure):
    _fields_ = [("etavp_new", c_void_p), ("plastic_strain_rate", c_void_p), ("plastic_yield_new", c_void_p),
                ("yield_error", c_void_p), ("nnode_yield", c_int64), ("global_yield_error", c_double)]


class MarkerRheologyIn(C.Structure):
    """ebg_marker_rheology_in (include/earthbox_gpu.h)."""
    _fields_ = [("marknum", c_int64), ("inside_flags", c_void_p), ("matid", c_void_p),
                ("strain_plastic", c_void_p), ("fric_ini", c_void_p), ("melt_damage", c_void_p),
                ("sxx", c_void_p), ("sxy", c_void_p), ("exx", c_void_p), ("exy", c_void_p), ("TK", c_void_p),
                ("pr", c_void_p), ("GII", c_void_p), ("sr_ratio", c_void_p), ("nmat", c_int64),
                ("mat_flow_type", c_void_p), ("mat_flow", c_void_p), ("mat_plastic", c_void_p),
                ("mat_mu", c_void_p), ("timestep", c_double), ("temperature_top", c_double),
                ("viscosity_min", c_double), ("viscosity_max", c_double), ("powerlaw_stress_min", c_double),
                ("gas_constant", c_double), ("iuse_viscous_strain_soft", c_int), ("vsoftfac", c_double),
                ("itype_global", c_int), ("ntimestep", c_int64), ("stages", c_int)]


class MarkerRheologyOut(C.Structure):
    _fields_ = [("cohesion", c_void_p), ("fric", c_void_p), ("eta_flow", c_void_p), ("preexp", c_void_p),
                ("eta", c_void_p)]


class MarkerPlasticIn(C.Structure):
    """ebg_marker_plastic_in."""
    _fields_ = [("marknum", c_int64), ("inside_flags", c_void_p), ("y", c_void_p), ("matid", c_void_p),
                ("mat_mu", c_void_p), ("nmat", c_int64), ("cohesion", c_void_p), ("fric", c_void_p),
                ("exx", c_void_p), ("exy", c_void_p), ("sr_ratio", c_void_p), ("pr", c_void_p),
                ("forecast", C.c_int32), ("ntimestep", C.c_int32), ("timestep", c_double), ("y_sealevel", c_double),
                ("viscosity_min", c_double), ("viscosity_max", c_double), ("yield_stress_min", c_double),
                ("yield_stress_max", c_double), ("iuse_fluid_pressure_for_yield", C.c_int32),
                ("no_yielding_in_mobile_wall", C.c_int32), ("matid_sticky_air", C.c_int32),
                ("matid_mobile_wall", C.c_int32), ("matid_plate_extension", C.c_int32)]


class MarkerPlasticOut(C.Structure):
    _fields_ = [("sxx", c_void_p), ("sxy", c_void_p), ("eta", c_void_p), ("pfailure", c_void_p)]


class MarkerHeatProps(C.Structure):
    """ebg_marker_heat_props (include/earthbox_gpu.h)."""
    _fields_ = [("marknum", c_int64), ("x", c_void_p), ("y", c_void_p), ("inside_flags", c_void_p),
                ("TK", c_void_p), ("kt", c_void_p), ("rhocp", c_void_p), ("ha", c_void_p), ("matid", c_void_p),
                ("mat_hr", c_void_p), ("nmat", c_int64), ("serpentinization_heat_production", c_void_p)]


class HeatGridFields(C.Structure):
    """ebg_heat_grid_fields."""
    _fields_ = [("tk1", c_void_p), ("kt1", c_void_p), ("rhocp1", c_void_p), ("hr1", c_void_p), ("ha1", c_void_p),
                ("wtnodes", c_void_p), ("marker_xn", c_void_p), ("marker_yn", c_void_p)]


class ThermalBC(C.Structure):
    """ebg_thermal_bc."""
    _fields_ = [("btopt", c_void_p), ("bbottomt", c_void_p), ("bleftt", c_void_p), ("brightt", c_void_p)]


class HeatRhsIn(C.Structure):
    """ebg_heat_rhs_in."""
    _fields_ = [(k, c_void_p) for k in ("hr1", "ha1", "rho1", "vx1", "vy1", "sxy2", "etas1", "sxx2", "etan1")] + [
        ("gravity_x", c_double), ("gravity_y", c_double), ("iuse_adiabatic_heating", C.c_int32),
        ("iuse_shear_heating", C.c_int32)]


class HeatSystem(C.Structure):
    """ebg_heat_system."""
    _fields_ = [("RT1", c_void_p), ("tk0", c_void_p), ("rhocp1", c_void_p), ("kt1", c_void_p), ("bc", ThermalBC),
                ("timestep", c_double)]


class HeatSolveOpts(C.Structure):
    """ebg_heat_solve_opts (zeros select the defaults)."""
    _fields_ = [("restart", C.c_int32), ("max_iterations", C.c_int32), ("rtol", c_double),
                ("pre_sweeps", C.c_int32), ("post_sweeps", C.c_int32)]


class HeatLoopIn(C.Structure):
    """ebg_heat_loop_in."""
    _fields_ = [("RT1", c_void_p), ("rhocp1", c_void_p), ("kt1", c_void_p), ("bc", ThermalBC),
                ("timestep", c_double), ("max_temp_change", c_double), ("max_substeps", C.c_int32),
                ("tk1", c_void_p)]


class HeatLoopOut(C.Structure):
    """ebg_heat_loop_out."""
    _fields_ = [("tk0", c_void_p), ("tk2", c_void_p), ("rest", c_void_p), ("dtk1", c_void_p)]


class HeatLoopStats(C.Structure):
    """ebg_heat_loop_stats."""
    _fields_ = [("substeps", C.c_int32), ("solves", C.c_int32), ("iterations", C.c_int32),
                ("not_converged", C.c_int32), ("last_relres", c_double), ("timestep_last", c_double),
                ("timestep_sum", c_double)]


class MarkerCells(C.Structure):
    """ebg_marker_cells."""
    _fields_ = [("marknum", c_int64), ("x", c_void_p), ("y", c_void_p), ("inside_flags", c_void_p),
                ("marker_xn", c_void_p), ("marker_yn", c_void_p)]


class SubgridHeatIn(C.Structure):
    """ebg_subgrid_heat_in."""
    _fields_ = [("markers", MarkerCells), ("tk1", c_void_p), ("kt1", c_void_p), ("rhocp1", c_void_p),
                ("timestep", c_double), ("subgrid_diff_coef_temp", c_double)]


class SubgridStressIn(C.Structure):
    """ebg_subgrid_stress_in."""
    _fields_ = [("markers", MarkerCells), ("matid", c_void_p), ("mat_mu", c_void_p), ("nmat", c_int64),
                ("eta", c_void_p), ("sxy1", c_void_p), ("sxx1", c_void_p), ("timestep", c_double),
                ("subgrid_diff_coef_stress", c_double)]


class SubgridStressOut(C.Structure):
    """ebg_subgrid_stress_out."""
    _fields_ = [(k, c_void_p) for k in ("marker_sxy", "marker_sxx", "dsxy", "dsxx", "dsxyn", "dsxxn", "wtetas",
                                        "wtetan")]


class MarkerStokesProps(C.Structure):
    """ebg_marker_stokes_props (include/earthbox_gpu.h)."""
    _fields_ = [("marknum", c_int64), ("x", c_void_p), ("y", c_void_p), ("inside_flags", c_void_p),
                ("rho", c_void_p), ("cohesion", c_void_p), ("fric", c_void_p), ("eta_flow", c_void_p),
                ("eta", c_void_p), ("sxy", c_void_p), ("sxx", c_void_p), ("meltfrac", c_void_p),
                ("pfailure", c_void_p), ("dilatation", c_void_p), ("matid", c_void_p),
                ("mat_mu", c_void_p), ("nmat", c_int64)]


GRID_FIELD_NAMES = ("rho1", "rho1_vy", "etas1", "etan1", "eta_flow", "mus1", "mun1", "sxy1", "sxx1",
                    "plastics", "plasticn", "cohesion_grid", "fric_degrees_grid", "dilatation_grid",
                    "extractable_meltfrac_grid", "wtnodes", "wtetas", "wtetan", "wtnodes_vy",
                    "marker_xn", "marker_yn")


class StokesGridFields(C.Structure):
    """ebg_stokes_grid_fields (include/earthbox_gpu.h)."""
    _fields_ = [(k, c_void_p) for k in GRID_FIELD_NAMES]


class SolveOpts(C.Structure):
    """ebg_solve_opts (include/earthbox_gpu.h)."""
    _fields_ = [("precision", c_int), ("restart", c_int), ("max_iterations", c_int),
                ("rtol", c_double), ("nlevels", c_int), ("finest_sweeps", c_int),
                ("sweep_multiplier", c_int), ("sweep_cap", c_int), ("theta_s", c_double),
                ("theta_c", c_double), ("use_initial_guess", c_int), ("learned", c_int),
                ("bc_d", c_double * 4), ("smoother", c_int), ("local_sweeps", c_int),
                ("tile", c_int), ("flexible", c_int), ("coarse_unknowns", c_int),
                ("additive_tiles", c_int), ("learned_patience", c_int)]


class MGRunParams(C.Structure):
    """ebg_mg_run_params (include/earthbox_gpu.h)."""
    _fields_ = [("nvcycles", c_int), ("nvcycles_viscosity_jump", c_int), ("finest_sweeps", c_int),
                ("sweep_multiplier", c_int), ("sweep_cap", c_int), ("nviscosity_jumps", c_double),
                ("convergence_criterion", c_double), ("resmax", c_double), ("stokesscale", c_double),
                ("continscale", c_double), ("theta_s", c_double), ("theta_c", c_double),
                ("relax_v", c_double), ("relax_p", c_double), ("pressure_bc", c_double)]


class SolveStats(C.Structure):
    """ebg_solve_stats (include/earthbox_gpu.h)."""
    _fields_ = [("iterations", c_int), ("restarts", c_int), ("relres", c_double),
                ("relres_initial", c_double), ("seconds", c_double), ("status", c_int),
                ("learned_accepted", c_int), ("learned_evaluated", c_int)]


class EBGError(RuntimeError):
    """A hard error reported by the library (negative return code)."""


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() "
            "(or python earthbox.jl_amd/build_lib.py) first")
    # torch ships its own HIP runtime.  Loading it first lets the library bind
    # to that same runtime (same soname); loading the library first would map
    # a second HIP/HSA runtime into the process, and the one initialised
    # second sees no device.
    import torch  # noqa: F401
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if the export is missing
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc: int, what: str) -> int:
    """Raise EBGError for rc < 0; return rc otherwise (0 ok, >0 not converged)."""
    if rc < 0:
        msg = lib.ebg_last_error().decode(errors="replace")
        raise EBGError(f"{what} failed (rc={rc}): {msg}")
    return rc
