"""
    EarthBoxGPU

Thin `ccall` binding of `libearthbox_gpu.so` (C ABI: include/earthbox_gpu.h) for
EarthBox.jl.  It keeps the reference's entry points as the seams:

- `solve_system_gpu!(model, S)`        beside SystemSolver.solve_system_optimized
                                        (solver/system_solver/SystemSolver.jl:67-91)
- `interpolate_markers_gpu!(model, inside_flags)` beside Interpolation.initialize_bilinear_interpolation!
                                        + interpolate_marker_stokes_parameters_to_grid!
                                        (interpolation/Interpolation.jl:61-95)
- `advect_markers_gpu!(model)`         beside Advection.advect_markers_and_rotate_stress_tensor!
                                        (advection/Advection.jl:109-114)
- `run_multigrid_solver_gpu!(multigrid_data)` beside MultigridSolver.run_multigrid_solver
                                        (multigrid/multigrid_solver/MultigridSolver.jl:113-135, 3-D)
- `interpolate_marker_heat_parameters_to_grid_gpu!(model, inside_flags)` beside
                                        Interpolation.interpolate_marker_heat_parameters_to_grid!
                                        (interpolation/Interpolation.jl:77-85)
- `update_marker_temp_for_grid_and_subgrid_changes_gpu!(model, inside_flags)` beside
                                        HeatSolver.update_marker_temp_for_grid_and_subgrid_changes!
                                        (heat_solver/HeatSolver.jl:77-101)
- `apply_sticky_temperature_correction_gpu!(model, inside_flags, air, water)` beside
                                        HeatSolver.apply_sticky_temperature_correction!
- `update_subgrid_stress_gpu!(model, inside_flags)` beside SubGridStress.Update.update_subgrid_stress!
                                        (subgrid_stress/update/Update.jl:55-71)
- `update_plastic_yielding_gpu!(model, inside_flags, no_yielding_in_mobile_wall)` beside
                                        MarkerPlasticity.update_plastic_yielding!
                                        (rheology/plastic_failure/markers/MarkerPlasticity.jl:34-43)
- `interpolate_viscosity_to_transport_arrays_gpu!(model, inside_flags)` beside
                                        backup_and_clear_viscosity_transport_arrays! +
                                        MarkersToStokesArrays.interpolate_viscosity_to_transport_arrays!
                                        (PlasticityLoopMarkers.jl:50-52)

Every call copies the needed ModelData arrays to device buffers owned by a
`GPUState`, runs the kernels and copies results back, so ModelData stays the
authoritative state (checkpoints need nothing from the device).
"""
module EarthBoxGPU

const LIB = get(ENV, "EARTHBOX_GPU_LIB",
                joinpath(@__DIR__, "..", "earthbox_gpu", "libearthbox_gpu.so"))

struct SolveOpts
    precision::Cint
    restart::Cint
    max_iterations::Cint
    rtol::Cdouble
    nlevels::Cint
    finest_sweeps::Cint
    sweep_multiplier::Cint
    sweep_cap::Cint
    theta_s::Cdouble
    theta_c::Cdouble
    use_initial_guess::Cint
    learned::Cint
    bc_d::NTuple{4,Cdouble}
    smoother::Cint
    local_sweeps::Cint
    tile::Cint
    flexible::Cint
    coarse_unknowns::Cint
    additive_tiles::Cint
    learned_patience::Cint
end

struct SolveStats
    iterations::Cint
    restarts::Cint
    relres::Cdouble
    relres_initial::Cdouble
    seconds::Cdouble
    status::Cint
    learned_accepted::Cint
    learned_evaluated::Cint
end

"Field order of ebg_mg_run_params (include/earthbox_gpu.h)."
struct MGRunParams
    nvcycles::Cint
    nvcycles_viscosity_jump::Cint
    finest_sweeps::Cint
    sweep_multiplier::Cint
    sweep_cap::Cint
    nviscosity_jumps::Cdouble
    convergence_criterion::Cdouble
    resmax::Cdouble
    stokesscale::Cdouble
    continscale::Cdouble
    theta_s::Cdouble
    theta_c::Cdouble
    relax_v::Cdouble
    relax_p::Cdouble
    pressure_bc::Cdouble
end

check(rc, what) = rc < 0 ? error("$what failed: ",
    unsafe_string(ccall((:ebg_last_error, LIB), Cstring, ()))) : rc

mutable struct GPUState
    ctx::Ptr{Cvoid}
    buffers::Dict{Symbol,Ptr{Cvoid}}
end

function GPUState(device::Integer=0)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:ebg_ctx_create, LIB), Cint, (Ptr{Ptr{Cvoid}}, Cint), ctx, device), "ebg_ctx_create")
    st = GPUState(ctx[], Dict{Symbol,Ptr{Cvoid}}())
    finalizer(st) do s
        for p in values(s.buffers)
            ccall((:ebg_device_free, LIB), Cint, (Ptr{Cvoid},), p)
        end
        ccall((:ebg_ctx_destroy, LIB), Cint, (Ptr{Cvoid},), s.ctx)
    end
    return st
end

"Device mirror of a host array (reallocated when the size changes)."
function dev!(st::GPUState, key::Symbol, a::AbstractArray)
    nbytes = sizeof(a)
    p = get(st.buffers, key, C_NULL)
    if p == C_NULL || get(st.buffers, Symbol(key, :_n), C_NULL) != Ptr{Cvoid}(nbytes)
        p != C_NULL && ccall((:ebg_device_free, LIB), Cint, (Ptr{Cvoid},), p)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:ebg_device_alloc, LIB), Cint, (Ptr{Ptr{Cvoid}}, Csize_t), r, nbytes), "alloc")
        p = r[]
        st.buffers[key] = p
        st.buffers[Symbol(key, :_n)] = Ptr{Cvoid}(nbytes)
    end
    return p
end
h2d!(st, key, a) = (p = dev!(st, key, a);
                    GC.@preserve a check(ccall((:ebg_memcpy_h2d, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), p, pointer(a), sizeof(a)), "h2d"); p)
d2h!(a, p) = GC.@preserve a check(ccall((:ebg_memcpy_d2h, LIB), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), pointer(a), p, sizeof(a)), "d2h")

function set_grid!(st::GPUState, model)
    g = model.grids.arrays
    geo = model.grids.parameters.geometry
    check(ccall((:ebg_grid2d_set, LIB), Cint,
        (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble),
        st.ctx, geo.xnum.value, geo.ynum.value,
        g.basic.gridx_b.array, g.basic.gridy_b.array, g.basic.xstp_b.array, g.basic.ystp_b.array,
        g.staggered_vy.gridx_vy.array, g.staggered_vy.xstp_vy.array,
        g.staggered_vx.gridy_vx.array, g.staggered_vx.ystp_vx.array,
        geo.xstpavg.value, geo.ystpavg.value), "ebg_grid2d_set")
end

"""
    solve_system_gpu!(st, model; mixed=true, rtol=1e-8) -> (S, stats)

GPU replacement of SystemSolver.solve_system_optimized: solves the assembled
system without building the COO, returning S in the reference's unknown order.
"""
function solve_system_gpu!(st::GPUState, model; mixed::Bool=true, rtol::Float64=1e-8,
                           initial_guess::Union{Nothing, Tuple{Vector{Float64}, Float64}}=nothing)
    set_grid!(st, model)
    sc = model.stokes_continuity
    bcs = model.bcs.arrays
    p = Dict(k => h2d!(st, k, a) for (k, a) in (
        :etas0 => sc.arrays.viscosity.etas0.array, :etan0 => sc.arrays.viscosity.etan0.array,
        :rho1_vy => sc.arrays.density.rho1_vy.array, :RX1 => sc.arrays.rhs.RX1.array,
        :RY1 => sc.arrays.rhs.RY1.array, :RC1 => sc.arrays.rhs.RC1.array))
    vc = bcs.vel_comp
    check(ccall((:ebg_stokes_set_system, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Ptr{Float64}, Cint, Cdouble, Cdouble, Cdouble,
         Cdouble, Cint),
        st.ctx, p[:etas0], p[:etan0], p[:rho1_vy], p[:RX1], p[:RY1], p[:RC1],
        vc.btopx.array, vc.btopy.array, vc.bbottomx.array, vc.bbottomy.array,
        vc.bleftx.array, vc.blefty.array, vc.brightx.array, vc.brighty.array,
        bcs.internal.bintern_zone.array, bcs.internal.bintern_velocity.array,
        model.bcs.parameters.bc_options.pressure_bc_mode.value,
        model.bcs.parameters.pressure.pressure_bc.value, sc.parameters.build.pscale.value,
        model.timestep.parameters.main_time_loop.timestep.value,
        model.gravity.parameters.gravity_y.value,
        sc.parameters.build.iuse_interface_stabilization.value), "ebg_stokes_set_system")
    N = sc.parameters.build.N.value
    # initial_guess = (S, pscale) of an earlier solve (the previous Picard
    # iterate or time step): a warm start; its scaled pressures (every third
    # unknown, p / pscale) are rescaled to this system's pscale
    pscale = sc.parameters.build.pscale.value
    S = initial_guess === nothing ? zeros(Float64, N) : copy(initial_guess[1])
    if initial_guess !== nothing && initial_guess[2] != pscale
        S[3:3:end] .*= initial_guess[2] / pscale
    end
    dS = h2d!(st, :S, S)
    o = Ref{SolveOpts}()
    ccall((:ebg_stokes_solve_opts_default, LIB), Cint, (Ref{SolveOpts},), o)
    # restart 100: the basis holds the plastic systems without a restart
    o[] = SolveOpts(mixed ? 1 : 0, 100, o[].max_iterations, rtol, o[].nlevels,
                    o[].finest_sweeps, o[].sweep_multiplier, o[].sweep_cap, o[].theta_s,
                    o[].theta_c, initial_guess === nothing ? 0 : 1, 0,
                    (vc.btopx.array[2, 2], vc.bbottomx.array[2, 2], vc.blefty.array[2, 2],
                     vc.brighty.array[2, 2]), o[].smoother, o[].local_sweeps, o[].tile,
                    o[].flexible, o[].coarse_unknowns, o[].additive_tiles, o[].learned_patience)
    stats = Ref{SolveStats}()
    rc = check(ccall((:ebg_stokes_solve, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ref{SolveOpts}, Ref{SolveStats}), st.ctx, dS, o, stats),
        "ebg_stokes_solve")
    d2h!(S, dS)
    # rc > 0: not converged; the caller falls back to the CPU path, as the
    # reference does after MUMPS retries (ParallelSolver.jl:75-124)
    return S, stats[], rc
end

"Field order of ebg_viscoelastic_in / ebg_viscoelastic_out (include/earthbox_gpu.h)."
struct ViscoelasticIn
    etas1::Ptr{Cvoid}; mus1::Ptr{Cvoid}; sxy1::Ptr{Cvoid}; rho1::Ptr{Cvoid}
    etan1::Ptr{Cvoid}; mun1::Ptr{Cvoid}; sxx1::Ptr{Cvoid}
    rho1_vy::Ptr{Cvoid}
    dilatation_grid::Ptr{Cvoid}; eii_plastic_pressure::Ptr{Cvoid}
    timestep::Cdouble; gravity_x::Cdouble; gravity_y::Cdouble
    iuse_interface_stabilization::Cint
end
struct ViscoelasticOut
    etas0::Ptr{Cvoid}; sxy0::Ptr{Cvoid}; etan0::Ptr{Cvoid}; sxx0::Ptr{Cvoid}
    RX1::Ptr{Cvoid}; RY1::Ptr{Cvoid}; RC1::Ptr{Cvoid}
end

"""
    calc_viscoelastic_terms_and_rhs_gpu!(st, model)

GPU replacement of the back-to-back calls
StokesViscoelasticTerms.calc_viscoelastic_terms!(model) and
StokesContinuityRhs.viscoelastic_rhs!(model) (StokesContinuitySolver.jl:189-190):
fills etas0, sxy0, etan0, sxx0, RX1, RY1 and RC1 of model.stokes_continuity.
"""
function calc_viscoelastic_terms_and_rhs_gpu!(st::GPUState, model)
    set_grid!(st, model)
    a = model.stokes_continuity.arrays
    up(k, x) = h2d!(st, k, x)
    vin = ViscoelasticIn(up(:etas1, a.viscosity.etas1.array), up(:mus1, a.shear_modulus.mus1.array),
        up(:sxy1, a.stress.sxy1.array), up(:rho1, a.density.rho1.array),
        up(:etan1, a.viscosity.etan1.array), up(:mun1, a.shear_modulus.mun1.array),
        up(:sxx1, a.stress.sxx1.array), up(:rho1_vy, a.density.rho1_vy.array),
        up(:dil, a.plastic_def.dilatation_grid.array),
        up(:eiip, a.strain_rate_and_spin.eii_plastic_pressure.array),
        model.timestep.parameters.main_time_loop.timestep.value,
        model.gravity.parameters.gravity_x.value, model.gravity.parameters.gravity_y.value,
        model.stokes_continuity.parameters.build.iuse_interface_stabilization.value)
    outs = (a.viscosity.etas0.array, a.stress.sxy0.array, a.viscosity.etan0.array, a.stress.sxx0.array,
            a.rhs.RX1.array, a.rhs.RY1.array, a.rhs.RC1.array)
    keys = (:etas0, :sxy0, :etan0, :sxx0, :RX1, :RY1, :RC1)
    vout = ViscoelasticOut((dev!(st, k, x) for (k, x) in zip(keys, outs))...)
    check(ccall((:ebg_stokes_viscoelastic_rhs, LIB), Cint,
        (Ptr{Cvoid}, Ref{ViscoelasticIn}, Ref{ViscoelasticOut}), st.ctx, vin, vout),
        "ebg_stokes_viscoelastic_rhs")
    check(ccall((:ebg_ctx_synchronize, LIB), Cint, (Ptr{Cvoid},), st.ctx), "sync")
    for (k, x) in zip(keys, outs)
        d2h!(x, st.buffers[k])
    end
    return nothing
end

"Field order of ebg_strain_stress_in / ebg_strain_stress_out (include/earthbox_gpu.h)."
struct StrainStressIn
    vx1::Ptr{Cvoid}; vy1::Ptr{Cvoid}
    etas1::Ptr{Cvoid}; mus1::Ptr{Cvoid}; sxy1::Ptr{Cvoid}
    etan1::Ptr{Cvoid}; mun1::Ptr{Cvoid}; sxx1::Ptr{Cvoid}
    timestep::Cdouble
end
struct StrainStressOut
    exx::Ptr{Cvoid}; exy::Ptr{Cvoid}; eii::Ptr{Cvoid}
    sxy2::Ptr{Cvoid}; sxx2::Ptr{Cvoid}; dsxy::Ptr{Cvoid}; dsxx::Ptr{Cvoid}
end

"""
    update_grid_strain_rate_and_stress_gpu!(st, model)

GPU replacement of the grid part of stokes_solution_processing!
(StokesContinuitySolver.jl:221, 273-274): GridStrainRate's exx, exy and eii,
GridStress.forecast_viscoelastic_grid_stress! (sxy2, sxx2) and the stress
change (dsxy, dsxx).  Spin: ebg_calculate_spin.
"""
function update_grid_strain_rate_and_stress_gpu!(st::GPUState, model)
    set_grid!(st, model)
    a = model.stokes_continuity.arrays
    up(k, x) = h2d!(st, k, x)
    sin_ = StrainStressIn(up(:vx1, a.staggered_grid_velocity.vx1.array), up(:vy1, a.staggered_grid_velocity.vy1.array),
        up(:etas1, a.viscosity.etas1.array), up(:mus1, a.shear_modulus.mus1.array), up(:sxy1, a.stress.sxy1.array),
        up(:etan1, a.viscosity.etan1.array), up(:mun1, a.shear_modulus.mun1.array), up(:sxx1, a.stress.sxx1.array),
        model.timestep.parameters.main_time_loop.timestep.value)
    sr = a.strain_rate_and_spin
    outs = (sr.exx.array, sr.exy.array, sr.eii.array, a.stress.sxy2.array, a.stress.sxx2.array,
            a.stress_change.dsxy.array, a.stress_change.dsxx.array)
    keys = (:exx, :exy, :eii, :sxy2, :sxx2, :dsxy, :dsxx)
    sout = StrainStressOut((dev!(st, k, x) for (k, x) in zip(keys, outs))...)
    check(ccall((:ebg_grid_strain_rate_and_stress, LIB), Cint,
        (Ptr{Cvoid}, Ref{StrainStressIn}, Ref{StrainStressOut}), st.ctx, sin_, sout),
        "ebg_grid_strain_rate_and_stress")
    check(ccall((:ebg_ctx_synchronize, LIB), Cint, (Ptr{Cvoid},), st.ctx), "sync")
    for (k, x) in zip(keys, outs)
        d2h!(x, st.buffers[k])
    end
    return nothing
end

"Field order of ebg_plastic_nodes_in / ebg_plastic_nodes_out (include/earthbox_gpu.h)."
struct PlasticNodesIn
    plastic_yield::Ptr{Cvoid}; etas1::Ptr{Cvoid}; eta_flow::Ptr{Cvoid}; mus1::Ptr{Cvoid}
    cohesion_grid::Ptr{Cvoid}; fric_degrees_grid::Ptr{Cvoid}; sxy2::Ptr{Cvoid}
    sxx2::Ptr{Cvoid}; pr1::Ptr{Cvoid}
    timestep::Cdouble; timesum::Cdouble; ntimestep::Int64
    viscosity_min::Cdouble; viscosity_max::Cdouble
    iuse_fluid_pressure_for_yield::Cint; y_sealevel::Cdouble
    no_yielding_in_mobile_wall::Cint; no_yielding_in_plate_extension::Cint
    velocity_internal_x::Cdouble; x_left_mobile_wall::Cdouble; x_right_mobile_wall::Cdouble
    y_top_mobile_wall::Cdouble; y_bottom_mobile_wall::Cdouble
    plate_extension_thickness::Cdouble; plate_extension_width::Cdouble
end
mutable struct PlasticNodesOut
    etavp_new::Ptr{Cvoid}; plastic_strain_rate::Ptr{Cvoid}; plastic_yield_new::Ptr{Cvoid}
    yield_error::Ptr{Cvoid}; nnode_yield::Int64; global_yield_error::Cdouble
end

"""
    update_nodes_for_plastic_yielding_gpu!(st, model, no_yielding_in_mobile_wall,
        no_yielding_in_plate_extension=false) -> (etavp_new, plastic_strain_rate,
        plastic_yield_new, nnode_yield, global_yield_error)

GPU replacement of PlasticFailureNodes.update_nodes_for_plastic_yielding
(rheology/plastic_failure/nodes/PlasticFailureNodes.jl:24-319); returns the
reference's tuple, the three grids being the model's buffers, and updates
plastic_def.yield_error.
"""
function update_nodes_for_plastic_yielding_gpu!(st::GPUState, model, no_yielding_in_mobile_wall::Bool,
                                                no_yielding_in_plate_extension::Bool=false)
    set_grid!(st, model)
    a = model.stokes_continuity.arrays
    pd = a.plastic_def
    mw = model.geometry.parameters.mobile_wall
    up(k, x) = h2d!(st, k, x)
    pin = PlasticNodesIn(up(:pyield, pd.plastic_yield.array), up(:etas1, a.viscosity.etas1.array),
        up(:eta_flow, a.viscosity.eta_flow.array), up(:mus1, a.shear_modulus.mus1.array),
        up(:coh, pd.cohesion_grid.array), up(:fric, pd.fric_degrees_grid.array),
        up(:sxy2, a.stress.sxy2.array), up(:sxx2, a.stress.sxx2.array), up(:pr1, a.pressure.pr1.array),
        model.timestep.parameters.main_time_loop.timestep.value,
        model.timestep.parameters.main_time_loop.timesum.value,
        model.timestep.parameters.main_time_loop.ntimestep.value,
        model.materials.parameters.viscosity_limits.viscosity_min.value,
        model.materials.parameters.viscosity_limits.viscosity_max.value,
        model.materials.parameters.stress_limits_yield.iuse_fluid_pressure_for_yield.value,
        model.topography.parameters.sealevel.y_sealevel.value,
        no_yielding_in_mobile_wall, no_yielding_in_plate_extension,
        model.bcs.parameters.velocity.velocity_internal_x.value,
        mw.x_left_mobile_wall.value, mw.x_right_mobile_wall.value, mw.y_top_mobile_wall.value,
        mw.y_bottom_mobile_wall.value, mw.plate_extension_thickness.value, mw.plate_extension_width.value)
    outs = (pd.etavp_buffer.array, pd.plastic_strain_rate_buffer.array, pd.plastic_yield_buffer.array)
    keys = (:etavp_new, :psr, :pyield_new)
    pout = PlasticNodesOut(dev!(st, keys[1], outs[1]), dev!(st, keys[2], outs[2]), dev!(st, keys[3], outs[3]),
                           up(:yield_error, pd.yield_error.array), 0, 0.0)
    check(ccall((:ebg_update_nodes_for_plastic_yielding, LIB), Cint,
        (Ptr{Cvoid}, Ref{PlasticNodesIn}, Ref{PlasticNodesOut}), st.ctx, pin, pout),
        "ebg_update_nodes_for_plastic_yielding")
    for (k, x) in zip(keys, outs)
        d2h!(x, st.buffers[k])
    end
    d2h!(pd.yield_error.array, st.buffers[:yield_error])
    return (outs..., pout.nnode_yield, pout.global_yield_error)
end

"""
    advect_markers_gpu!(st, model, inside_flags)

GPU replacement of Advection.advect_markers_and_rotate_stress_tensor!.
"""
function advect_markers_gpu!(st::GPUState, model, inside_flags::Vector{Int8})
    set_grid!(st, model)
    m = model.markers.arrays
    sc = model.stokes_continuity.arrays
    n = model.markers.parameters.distribution.marknum.value
    dx = h2d!(st, :mx, m.location.marker_x.array)
    dy = h2d!(st, :my, m.location.marker_y.array)
    dxn = h2d!(st, :xn, m.grid_marker_relationship.marker_xn.array)
    dyn = h2d!(st, :yn, m.grid_marker_relationship.marker_yn.array)
    dfl = h2d!(st, :flags, inside_flags)
    dvx1 = h2d!(st, :vx1, sc.staggered_grid_velocity.vx1.array)
    dvy1 = h2d!(st, :vy1, sc.staggered_grid_velocity.vy1.array)
    desp = h2d!(st, :esp, sc.strain_rate_and_spin.esp.array)
    dmvx = dev!(st, :mvx, m.advection.marker_vx.array)
    dmvy = dev!(st, :mvy, m.advection.marker_vy.array)
    dmsp = dev!(st, :msp, m.advection.marker_spin.array)
    dsxx = h2d!(st, :sxx, m.stress.marker_sxx.array)
    dsxy = h2d!(st, :sxy, m.stress.marker_sxy.array)
    order = 4
    check(ccall((:ebg_rk4_advect, LIB), Cint,
        (Ptr{Cvoid}, Int64, Cint, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
         Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
         Ptr{Cvoid}, Ptr{Cvoid}, Cint),
        st.ctx, n, order, model.timestep.parameters.main_time_loop.timestep.value,
        dx, dy, dxn, dyn, dfl, dvx1, dvy1, desp, dmvx, dmvy, dmsp, dsxx, dsxy, 1), "ebg_rk4_advect")
    check(ccall((:ebg_ctx_synchronize, LIB), Cint, (Ptr{Cvoid},), st.ctx), "sync")
    d2h!(m.location.marker_x.array, dx)
    d2h!(m.location.marker_y.array, dy)
    d2h!(m.advection.marker_vx.array, dmvx)
    d2h!(m.advection.marker_vy.array, dmvy)
    d2h!(m.advection.marker_spin.array, dmsp)
    d2h!(m.stress.marker_sxx.array, dsxx)
    d2h!(m.stress.marker_sxy.array, dsxy)
    return nothing
end

"Field order of ebg_marker_stokes_props (include/earthbox_gpu.h)."
struct MarkerStokesProps
    marknum::Int64
    x::Ptr{Cvoid}; y::Ptr{Cvoid}; inside_flags::Ptr{Cvoid}
    rho::Ptr{Cvoid}; cohesion::Ptr{Cvoid}; fric::Ptr{Cvoid}; eta_flow::Ptr{Cvoid}; eta::Ptr{Cvoid}
    sxy::Ptr{Cvoid}; sxx::Ptr{Cvoid}; meltfrac::Ptr{Cvoid}
    pfailure::Ptr{Cvoid}; dilatation::Ptr{Cvoid}
    matid::Ptr{Cvoid}; mat_mu::Ptr{Cvoid}
    nmat::Int64
end

"Field order of ebg_stokes_grid_fields (include/earthbox_gpu.h)."
struct StokesGridFields
    rho1::Ptr{Cvoid}; rho1_vy::Ptr{Cvoid}; etas1::Ptr{Cvoid}; etan1::Ptr{Cvoid}; eta_flow::Ptr{Cvoid}
    mus1::Ptr{Cvoid}; mun1::Ptr{Cvoid}; sxy1::Ptr{Cvoid}; sxx1::Ptr{Cvoid}
    plastics::Ptr{Cvoid}; plasticn::Ptr{Cvoid}; cohesion_grid::Ptr{Cvoid}; fric_degrees_grid::Ptr{Cvoid}
    dilatation_grid::Ptr{Cvoid}; extractable_meltfrac_grid::Ptr{Cvoid}
    wtnodes::Ptr{Cvoid}; wtetas::Ptr{Cvoid}; wtetan::Ptr{Cvoid}; wtnodes_vy::Ptr{Cvoid}
    marker_xn::Ptr{Cvoid}; marker_yn::Ptr{Cvoid}
end

"""
    interpolate_markers_gpu!(st, model, inside_flags)

GPU replacement of Interpolation.initialize_bilinear_interpolation! followed by
Interpolation.interpolate_marker_stokes_parameters_to_grid!
(interpolation/Interpolation.jl:61-95), called from TimeLoop.jl:219,226.  It
reads the marker arrays BilinearNumerator.jl:205-300 reads, writes the Stokes
transport grids (rho1, rho1_vy, etas1, etan1, eta_flow, mus1, mun1, sxy1,
sxx1, plastics, plasticn, cohesion_grid, fric_degrees_grid, dilatation_grid,
extractable_meltfrac_grid), the weight sums (wtnodes, wtetas, wtetan,
wtnodes_vy) and marker_xn/yn.  The grids travel both ways: nodes without
marker weight keep their previous value (BackupTransportArrays.jl:33-104).
The per-marker distances and weights the reference stores
(marker_dx/dy, marker_wtfor*node) are not written: the GPU keeps them in
registers.  Node sums follow the serial marker order, so results equal the
reference's loops bit for bit (the friction angle's asin aside).
"""
function interpolate_markers_gpu!(st::GPUState, model, inside_flags::Vector{Int8})
    set_grid!(st, model)
    m = model.markers.arrays
    sc = model.stokes_continuity.arrays
    iw = model.interpolation.arrays.grid_weights
    n = model.markers.parameters.distribution.marknum.value
    mat_mu = model.materials.arrays.mat_mu.array
    up(k, x) = h2d!(st, k, x)
    props = MarkerStokesProps(n, up(:mx, m.location.marker_x.array), up(:my, m.location.marker_y.array),
        up(:flags, inside_flags),
        up(:mrho, m.material.marker_rho.array), up(:mcoh, m.rheology.marker_cohesion.array),
        up(:mfric, m.rheology.marker_fric.array), up(:metaf, m.rheology.marker_eta_flow.array),
        up(:meta, m.rheology.marker_eta.array), up(:msxy, m.stress.marker_sxy.array),
        up(:msxx, m.stress.marker_sxx.array), up(:mmelt, m.melt.marker_extractable_meltfrac.array),
        up(:mpf, m.rheology.marker_pfailure.array), up(:mdil, m.rheology.marker_dilatation_angle.array),
        up(:matid, m.material.marker_matid.array), up(:mat_mu, mat_mu), length(mat_mu))
    grids = (sc.density.rho1.array, sc.density.rho1_vy.array, sc.viscosity.etas1.array,
             sc.viscosity.etan1.array, sc.viscosity.eta_flow.array, sc.shear_modulus.mus1.array,
             sc.shear_modulus.mun1.array, sc.stress.sxy1.array, sc.stress.sxx1.array,
             sc.plastic_def.plastics.array, sc.plastic_def.plasticn.array, sc.plastic_def.cohesion_grid.array,
             sc.plastic_def.fric_degrees_grid.array, sc.plastic_def.dilatation_grid.array,
             sc.plastic_def.extractable_meltfrac_grid.array,
             iw.wtnodes.array, iw.wtetas.array, iw.wtetan.array, iw.wtnodes_vy.array,
             m.grid_marker_relationship.marker_xn.array, m.grid_marker_relationship.marker_yn.array)
    keys = (:rho1, :rho1_vy, :etas1, :etan1, :eta_flow, :mus1, :mun1, :sxy1, :sxx1, :plastics, :plasticn,
            :cohesion_grid, :fric_degrees_grid, :dilatation_grid, :meltfrac_grid,
            :wtnodes, :wtetas, :wtetan, :wtnodes_vy, :xn, :yn)
    # previous values are the backup where no marker weight lands
    ptrs = [h2d!(st, k, x) for (k, x) in zip(keys, grids)]
    f = StokesGridFields(ptrs...)
    check(ccall((:ebg_interpolate_marker_stokes_parameters_to_grid, LIB), Cint,
        (Ptr{Cvoid}, Ref{MarkerStokesProps}, Ref{StokesGridFields}), st.ctx, props, f),
        "ebg_interpolate_marker_stokes_parameters_to_grid")
    check(ccall((:ebg_ctx_synchronize, LIB), Cint, (Ptr{Cvoid},), st.ctx), "sync")
    for (k, x) in zip(keys, grids)
        d2h!(x, st.buffers[k])
    end
    return nothing
end

"Field order of ebg_marker_gather_in / ebg_marker_gather_out (include/earthbox_gpu.h)."
struct MarkerGatherIn
    marknum::Int64
    x::Ptr{Cvoid}; y::Ptr{Cvoid}; marker_xn::Ptr{Cvoid}; marker_yn::Ptr{Cvoid}; inside_flags::Ptr{Cvoid}
    exx::Ptr{Cvoid}; exy::Ptr{Cvoid}; pr1::Ptr{Cvoid}
    sxx2::Ptr{Cvoid}; sxy2::Ptr{Cvoid}; dsxx::Ptr{Cvoid}; dsxy::Ptr{Cvoid}
    eta::Ptr{Cvoid}; matid::Ptr{Cvoid}; mat_mu::Ptr{Cvoid}
    nmat::Int64
    timestep::Cdouble
end
struct MarkerGatherOut
    marker_exx::Ptr{Cvoid}; marker_exy::Ptr{Cvoid}; marker_pr::Ptr{Cvoid}; marker_sr_ratio::Ptr{Cvoid}
    marker_sxx::Ptr{Cvoid}; marker_sxy::Ptr{Cvoid}
end

"""
    update_markers_after_stokes_gpu!(st, model, inside_flags)

GPU replacement of the marker steps of execute_post_stokes_solver_steps!
(StokesContinuitySolver.jl:272-279): MarkerStrainRate.calculate_marker_strain_rate!,
MarkerStrainRateRatio.calculate_marker_strain_rate_ratio! and
MarkerPressure.calculate_marker_pressure!, in one pass.  The grid steps before
them (update_grid_stress!, calculate_grid_stress_change!) are
update_grid_strain_rate_and_stress_gpu!.  With correct_stress=true the same
pass also runs StressCorrection.apply_marker_stress_correction_for_remaining_grid_stress!,
after update_subgrid_stress_gpu! when subgrid_diff_coef_stress > 0: together
the whole of update_marker_stress_for_grid_and_subgrid_changes!.
"""
function update_markers_after_stokes_gpu!(st::GPUState, model, inside_flags::Vector{Int8};
                                          correct_stress::Bool=false)
    if correct_stress
        # update_marker_stress_for_grid_and_subgrid_changes! (StokesContinuitySolver.jl:365-381)
        # runs update_subgrid_stress! before the correction when the coefficient is positive
        coef = model.markers.parameters.subgrid_diffusion.subgrid_diff_coef_stress.value
        coef > 0 && update_subgrid_stress_gpu!(st, model, inside_flags)
    end
    set_grid!(st, model)
    m = model.markers.arrays
    sc = model.stokes_continuity.arrays
    n = model.markers.parameters.distribution.marknum.value
    mat_mu = model.materials.arrays.mat_mu.array
    up(k, x) = h2d!(st, k, x)
    gin = MarkerGatherIn(n, up(:mx, m.location.marker_x.array), up(:my, m.location.marker_y.array),
        up(:xn, m.grid_marker_relationship.marker_xn.array), up(:yn, m.grid_marker_relationship.marker_yn.array),
        up(:flags, inside_flags),
        up(:exx, sc.strain_rate_and_spin.exx.array), up(:exy, sc.strain_rate_and_spin.exy.array),
        up(:pr1, sc.pressure.pr1.array), up(:sxx2, sc.stress.sxx2.array), up(:sxy2, sc.stress.sxy2.array),
        up(:dsxx, sc.stress_change.dsxx.array), up(:dsxy, sc.stress_change.dsxy.array),
        up(:meta, m.rheology.marker_eta.array), up(:matid, m.material.marker_matid.array), up(:mat_mu, mat_mu),
        length(mat_mu), model.timestep.parameters.main_time_loop.timestep.value)
    outs = (m.strain.marker_exx.array, m.strain.marker_exy.array, m.pressure.marker_pr.array,
            m.strain.marker_sr_ratio.array)
    keys = (:mexx, :mexy, :mpr, :msr)
    if correct_stress
        outs = (outs..., m.stress.marker_sxx.array, m.stress.marker_sxy.array)
        keys = (keys..., :sxx, :sxy)
    end
    # outside markers keep their values, so the outputs travel both ways
    ptrs = [h2d!(st, k, x) for (k, x) in zip(keys, outs)]
    gout = MarkerGatherOut(ptrs..., ntuple(_ -> C_NULL, 6 - length(ptrs))...)
    check(ccall((:ebg_marker_strain_rate_and_pressure, LIB), Cint,
        (Ptr{Cvoid}, Ref{MarkerGatherIn}, Ref{MarkerGatherOut}), st.ctx, gin, gout),
        "ebg_marker_strain_rate_and_pressure")
    check(ccall((:ebg_ctx_synchronize, LIB), Cint, (Ptr{Cvoid},), st.ctx), "sync")
    for (k, x) in zip(keys, outs)
        d2h!(x, st.buffers[k])
    end
    return nothing
end

"""
    run_multigrid_solver_gpu!(st, multigrid_data) -> (cycles, converged)

GPU replacement of MultigridSolver.run_multigrid_solver for MultigridData3d
(multigrid_solver/MultigridSolver.jl:113-135).  Uploads level_vector[1]
(viscosity, right-hand sides, initial solution), runs the continuation and
V-cycles on the device, and writes back the solution, the scaled viscosity
and the principal mean residuals, as the reference leaves them.  With
use_multimulti the initial guess and right-hand sides come from level0, the
device runs the multimulti driver (ebg_mg3d_run_multimulti) and the global
solution and global mean residuals return to level0 and mean_residuals.
"""
function run_multigrid_solver_gpu!(st::GPUState, md)
    mm = md.vcycle.use_multimulti
    lv = md.level_vector[1]
    # with multimulti the initial guess and right-hand sides live on level 0
    # (MultigridDataManager.jl:376-421); the viscosity is level 1's
    src = mm ? md.level0 : lv
    geo = lv.grid.parameters.geometry
    nlev = Ref{Cint}(0)
    check(ccall((:ebg_mg3d_create, LIB), Cint,
        (Ptr{Cvoid}, Int64, Int64, Int64, Cdouble, Cdouble, Cdouble, Cint, Ref{Cint}),
        st.ctx, geo.xnum.value, geo.ynum.value, geo.znum.value, geo.xsize.value, geo.ysize.value,
        geo.zsize.value, length(md.level_vector), nlev), "ebg_mg3d_create")
    ids = (etan=0, etaxy=1, etaxz=2, etayz=3, vx=4, vy=5, vz=6, pr=7, RX=8, RY=9, RZ=10, RC=11)
    copy_field(id, dst, src) = check(ccall((:ebg_mg3d_copy_field, LIB), Cint,
        (Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}), st.ctx, 0, id, dst, src), "ebg_mg3d_copy_field")
    for (name, id) in pairs(ids)
        from = name in (:etan, :etaxy, :etaxz, :etayz) ? lv : src
        hasproperty(from, name) || continue
        copy_field(id, C_NULL, h2d!(st, Symbol(:mg3_, name), getproperty(from, name).array))
    end
    s = md.smoothing_iterations
    cap = length(s) > 1 ? maximum(s[2:end]) : 0
    mult = length(s) > 1 ? div(s[2], s[1]) : 2
    vs, rx = md.viscosity_scaling, md.relaxation
    p = Ref(MGRunParams(md.vcycle.nvcycles, vs.nvcycles_viscosity_jump, s[1], mult, cap,
                        vs.nviscosity_jumps, md.vcycle.convergence_criterion, vs.resmax,
                        md.residual_scaling_factors.stokesscale, md.residual_scaling_factors.continscale,
                        rx.relax_stokes, rx.relax_continuity, rx.relax_velocity, rx.relax_pressure,
                        md.pressure_bc))
    cycles, conv = Ref{Cint}(0), Ref{Cint}(0)
    hist = zeros(Float64, 4 * max(md.vcycle.nvcycles, 1))
    nglobal = Ref{Cint}(0)
    ghist = zeros(Float64, 4 * max(md.vcycle.nvcycles, 1))
    if mm
        check(ccall((:ebg_mg3d_run_multimulti, LIB), Cint,
            (Ptr{Cvoid}, Ref{MGRunParams}, Ref{Cint}, Ref{Cint}, Ptr{Float64}, Ref{Cint}, Ptr{Float64}, Cint),
            st.ctx, p, cycles, conv, hist, nglobal, ghist, max(md.vcycle.nvcycles, 1)), "ebg_mg3d_run_multimulti")
    else
        check(ccall((:ebg_mg3d_run, LIB), Cint, (Ptr{Cvoid}, Ref{MGRunParams}, Ref{Cint}, Ref{Cint}, Ptr{Float64}),
                    st.ctx, p, cycles, conv, hist), "ebg_mg3d_run")
    end
    # the solution returns to level 1, or with multimulti to level 0 (the
    # global solution, which the reference reports)
    for name in (:vx, :vy, :vz, :pr, :etan, :etaxy, :etaxz, :etayz)
        a = getproperty(name in (:vx, :vy, :vz, :pr) ? src : lv, name).array
        dp = dev!(st, Symbol(:mg3_, name), a)
        copy_field(ids[name], dp, C_NULL)
        check(ccall((:ebg_ctx_synchronize, LIB), Cint, (Ptr{Cvoid},), st.ctx), "sync")
        d2h!(a, dp)
    end
    mr = md.mean_residuals
    for c in 1:cycles[]
        mr.resx_principle[c] = hist[4c - 3]
        mr.resy_principle[c] = hist[4c - 2]
        mr.resz_principle[c] = hist[4c - 1]
        mr.resc_principle[c] = hist[4c]
    end
    for c in 1:nglobal[]
        mr.resx_global[c] = ghist[4c - 3]
        mr.resy_global[c] = ghist[4c - 2]
        mr.resz_global[c] = ghist[4c - 1]
        mr.resc_global[c] = ghist[4c]
    end
    return Int(cycles[]), conv[] != 0
end

# ---------------------------------------------------------------- heat path

"Field order of ebg_marker_heat_props / ebg_heat_grid_fields / ebg_thermal_bc (include/earthbox_gpu.h)."
struct MarkerHeatProps
    marknum::Int64
    x::Ptr{Cvoid}; y::Ptr{Cvoid}; inside_flags::Ptr{Cvoid}
    TK::Ptr{Cvoid}; kt::Ptr{Cvoid}; rhocp::Ptr{Cvoid}; ha::Ptr{Cvoid}
    matid::Ptr{Cvoid}; mat_hr::Ptr{Cvoid}; nmat::Int64
    serpentinization_heat_production::Ptr{Cvoid}
end
struct HeatGridFields
    tk1::Ptr{Cvoid}; kt1::Ptr{Cvoid}; rhocp1::Ptr{Cvoid}; hr1::Ptr{Cvoid}; ha1::Ptr{Cvoid}
    wtnodes::Ptr{Cvoid}; marker_xn::Ptr{Cvoid}; marker_yn::Ptr{Cvoid}
end
struct ThermalBC
    btopt::Ptr{Cvoid}; bbottomt::Ptr{Cvoid}; bleftt::Ptr{Cvoid}; brightt::Ptr{Cvoid}
end

function thermal_bc!(st::GPUState, model)
    b = model.bcs.arrays.temperature
    return ThermalBC(h2d!(st, :btopt, b.btopt.array), h2d!(st, :bbottomt, b.bbottomt.array),
                     h2d!(st, :bleftt, b.bleftt.array), h2d!(st, :brightt, b.brightt.array))
end

sync!(st::GPUState) = check(ccall((:ebg_ctx_synchronize, LIB), Cint, (Ptr{Cvoid},), st.ctx), "sync")

"""
    interpolate_marker_heat_parameters_to_grid_gpu!(st, model, inside_flags)

GPU replacement of Interpolation.interpolate_marker_heat_parameters_to_grid!
(interpolation/Interpolation.jl:77-85, called at TimeLoop.jl:220): the heat
backups (tk1 falls back to tk2, BackupTransportArrays.jl:154-163), the
inclusive scatter of marker_TK, marker_kt, marker_rhocp, the heat production
and marker_ha (BilinearNumerator.jl:31-141), the averages and the thermal
boundary conditions on tk1 (ApplyThermalBC.jl:20-44).  Bit-identical.
"""
function interpolate_marker_heat_parameters_to_grid_gpu!(st::GPUState, model, inside_flags::Vector{Int8})
    set_grid!(st, model)
    m = model.markers.arrays
    he = model.heat_equation.arrays
    n = model.markers.parameters.distribution.marknum.value
    mat_hr = model.materials.arrays.mat_hr.array
    up(k, x) = h2d!(st, k, x)
    serp = model.materials.parameters.serpentinization.iuse_serpentinization.value == 1 ?
           up(:mserp, m.material.marker_serpentinization_heat_production.array) : C_NULL
    props = MarkerHeatProps(n, up(:mx, m.location.marker_x.array), up(:my, m.location.marker_y.array),
        up(:flags, inside_flags), up(:mTK, m.thermal.marker_TK.array), up(:mkt, m.thermal.marker_kt.array),
        up(:mrhocp, m.thermal.marker_rhocp.array), up(:mha, m.thermal.marker_ha.array),
        up(:matid, m.material.marker_matid.array), up(:mat_hr, mat_hr), length(mat_hr), serp)
    # tk1 receives tk2 as its backup; the others back up their previous values
    grids = (he.temperature.tk1.array, he.thermal_conductivity.kt1.array, he.rhocp.rhocp1.array,
             he.radiogenic_production.hr1.array, he.adiabatic_production.ha1.array,
             model.interpolation.arrays.grid_weights.wtnodes.array)
    keys = (:tk1, :kt1, :rhocp1, :hr1, :ha1, :wtnodes)
    ptrs = [h2d!(st, k, x) for (k, x) in zip(keys, grids)]
    h2d!(st, :tk1, he.temperature.tk2.array)
    f = HeatGridFields(ptrs..., C_NULL, C_NULL)
    check(ccall((:ebg_interpolate_marker_heat_parameters_to_grid, LIB), Cint,
        (Ptr{Cvoid}, Ref{MarkerHeatProps}, Ref{HeatGridFields}), st.ctx, props, f),
        "ebg_interpolate_marker_heat_parameters_to_grid")
    bc = thermal_bc!(st, model)
    check(ccall((:ebg_apply_thermal_bc, LIB), Cint, (Ptr{Cvoid}, Ref{ThermalBC}, Ptr{Cvoid}),
        st.ctx, bc, st.buffers[:tk1]), "ebg_apply_thermal_bc")
    sync!(st)
    for (k, x) in zip(keys, grids)
        d2h!(x, st.buffers[k])
    end
    return nothing
end

"Field order of ebg_heat_rhs_in, ebg_heat_loop_in/out/stats, ebg_marker_cells, ebg_subgrid_heat_in."
struct HeatRhsIn
    hr1::Ptr{Cvoid}; ha1::Ptr{Cvoid}; rho1::Ptr{Cvoid}; vx1::Ptr{Cvoid}; vy1::Ptr{Cvoid}
    sxy2::Ptr{Cvoid}; etas1::Ptr{Cvoid}; sxx2::Ptr{Cvoid}; etan1::Ptr{Cvoid}
    gravity_x::Cdouble; gravity_y::Cdouble
    iuse_adiabatic_heating::Int32; iuse_shear_heating::Int32
end
struct HeatLoopIn
    RT1::Ptr{Cvoid}; rhocp1::Ptr{Cvoid}; kt1::Ptr{Cvoid}
    bc::ThermalBC
    timestep::Cdouble; max_temp_change::Cdouble; max_substeps::Int32
    tk1::Ptr{Cvoid}
end
struct HeatLoopOut
    tk0::Ptr{Cvoid}; tk2::Ptr{Cvoid}; rest::Ptr{Cvoid}; dtk1::Ptr{Cvoid}
end
mutable struct HeatLoopStats
    substeps::Int32; solves::Int32; iterations::Int32; not_converged::Int32
    last_relres::Cdouble; timestep_last::Cdouble; timestep_sum::Cdouble
    HeatLoopStats() = new(0, 0, 0, 0, 0.0, 0.0, 0.0)
end
struct MarkerCells
    marknum::Int64
    x::Ptr{Cvoid}; y::Ptr{Cvoid}; inside_flags::Ptr{Cvoid}; marker_xn::Ptr{Cvoid}; marker_yn::Ptr{Cvoid}
end
struct SubgridHeatIn
    markers::MarkerCells
    tk1::Ptr{Cvoid}; kt1::Ptr{Cvoid}; rhocp1::Ptr{Cvoid}
    timestep::Cdouble; subgrid_diff_coef_temp::Cdouble
end

function marker_cells!(st::GPUState, model, inside_flags)
    m = model.markers.arrays
    n = model.markers.parameters.distribution.marknum.value
    return MarkerCells(n, h2d!(st, :mx, m.location.marker_x.array), h2d!(st, :my, m.location.marker_y.array),
        h2d!(st, :flags, inside_flags), h2d!(st, :xn, m.grid_marker_relationship.marker_xn.array),
        h2d!(st, :yn, m.grid_marker_relationship.marker_yn.array))
end

"""
    update_marker_temp_for_grid_and_subgrid_changes_gpu!(st, model, inside_flags) -> stats

GPU replacement of HeatSolver.update_marker_temp_for_grid_and_subgrid_changes!
(heat_solver/HeatSolver.jl:77-101, TimeLoop.jl:250-257): the right-hand side
(HeatRhs.jl), the adaptive heat loop around the heat solve (HeatLoop.jl:12-78;
the direct solve of SystemSolver.jl:32-66 becomes a multigrid-preconditioned
FGMRES to ||R - L T|| / ||R|| <= 1e-12), dtk1 = tk2 - tk1, the sub-grid
diffusion when subgrid_diff_coef_temp > 0 (SubgridDiffusion.jl:52-185) and
the marker temperature update (MarkerTemperature.jl:18-45).  Writes back RT1,
tk0, tk2, rest, dtk1, dtkn and marker_TK; nothing when the heat solver is off
(HeatSolver.use_heat_solver).
"""
function update_marker_temp_for_grid_and_subgrid_changes_gpu!(st::GPUState, model, inside_flags::Vector{Int8})
    he = model.heat_equation.arrays
    hp = model.heat_equation.parameters
    dt = model.timestep.parameters.main_time_loop.timestep.value
    limit = hp.temp_change_limit.max_temp_change.value
    (dt > 0 && limit > 0 && hp.heat_options.iuse_heat.value == 1) || return nothing
    set_grid!(st, model)
    sc = model.stokes_continuity.arrays
    gp = model.gravity.parameters
    up(k, x) = h2d!(st, k, x)
    rin = HeatRhsIn(up(:hr1, he.radiogenic_production.hr1.array), up(:ha1, he.adiabatic_production.ha1.array),
        up(:rho1, sc.density.rho1.array), up(:vx1, sc.staggered_grid_velocity.vx1.array),
        up(:vy1, sc.staggered_grid_velocity.vy1.array), up(:sxy2, sc.stress.sxy2.array),
        up(:etas1, sc.viscosity.etas1.array), up(:sxx2, sc.stress.sxx2.array), up(:etan1, sc.viscosity.etan1.array),
        gp.gravity_x.value, gp.gravity_y.value, Int32(hp.heat_options.iuse_adiabatic_heating.value),
        Int32(hp.heat_options.iuse_shear_heating.value))
    RT1 = up(:RT1, he.rhs.RT1.array)
    check(ccall((:ebg_heat_rhs, LIB), Cint, (Ptr{Cvoid}, Ref{HeatRhsIn}, Ptr{Cvoid}), st.ctx, rin, RT1),
          "ebg_heat_rhs")
    tk1 = up(:tk1, he.temperature.tk1.array)
    lin = HeatLoopIn(RT1, up(:rhocp1, he.rhocp.rhocp1.array), up(:kt1, he.thermal_conductivity.kt1.array),
                     thermal_bc!(st, model), dt, limit, Int32(0), tk1)
    # initialize_old_temperature_array!: tk0 starts as tk1
    lout = HeatLoopOut(up(:tk0, he.temperature.tk1.array), up(:tk2, he.temperature.tk2.array),
                       up(:rest, he.residuals.rest.array), up(:dtk1, he.temperature.dtk1.array))
    stats = HeatLoopStats()
    rc = check(ccall((:ebg_run_conductive_heat_solver_loop, LIB), Cint,
        (Ptr{Cvoid}, Ref{HeatLoopIn}, Ref{HeatLoopOut}, Ptr{Cvoid}, Ref{HeatLoopStats}),
        st.ctx, lin, lout, C_NULL, stats), "ebg_run_conductive_heat_solver_loop")
    rc == 0 || @warn "heat solve missed its tolerance" stats.last_relres
    cells = marker_cells!(st, model, inside_flags)
    TK = up(:mTK, model.markers.arrays.thermal.marker_TK.array)
    coef = model.markers.parameters.subgrid_diffusion.subgrid_diff_coef_temp.value
    if coef > 0
        sin = SubgridHeatIn(cells, tk1, st.buffers[:kt1], st.buffers[:rhocp1], dt, coef)
        check(ccall((:ebg_subgrid_heat_diffusion_update, LIB), Cint,
            (Ptr{Cvoid}, Ref{SubgridHeatIn}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
            st.ctx, sin, TK, up(:dtkn, he.temperature.dtkn.array), st.buffers[:dtk1],
            up(:wtnodes, model.interpolation.arrays.grid_weights.wtnodes.array)),
            "ebg_subgrid_heat_diffusion_update")
    end
    check(ccall((:ebg_grid_to_marker_temperature, LIB), Cint,
        (Ptr{Cvoid}, Ref{MarkerCells}, Ptr{Cvoid}, Cint, Ptr{Cvoid}), st.ctx, cells, st.buffers[:dtk1], 1, TK),
        "ebg_grid_to_marker_temperature")
    sync!(st)
    for (k, x) in ((:RT1, he.rhs.RT1.array), (:tk0, he.temperature.tk0.array), (:tk2, he.temperature.tk2.array),
                   (:rest, he.residuals.rest.array), (:dtk1, he.temperature.dtk1.array),
                   (:mTK, model.markers.arrays.thermal.marker_TK.array))
        d2h!(x, st.buffers[k])
    end
    if coef > 0
        d2h!(he.temperature.dtkn.array, st.buffers[:dtkn])
        d2h!(model.interpolation.arrays.grid_weights.wtnodes.array, st.buffers[:wtnodes])
    end
    return stats
end

"""
    apply_sticky_temperature_correction_gpu!(st, model, inside_flags, matid_sticky_air, matid_sticky_water)

GPU replacement of MarkerTemperature.apply_sticky_temperature_correction!
(MarkerTemperature.jl:52-72); the caller passes the ids of
MaterialGroupIDs.get_sticky_material_ids(model).
"""
function apply_sticky_temperature_correction_gpu!(st::GPUState, model, inside_flags::Vector{Int8},
                                                  matid_sticky_air::Integer, matid_sticky_water::Integer)
    model.heat_equation.parameters.heat_options.iuse_sticky_correction.value == 1 || return nothing
    set_grid!(st, model)
    m = model.markers.arrays
    cells = marker_cells!(st, model, inside_flags)
    TK = h2d!(st, :mTK, m.thermal.marker_TK.array)
    check(ccall((:ebg_apply_sticky_temperature_correction, LIB), Cint,
        (Ptr{Cvoid}, Ref{MarkerCells}, Ptr{Cvoid}, Cint, Cint, Cdouble, Ptr{Cvoid}), st.ctx, cells,
        h2d!(st, :matid, m.material.marker_matid.array), matid_sticky_air, matid_sticky_water,
        model.bcs.parameters.temperature.temperature_top.value, TK), "ebg_apply_sticky_temperature_correction")
    sync!(st)
    d2h!(m.thermal.marker_TK.array, TK)
    return nothing
end

"Field order of ebg_subgrid_stress_in / ebg_subgrid_stress_out."
struct SubgridStressIn
    markers::MarkerCells
    matid::Ptr{Cvoid}; mat_mu::Ptr{Cvoid}; nmat::Int64; eta::Ptr{Cvoid}
    sxy1::Ptr{Cvoid}; sxx1::Ptr{Cvoid}
    timestep::Cdouble; subgrid_diff_coef_stress::Cdouble
end
struct SubgridStressOut
    marker_sxy::Ptr{Cvoid}; marker_sxx::Ptr{Cvoid}; dsxy::Ptr{Cvoid}; dsxx::Ptr{Cvoid}
    dsxyn::Ptr{Cvoid}; dsxxn::Ptr{Cvoid}; wtetas::Ptr{Cvoid}; wtetan::Ptr{Cvoid}
end

"""
    update_subgrid_stress_gpu!(st, model, inside_flags)

GPU replacement of SubGridStress.Update.update_subgrid_stress!
(subgrid_stress/update/Update.jl:55-71): relaxed marker stresses, their
scatter into dsxyn and dsxxn and the remaining changes dsxy, dsxx.
"""
function update_subgrid_stress_gpu!(st::GPUState, model, inside_flags::Vector{Int8})
    set_grid!(st, model)
    m = model.markers.arrays
    sc = model.stokes_continuity.arrays
    mat_mu = model.materials.arrays.mat_mu.array
    up(k, x) = h2d!(st, k, x)
    sin = SubgridStressIn(marker_cells!(st, model, inside_flags), up(:matid, m.material.marker_matid.array),
        up(:mat_mu, mat_mu), length(mat_mu), up(:meta, m.rheology.marker_eta.array),
        up(:sxy1, sc.stress.sxy1.array), up(:sxx1, sc.stress.sxx1.array),
        model.timestep.parameters.main_time_loop.timestep.value,
        model.markers.parameters.subgrid_diffusion.subgrid_diff_coef_stress.value)
    iw = model.interpolation.arrays.grid_weights
    outs = (m.stress.marker_sxy.array, m.stress.marker_sxx.array, sc.stress_change.dsxy.array,
            sc.stress_change.dsxx.array, sc.stress_change.dsxyn.array, sc.stress_change.dsxxn.array,
            iw.wtetas.array, iw.wtetan.array)
    keys = (:sxy, :sxx, :dsxy, :dsxx, :dsxyn, :dsxxn, :wtetas, :wtetan)
    sout = SubgridStressOut([h2d!(st, k, x) for (k, x) in zip(keys, outs)]...)
    check(ccall((:ebg_update_subgrid_stress, LIB), Cint,
        (Ptr{Cvoid}, Ref{SubgridStressIn}, Ref{SubgridStressOut}), st.ctx, sin, sout), "ebg_update_subgrid_stress")
    sync!(st)
    for (k, x) in zip(keys, outs)
        d2h!(x, st.buffers[k])
    end
    return nothing
end

"Field order of ebg_marker_plastic_in / ebg_marker_plastic_out."
struct MarkerPlasticIn
    marknum::Int64
    inside_flags::Ptr{Cvoid}; y::Ptr{Cvoid}; matid::Ptr{Cvoid}; mat_mu::Ptr{Cvoid}
    nmat::Int64
    cohesion::Ptr{Cvoid}; fric::Ptr{Cvoid}; exx::Ptr{Cvoid}; exy::Ptr{Cvoid}; sr_ratio::Ptr{Cvoid}; pr::Ptr{Cvoid}
    forecast::Int32; ntimestep::Int32
    timestep::Cdouble; y_sealevel::Cdouble
    viscosity_min::Cdouble; viscosity_max::Cdouble; yield_stress_min::Cdouble; yield_stress_max::Cdouble
    iuse_fluid_pressure_for_yield::Int32; no_yielding_in_mobile_wall::Int32
    matid_sticky_air::Int32; matid_mobile_wall::Int32; matid_plate_extension::Int32
end
struct MarkerPlasticOut
    sxx::Ptr{Cvoid}; sxy::Ptr{Cvoid}; eta::Ptr{Cvoid}; pfailure::Ptr{Cvoid}
end

"""
    update_plastic_yielding_gpu!(st, model, inside_flags, no_yielding_in_mobile_wall)

GPU replacement of MarkerPlasticity.update_plastic_yielding!
(rheology/plastic_failure/markers/MarkerPlasticity.jl:34-43), the marker step
of PlasticityLoopMarkers.picard_loop (PlasticityLoopMarkers.jl:51): the
viscoelastic (itype_plasticity 0) or purely elastic (1) stress forecast,
the yield test and the corrected marker viscosity (and, for the viscoelastic
forecast, stresses); marker_pfailure keeps 1 once set.
"""
function update_plastic_yielding_gpu!(st::GPUState, model, inside_flags::Vector{Int8},
                                      no_yielding_in_mobile_wall::Bool)
    m = model.markers.arrays
    mp = model.materials.parameters
    tl = model.timestep.parameters.main_time_loop
    mat_mu = model.materials.arrays.mat_mu.array
    dom = model.materials.dicts.matid_domains
    n = model.markers.parameters.distribution.marknum.value
    up(k, x) = h2d!(st, k, x)
    pin = MarkerPlasticIn(n, up(:flags, inside_flags), up(:my, m.location.marker_y.array),
        up(:matid, m.material.marker_matid.array), up(:mat_mu, mat_mu), length(mat_mu),
        up(:mcoh, m.rheology.marker_cohesion.array), up(:mfric, m.rheology.marker_fric.array),
        up(:mexx, m.strain.marker_exx.array), up(:mexy, m.strain.marker_exy.array),
        up(:msrr, m.strain.marker_sr_ratio.array), up(:mpr, m.pressure.marker_pr.array),
        Int32(mp.material_description.itype_plasticity.value), Int32(tl.ntimestep.value), tl.timestep.value,
        model.topography.parameters.sealevel.y_sealevel.value,
        mp.viscosity_limits.viscosity_min.value, mp.viscosity_limits.viscosity_max.value,
        mp.stress_limits_yield.yield_stress_min.value, mp.stress_limits_yield.yield_stress_max.value,
        Int32(mp.stress_limits_yield.iuse_fluid_pressure_for_yield.value), Int32(no_yielding_in_mobile_wall),
        Int32(dom["Atmosphere"]), Int32(dom["MobileWall"]), Int32(dom["PlateExtension"]))
    outs = (m.stress.marker_sxx.array, m.stress.marker_sxy.array, m.rheology.marker_eta.array,
            m.rheology.marker_pfailure.array)
    keys = (:msxx, :msxy, :meta, :mpf)
    pout = MarkerPlasticOut([h2d!(st, k, x) for (k, x) in zip(keys, outs)]...)
    check(ccall((:ebg_update_marker_plastic_yielding, LIB), Cint,
        (Ptr{Cvoid}, Ref{MarkerPlasticIn}, Ref{MarkerPlasticOut}), st.ctx, pin, pout),
        "ebg_update_marker_plastic_yielding")
    sync!(st)
    for (k, x) in zip(keys, outs)
        d2h!(x, st.buffers[k])
    end
    return nothing
end

"""
    interpolate_viscosity_to_transport_arrays_gpu!(st, model, inside_flags)

GPU replacement of BackupTransportArrays.backup_and_clear_viscosity_transport_arrays!
followed by MarkersToStokesArrays.interpolate_viscosity_to_transport_arrays!
(MarkersToStokesArrays.jl:60-64; BilinearNumerator.jl:470-533,
BilinearAverage.jl:213-227), PlasticityLoopMarkers.jl:50,52: marker_eta alone
scattered to etas1 (exclusive rule) and etan1 (central rule), the previous
values kept (and copied to etas0, etan0) where a node receives no weight.
The weight sums are recomputed from the same markers, so they equal the
wtetas, wtetan of the full scatter.
"""
function interpolate_viscosity_to_transport_arrays_gpu!(st::GPUState, model, inside_flags::Vector{Int8})
    set_grid!(st, model)
    m = model.markers.arrays
    v = model.stokes_continuity.arrays.viscosity
    copyto!(v.etas0.array, v.etas1.array)
    copyto!(v.etan0.array, v.etan1.array)
    n = model.markers.parameters.distribution.marknum.value
    up(k, x) = h2d!(st, k, x)
    N = C_NULL
    props = MarkerStokesProps(n, up(:mx, m.location.marker_x.array), up(:my, m.location.marker_y.array),
        up(:flags, inside_flags), N, N, N, N, up(:meta, m.rheology.marker_eta.array), N, N, N, N, N, N, N, 0)
    es, en = up(:etas1, v.etas1.array), up(:etan1, v.etan1.array)
    f = StokesGridFields(N, N, es, en, N, N, N, N, N, N, N, N, N, N, N, N, N, N, N, N, N)
    check(ccall((:ebg_interpolate_marker_stokes_parameters_to_grid, LIB), Cint,
        (Ptr{Cvoid}, Ref{MarkerStokesProps}, Ref{StokesGridFields}), st.ctx, props, f),
        "ebg_interpolate_marker_stokes_parameters_to_grid")
    sync!(st)
    d2h!(v.etas1.array, es)
    d2h!(v.etan1.array, en)
    return nothing
end

end # module
