/*
 * earthbox_gpu.h -- C ABI of the MI355X-native EarthBox hot-path library.
 *
 * Conventions shared by every entry point:
 *   - All array arguments are DEVICE pointers (allocate with ebg_device_alloc or
 *     any HIP allocator).  ebg_memcpy_h2d / ebg_memcpy_d2h move host arrays, so a
 *     caller needs nothing but this header (the Julia side binds it with ccall).
 *   - 2-D arrays are column-major with the y (row) index fastest, exactly as the
 *     reference's Julia Matrix{Float64}; shapes are the reference's:
 *       basic grid ynum x xnum, vx grid (ynum+1) x xnum, vy grid ynum x (xnum+1),
 *       pressure grid (ynum-1) x (xnum-1).
 *   - Marker arrays are SoA vectors of length marknum; marker cell indices are
 *     1-based Int32 as in the reference (marker_xn / marker_yn).
 *   - Work is enqueued on the context's stream; call ebg_ctx_synchronize (or
 *     synchronise the stream) before reading results on the host.
 *   - Return value: EBG_OK (0) on success, > 0 "completed but not converged"
 *     (solver only; outputs are still written), < 0 hard error with a message
 *     available from ebg_last_error().  The library never aborts the process.
 */
#ifndef EARTHBOX_GPU_H
#define EARTHBOX_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ebg_ctx ebg_ctx;

enum {
    EBG_OK = 0,
    EBG_NOT_CONVERGED = 1,
    EBG_ERR_ARG = -1,
    EBG_ERR_HIP = -2,
    EBG_ERR_STATE = -3,
    EBG_ERR_NOMEM = -4,
    EBG_ERR_NUMERIC = -5
};

/* ---- context, memory, errors (no reference counterpart: plumbing) ---- */
int ebg_ctx_create(ebg_ctx **out, int device);
int ebg_ctx_destroy(ebg_ctx *ctx);
/* Enqueue all later work on an external hipStream_t (e.g. the caller's current
 * stream); NULL selects the default (null) stream.  The context's private
 * non-blocking stream, created by ebg_ctx_create, is released. */
int ebg_ctx_set_stream(ebg_ctx *ctx, void *hip_stream);
int ebg_ctx_synchronize(ebg_ctx *ctx);
const char *ebg_last_error(void);
int ebg_device_alloc(void **ptr, size_t bytes);
int ebg_device_free(void *ptr);
int ebg_memcpy_h2d(void *dst_dev, const void *src_host, size_t bytes);
int ebg_memcpy_d2h(void *dst_host, const void *src_dev, size_t bytes);

/* ---- staggered grid ----
 * Replaces the grid arrays the reference reads from ModelData
 * (staggered_grid/types_manager/uniform/core/Initialization.jl:27-43,
 *  staggered_grid/utils/VxGridTools.jl:40-99, VyGridTools.jl:41-100).
 * HOST pointers: gridx_b[xnum], gridy_b[ynum], xstp_b[xnum-1], ystp_b[ynum-1],
 * gridx_vy[xnum+1], xstp_vy[xnum], gridy_vx[ynum+1], ystp_vx[ynum].
 * The domain limits used for inside flags are gridx_b[1], gridx_b[xnum],
 * gridy_b[1], gridy_b[ynum] (staggered_grid/utils/Dimensions.jl:6-19).
 * Basic coordinates must increase and every step must lie in
 * [2^-100, 2^100] (EBG_ERR_ARG otherwise): the marker kernels' correctly
 * rounded division by the steps relies on it. */
int ebg_grid2d_set(ebg_ctx *ctx, int64_t xnum, int64_t ynum,
                   const double *gridx_b, const double *gridy_b,
                   const double *xstp_b, const double *ystp_b,
                   const double *gridx_vy, const double *xstp_vy,
                   const double *gridy_vx, const double *ystp_vx,
                   double xstpavg, double ystpavg);

/* ---- markers: inside flags and cell indices ----
 * Replaces GridFuncs.get_marker_inside_flags (src/utils/GridFuncs.jl:54-80). */
int ebg_marker_inside_flags(ebg_ctx *ctx, int64_t marknum, const double *marker_x,
                            const double *marker_y, int8_t *inside_flags);

/* Replaces MarkerGridMapping.calc_upper_left_basic_grid_cell_indices_for_markers!
 * (interpolation/mapping/MarkerGridMapping.jl:99-129): bit-exact
 * clamp(searchsortedlast(grid, x), 1, n-1).  Markers with flag != 1 keep their
 * previous xn/yn. */
int ebg_marker_cell_indices(ebg_ctx *ctx, int64_t marknum, const double *marker_x,
                            const double *marker_y, const int8_t *inside_flags,
                            int32_t *marker_xn, int32_t *marker_yn);

/* ---- RK4 advection ----
 * Replaces Advection.advect_markers_and_rotate_stress_tensor!
 * (advection/Advection.jl:109-114), i.e. in one fused pass:
 *   VelocityAndSpin.update_marker_velocity_and_spin_using_runge_kutta!
 *     (advection/velocity_and_spin/VelocityAndSpin.jl:45-129),
 *   MarkersLocation.update! (advection/location/MarkersLocation.jl:11-41) when
 *     update_location != 0,
 *   MarkersStressRotation.update! (advection/stress_rotation/
 *     MarkersStressRotation.jl:15-74) when marker_sxx and marker_sxy are non-NULL.
 * order: runge_kutta_order_max (4 = RK4, otherwise stage-1 velocity).
 * marker_xn/yn: stage-1 cell indices as stored by the scatter step; NULL means
 *   recompute them from marker_x/y (identical result, 8 B/marker less traffic).
 * inside_flags: NULL means recompute from the grid limits.
 * marker_vx/vy/spin: optional outputs (NULL = not written).
 * Arithmetic (ebg_rk4_set_arithmetic):
 *   EBG_RK4_EXACT (default): FP64 with the reference's evaluation order (no
 *     contraction, correctly rounded divisions), so velocities, spins and
 *     positions are bit-identical to the reference loops;
 *   EBG_RK4_TOLERANCE: FP64 with fused multiply-adds in cell coordinates
 *     from the axes' nominal uniform origin and step (reciprocal step
 *     instead of division), used when both axes are uniform to 1e-9 of a
 *     step (else the exact kernel runs); each wavefront's patch of the
 *     grids is staged in LDS (ebg_rk4_set_tiling).  Positions agree with the
 *     exact mode to far below 1e-10 of the domain size; the stage-1 cells
 *     given in marker_xn/yn are used as given. */
int ebg_rk4_advect(ebg_ctx *ctx, int64_t marknum, int order, double timestep,
                   double *marker_x, double *marker_y,
                   const int32_t *marker_xn, const int32_t *marker_yn,
                   const int8_t *inside_flags,
                   const double *vx1, const double *vy1, const double *esp,
                   double *marker_vx, double *marker_vy, double *marker_spin,
                   double *marker_sxx, double *marker_sxy, int update_location);

/* Selects the arithmetic of ebg_rk4_advect on this context (no counterpart
 * in the reference, whose RK4 has one arithmetic). */
enum { EBG_RK4_EXACT = 0, EBG_RK4_TOLERANCE = 1 };
int ebg_rk4_set_arithmetic(ebg_ctx *ctx, int mode);

/* EBG_RK4_TOLERANCE only: on = 1 (default) stages each wavefront's patch of
 * the velocity (and spin) grids in LDS and serves the stage reads from it;
 * on = 0 reads every stage value from global memory.  Both give identical
 * results.  No counterpart in the reference (a tuning switch of this GPU
 * kernel). */
int ebg_rk4_set_tiling(ebg_ctx *ctx, int on);

/* 1 when the current grid lets EBG_RK4_TOLERANCE use its fast kernel (both
 * axes uniform to 1e-9 of a step), else 0; -1 without a grid. */
int ebg_rk4_tolerance_kernel_available(ebg_ctx *ctx);

/* Replaces GridStrainRate.calculate_spin!
 * (stokes_continuity_solver/utils/strain_rate/GridStrainRate.jl:161-182). */
int ebg_calculate_spin(ebg_ctx *ctx, const double *vx1, const double *vy1, double *esp);

/* ---- marker -> grid interpolation of the Stokes transport properties ----
 * Per-marker DEVICE arrays (length marknum); any property pointer may be NULL
 * (its contribution is then 0).  matid (Int16, 1-based) and mat_mu (DEVICE,
 * nmat shear moduli, the reference's materials.arrays.mat_mu) go together. */
typedef struct {
    int64_t marknum;
    const double *x, *y;
    const int8_t *inside_flags;      /* NULL: computed from the grid limits */
    const double *rho, *cohesion, *fric, *eta_flow, *eta, *sxy, *sxx, *meltfrac;
    const float *pfailure, *dilatation;  /* Float32 as in the reference */
    const int16_t *matid;
    const double *mat_mu;
    int64_t nmat;
} ebg_marker_stokes_props;

/* DEVICE grid arrays, in/out: on entry the previous values (the backup used
 * where a node receives no marker weight, BackupTransportArrays.jl:33-104),
 * on exit the new averages.  Any may be NULL (not written).  Shapes: basic
 * (ynum,xnum): rho1 etas1 eta_flow mus1 sxy1 plastics cohesion_grid
 * fric_degrees_grid; pressure (ynum-1,xnum-1): etan1 mun1 sxx1 plasticn
 * dilatation_grid extractable_meltfrac_grid; vy (ynum,xnum+1): rho1_vy.
 * Weight sums (optional outputs): wtnodes, wtetas (ynum,xnum), wtetan
 * (ynum-1,xnum-1), wtnodes_vy (ynum,xnum+1).  marker_xn/yn (optional) receive
 * the bit-exact 1-based cell indices of inside markers. */
typedef struct {
    double *rho1, *rho1_vy, *etas1, *etan1, *eta_flow, *mus1, *mun1, *sxy1, *sxx1;
    double *plastics, *plasticn, *cohesion_grid, *fric_degrees_grid, *dilatation_grid;
    double *extractable_meltfrac_grid;
    double *wtnodes, *wtetas, *wtetan, *wtnodes_vy;
    int32_t *marker_xn, *marker_yn;
} ebg_stokes_grid_fields;

/* Replaces Interpolation.initialize_bilinear_interpolation! +
 * interpolate_marker_stokes_parameters_to_grid! (interpolation/Interpolation.jl:61-95):
 * MarkerGridMapping.jl:99-265, MarkerGridNodeWeights.jl:68-150,
 * BilinearDenominator.jl:34-182, BilinearNumerator.jl:205-454, WeightFuncs.jl:67-190,
 * BilinearAverage.jl:57-178.  Every node sum is accumulated in the reference's
 * serial marker order (no atomics): results equal the serial loops bit for bit
 * (the friction angle's asin aside). */
int ebg_interpolate_marker_stokes_parameters_to_grid(ebg_ctx *ctx,
                                                     const ebg_marker_stokes_props *props,
                                                     ebg_stokes_grid_fields *fields);

/* Summation order of both scatters on this context (no counterpart in the
 * reference, whose serial loops have one order):
 *   EBG_SCATTER_SERIAL_ORDER (default): every node sum in the reference's
 *     serial marker order, bit-identical to the serial loops;
 *   EBG_SCATTER_CELL_ORDER: each cell's markers summed privately in marker
 *     order into per-corner partial sums, each node sum the fixed-order sum
 *     of its neighbouring cells' partials; deterministic, no floating-point
 *     atomics, within a few ulps of the sum of the absolute terms of the
 *     serial order (pressure nodes bit-identical), several times faster. */
enum { EBG_SCATTER_SERIAL_ORDER = 0, EBG_SCATTER_CELL_ORDER = 1 };
int ebg_scatter_set_order(ebg_ctx *ctx, int order);

/* ---- marker -> grid heat-transport properties ----
 * DEVICE marker arrays (length marknum): x, y; inside_flags optional (NULL:
 * computed from the grid limits); TK, kt, rhocp, ha; matid (1-based) with the
 * material table mat_hr[nmat] for the heat production, plus
 * serpentinization_heat_production (optional: added when given, as with
 * iuse_serpentinization = 1, BilinearNumerator.jl:107-113).  A property may be NULL when its output is. */
typedef struct {
    int64_t marknum;
    const double *x, *y;
    const int8_t *inside_flags;
    const double *TK, *kt, *rhocp, *ha;
    const int16_t *matid;
    const double *mat_hr;
    int64_t nmat;
    const double *serpentinization_heat_production;
} ebg_marker_heat_props;

/* DEVICE basic-grid arrays (ynum,xnum), in/out: on entry the backups kept where
 * a node receives no marker weight, on exit the averages.  The reference backs
 * up tk1 from the previous heat solution tk2 (BackupTransportArrays.jl:154-163),
 * so tk1 must hold tk2's values on entry; kt1, rhocp1, hr1 and ha1 hold their
 * previous values.  Any may be NULL (not computed).  wtnodes (optional out):
 * the inclusive weight sums; marker_xn/yn (optional out): cell indices. */
typedef struct {
    double *tk1, *kt1, *rhocp1, *hr1, *ha1;
    double *wtnodes;
    int32_t *marker_xn, *marker_yn;
} ebg_heat_grid_fields;

/* Replaces MarkerToHeatArrays.interpolate_markers_to_transport_arrays!
 * without its last step (interpolation/utils/MarkerToHeatArrays.jl:29-35:
 * BackupTransportArrays.backup_and_clear_heat_transport_arrays!,
 * BilinearNumerator.calc_marker_weight_sums_for_basic_grid_heat!
 * (BilinearNumerator.jl:31-141) and BilinearAverage.calc_marker_average_at_basic_nodes_heat!
 * (BilinearAverage.jl:10-50)); the thermal boundary conditions follow with
 * ebg_apply_thermal_bc.  Node sums in serial marker order, bit-identical to
 * the serial loops.  Used with only tk1 requested it is also the scatter of a
 * per-marker quantity with the inclusive rule (the sub-grid temperature
 * change, SubgridDiffusion.jl:129-176). */
int ebg_interpolate_marker_heat_parameters_to_grid(ebg_ctx *ctx, const ebg_marker_heat_props *props,
                                                   ebg_heat_grid_fields *fields);

/* ---- viscoelastic terms and right-hand sides (K7) ----
 * DEVICE grid arrays with the reference shapes.  Inputs: etas1, mus1, sxy1,
 * rho1 (ynum,xnum); etan1, mun1, sxx1 (ynum-1,xnum-1); rho1_vy (ynum,xnum+1),
 * read only with interface stabilization; dilatation_grid and
 * eii_plastic_pressure (ynum-1,xnum-1), both NULL for RC1 = 0.  Outputs:
 * etas0, sxy0 (ynum,xnum); etan0, sxx0 (ynum-1,xnum-1); RX1 (ynum+1,xnum),
 * RY1 (ynum,xnum+1), RC1 (ynum-1,xnum-1), every entry written. */
typedef struct {
    const double *etas1, *mus1, *sxy1, *rho1;
    const double *etan1, *mun1, *sxx1;
    const double *rho1_vy;
    const double *dilatation_grid, *eii_plastic_pressure;
    double timestep, gravity_x, gravity_y;
    int iuse_interface_stabilization;
} ebg_viscoelastic_in;

typedef struct {
    double *etas0, *sxy0, *etan0, *sxx0;
    double *RX1, *RY1, *RC1;
} ebg_viscoelastic_out;

/* Replaces StokesViscoelasticTerms.calc_viscoelastic_terms! followed by
 * StokesContinuityRhs.viscoelastic_rhs! (solver/rhs/StokesViscoelasticTerms.jl:26-108,
 * solver/rhs/StokesContinuityRhs.jl:30-190, called at StokesContinuitySolver.jl:189-190)
 * on the context's grid; bit-identical except RC1, whose sine goes through the
 * device math library. */
int ebg_stokes_viscoelastic_rhs(ebg_ctx *ctx, const ebg_viscoelastic_in *in, ebg_viscoelastic_out *out);

/* ---- grid strain rate, forecast stress and stress change ----
 * Replaces GridStrainRate.calculate_normal_strain_rate!, calculate_shear_strain_rate!
 * and calculate_second_invariant! (stokes_continuity_solver/utils/strain_rate/
 * GridStrainRate.jl:33-159), GridStress.forecast_viscoelastic_grid_stress! and
 * calculate_deviatoric_grid_stress_change! (stokes_continuity_solver/utils/
 * GridStress.jl), the grid part of stokes_solution_processing!
 * (StokesContinuitySolver.jl:212-279); spin is ebg_calculate_spin.  DEVICE
 * arrays: in vx1 (ynum+1,xnum), vy1 (ynum,xnum+1), etas1, mus1, sxy1
 * (ynum,xnum), etan1, mun1, sxx1 (ynum-1,xnum-1); out exy, sxy2, dsxy
 * (ynum,xnum), exx, eii, sxx2, dsxx (ynum-1,xnum-1).  Bit-identical. */
typedef struct {
    const double *vx1, *vy1;
    const double *etas1, *mus1, *sxy1;
    const double *etan1, *mun1, *sxx1;
    double timestep;
} ebg_strain_stress_in;

typedef struct {
    double *exx, *exy, *eii;
    double *sxy2, *sxx2, *dsxy, *dsxx;
} ebg_strain_stress_out;

int ebg_grid_strain_rate_and_stress(ebg_ctx *ctx, const ebg_strain_stress_in *in, ebg_strain_stress_out *out);

/* ---- nodal plastic yielding (Picard loop over nodes) ----
 * DEVICE grid arrays, reference shapes: plastic_yield, etas1, eta_flow, mus1,
 * cohesion_grid, fric_degrees_grid, sxy2 (ynum,xnum); sxx2, pr1 (ynum-1,xnum-1).
 * Scalars as the reference reads them from ModelData (timestep, timesum,
 * ntimestep, viscosity limits, iuse_fluid_pressure_for_yield, y_sealevel,
 * mobile-wall geometry) and the loop's two no-yielding switches. */
typedef struct {
    const double *plastic_yield, *etas1, *eta_flow, *mus1, *cohesion_grid, *fric_degrees_grid, *sxy2;
    const double *sxx2, *pr1;
    double timestep, timesum;
    int64_t ntimestep;
    double viscosity_min, viscosity_max;
    int iuse_fluid_pressure_for_yield;
    double y_sealevel;
    int no_yielding_in_mobile_wall, no_yielding_in_plate_extension;
    double velocity_internal_x, x_left_mobile_wall, x_right_mobile_wall, y_top_mobile_wall,
        y_bottom_mobile_wall, plate_extension_thickness, plate_extension_width;
} ebg_plastic_nodes_in;

typedef struct {
    double *etavp_new, *plastic_strain_rate, *plastic_yield_new; /* DEVICE (ynum,xnum), written */
    double *yield_error;                                          /* DEVICE (ynum,xnum), updated */
    int64_t nnode_yield;                                          /* host result */
    double global_yield_error;                                    /* host result */
} ebg_plastic_nodes_out;

/* Replaces PlasticFailureNodes.update_nodes_for_plastic_yielding
 * (rheology/plastic_failure/nodes/PlasticFailureNodes.jl:24-319) on the
 * context's grid.  Per-node results follow the reference's expressions; the
 * sine, arcsine and cosine go through the device math library, and the RMS
 * of the yield errors is summed in a fixed tree order instead of node order. */
int ebg_update_nodes_for_plastic_yielding(ebg_ctx *ctx, const ebg_plastic_nodes_in *in,
                                          ebg_plastic_nodes_out *out);

/* Allocation-free variant of ebg_update_nodes_for_plastic_yielding (no
 * counterpart in the reference): for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same kernels, same results bit for bit.  Its reduction
 * scratch (partial sums and counts of at most 1024 blocks and the two results,
 * 2,050 doubles = 16.4 KB) is a static device array that each GPU holds until
 * the process ends; its device address is taken once per device by
 * ebg_ctx_create.  The node loop is grid-strided, so any grid fits the
 * scratch.  The call makes no hipMalloc/hipFree and calls no allocating
 * library routine; it still copies the count and RMS back to the host and
 * synchronizes the context's stream.  Every call on one GPU shares that
 * scratch: calls that overlap on one GPU (other streams or threads) race. */
int ebg_update_nodes_for_plastic_yielding_noalloc(ebg_ctx *ctx, const ebg_plastic_nodes_in *in,
                                                  ebg_plastic_nodes_out *out);

/* Replaces Interpolation.interpolate_basic_viscoplastic_viscosity_to_pressure_grid!
 * (interpolation/Interpolation.jl:97-109) with iuse_harmonic_avg_normal_viscosity = 1,
 * i.e. HarmonicAverage.normal_viscosity_from_harmonic_avg_shear_viscosity!: etan1
 * (ynum-1,xnum-1, DEVICE, updated) becomes the harmonic mean of the four etas1
 * around a cell where plasticn, or plastics or plastic_yield at one of its
 * corners, is positive (MathTools.harmonic4, utils/MathTools.jl:42-62).
 * Bit-identical. */
int ebg_harmonic_normal_viscosity(ebg_ctx *ctx, const double *etas1, const double *plastics,
                                  const double *plasticn, const double *plastic_yield, double *etan1);
/* Replaces PlasticStrainRateInterpolation.interpolate_at_pressure_nodes!
 * (interpolation/utils/stokes_funcs/PlasticStrainRateInterpolation.jl:18-41):
 * eii_plastic_pressure (ynum-1,xnum-1) = mean of the four eii_plastic_basic
 * corners.  Bit-identical. */
int ebg_plastic_strain_rate_at_pressure_nodes(ebg_ctx *ctx, const double *eii_plastic_basic,
                                              double *eii_plastic_pressure);

/* ---- marker rheology ----
 * Per-marker DEVICE arrays (length marknum); inside_flags NULL = every marker
 * inside.  Material tables are DEVICE, column-major with leading dimension
 * nmat, 1-based matid as the reference's materials.arrays: mat_flow_type
 * (Int64[nmat]), mat_flow (nmat x 12: 1 viscosity, 2 dislocation
 * pre-exponential, 3 stress exponent, 4 activation energy kJ/mol, 5
 * activation volume, 6 Peierls pre-exponential, 7-8 Peierls exponents m and
 * n, 9 Peierls stress MPa, 10 diffusion pre-exponential, 11 diffusion
 * activation energy, 12 diffusion activation volume), mat_plastic (nmat x 6:
 * cohesion initial/final, sine of friction initial/final, strain
 * initial/final), mat_mu (nmat).  `stages` selects which of the three
 * reference steps run, in this order, in one pass over the markers:
 *   EBG_RHEO_FAILURE: FailureProperties.calculate_failure_properties_loop!
 *     (rheology/plastic_failure/failure_properties/FailureProperties.jl):
 *     cohesion and sine of friction with linear strain weakening of the
 *     plastic strain, optional melt damage (melt_damage NULL = off);
 *     reads strain_plastic, fric_ini; writes cohesion, fric.
 *   EBG_RHEO_FLOW: CompositeViscosity.update_marker_flow_viscosity!
 *     (rheology/flow_viscosity/update/types/composite_viscosity/CompositeViscosity.jl):
 *     dislocation (power-law), Peierls and diffusion creep with the
 *     visco-elastic stress forecast found by bisection (20 steps, 1 Pa),
 *     optional viscous strain softening (ViscousStrainSoftening.jl), the
 *     result clipped to [viscosity_min, viscosity_max]; reads sxx, sxy,
 *     exx, exy, TK, pr, GII, sr_ratio; writes eta_flow, preexp.
 *   EBG_RHEO_VISCOPLASTIC: ViscoplasticViscosity.update_marker_viscoplastic_viscosity!
 *     (rheology/viscoplastic_viscosity/ViscoplasticViscosity.jl:21-61):
 *     eta = eta_flow for itype_global 0, or for itype_global 1 at
 *     ntimestep 0 or where eta is 0; reads and writes eta.
 * Same expressions and order as the reference; exp and pow are the device
 * library's, so values agree with the serial loops to rounding. */
enum { EBG_RHEO_FAILURE = 1, EBG_RHEO_FLOW = 2, EBG_RHEO_VISCOPLASTIC = 4 };
typedef struct {
    int64_t marknum;
    const int8_t *inside_flags;
    const int16_t *matid;
    const double *strain_plastic, *fric_ini, *melt_damage;
    const double *sxx, *sxy, *exx, *exy, *TK, *pr, *GII, *sr_ratio;
    int64_t nmat;
    const int64_t *mat_flow_type;
    const double *mat_flow, *mat_plastic, *mat_mu;
    double timestep, temperature_top, viscosity_min, viscosity_max, powerlaw_stress_min, gas_constant;
    int iuse_viscous_strain_soft;
    double vsoftfac;
    int itype_global;
    int64_t ntimestep;
    int stages;
} ebg_marker_rheology_in;

typedef struct {
    double *cohesion, *fric;     /* EBG_RHEO_FAILURE */
    double *eta_flow, *preexp;   /* EBG_RHEO_FLOW */
    double *eta;                 /* EBG_RHEO_VISCOPLASTIC (in/out) */
} ebg_marker_rheology_out;

/* Replaces Rheology.update_marker_rheology! (rheology/Rheology.jl:36-56, the
 * failure properties and the composite flow viscosity) and
 * Rheology.update_marker_viscoplastic_viscosity! (:70-80) with node-based
 * plasticity, for the composite-viscosity flow law (TimeLoop.jl:222-223). */
int ebg_update_marker_rheology(ebg_ctx *ctx, const ebg_marker_rheology_in *in, ebg_marker_rheology_out *out);

/* Replaces GridViscoplasticityToMarkers.calc_marker_viscoplastic_viscosity!
 * (rheology/plastic_failure/nodes/GridViscoplasticityToMarkers.jl), run after
 * the Picard loop over nodes (TimeLoop.jl:240): where a corner of the
 * marker's basic cell yields, eta becomes the yield-weighted harmonic average
 * of etas1 if below eta_flow (and pfailure 1), else eta_flow.  marker_xn/yn
 * are the 1-based basic cells the scatter stored; the distances are
 * recomputed as MarkerGridMapping.jl:185-218 does.  DEVICE: plastic_yield,
 * etas1 (ynum,xnum).  Bit-identical to the serial loop. */
int ebg_calc_marker_viscoplastic_viscosity(ebg_ctx *ctx, int64_t marknum, const double *marker_x,
                                           const double *marker_y, const int32_t *marker_xn,
                                           const int32_t *marker_yn, const int8_t *inside_flags,
                                           const double *plastic_yield, const double *etas1,
                                           const double *marker_eta_flow, double *marker_eta,
                                           float *marker_pfailure);

/* Replaces Advection.update_marker_strain! (advection/Advection.jl:169-178),
 * i.e. MarkersStrain.update_total_strain! and update_plastic_strain!
 * (advection/strain/MarkersStrain.jl): marker_GII += dt |e|, and where
 * pfailure is 1 the plastic strain rate 0.5 sii (1/eta - 1/eta_flow) (at
 * least 0) and the plastic strain with healing (at least 0).  Bit-identical. */
int ebg_update_marker_strain(ebg_ctx *ctx, int64_t marknum, const int8_t *inside_flags, double timestep,
                             double plastic_healing_rate, const double *marker_exx, const double *marker_exy,
                             const double *marker_sxx, const double *marker_sxy, const double *marker_eta,
                             const double *marker_eta_flow, const float *marker_pfailure, double *marker_GII,
                             double *marker_strain_plastic, double *marker_strain_rate_plastic);

/* ---- grid -> marker gathers after the Stokes solve ----
 * Per-marker DEVICE arrays (length marknum): x, y; marker_xn/yn (1-based basic
 * cell, as the scatter stores them); inside_flags (NULL: computed from the
 * grid limits); eta and matid (Int16, 1-based, with mat_mu of nmat shear
 * moduli), read only for the strain-rate ratio.  DEVICE grid arrays:
 * exy, sxy2, dsxy (ynum,xnum); exx, pr1, sxx2, dsxx (ynum-1,xnum-1). */
typedef struct {
    int64_t marknum;
    const double *x, *y;
    const int32_t *marker_xn, *marker_yn;
    const int8_t *inside_flags;
    const double *exx, *exy, *pr1;
    const double *sxx2, *sxy2, *dsxx, *dsxy;
    const double *eta;
    const int16_t *matid;
    const double *mat_mu;
    int64_t nmat;
    double timestep;
} ebg_marker_gather_in;

/* Per-marker DEVICE outputs; each may be NULL (not computed).  Markers whose
 * flag is not 1 keep their values.  marker_sxx and marker_sxy are updated in
 * place: the remaining grid stress change dsxx, dsxy at the marker is added. */
typedef struct {
    double *marker_exx, *marker_exy, *marker_pr, *marker_sr_ratio;
    double *marker_sxx, *marker_sxy;
} ebg_marker_gather_out;

/* Replaces, in one pass over the markers, the marker steps of
 * StokesContinuitySolver.execute_post_stokes_solver_steps!
 * (stokes_continuity_solver/StokesContinuitySolver.jl:272-279):
 *   MarkerStrainRate.calculate_marker_strain_rate!
 *     (stokes_continuity_solver/utils/strain_rate/MarkerStrainRate.jl:23-69),
 *   MarkerStrainRateRatio.calculate_marker_strain_rate_ratio!
 *     (stokes_continuity_solver/utils/strain_rate/MarkerStrainRateRatio.jl:71-188),
 *   MarkerPressure.calculate_marker_pressure!
 *     (stokes_continuity_solver/utils/MarkerPressure.jl:17-49),
 * then StressCorrection.apply_marker_stress_correction_for_remaining_grid_stress!
 *   (stokes_continuity_solver/subgrid_stress/stress_correction/StressCorrection.jl:27-83)
 *   when marker_sxx and marker_sxy are given (the step the reference runs
 *   alone when subgrid_diff_coef_stress is 0; the subgrid stress update it runs
 *   first otherwise is not part of this call),
 * with the normalized distances of
 * MarkerGridMapping.calc_normalized_distances_from_upper_left_basic_node!
 * (interpolation/mapping/MarkerGridMapping.jl:185-218) formed in-register.
 * marker_exx, marker_exy and marker_pr need exx, exy and pr1; marker_sr_ratio
 * needs exx, exy, sxx2, sxy2, dsxx, dsxy, eta, matid and mat_mu; marker_sxx
 * and marker_sxy need dsxx and dsxy.  FP64 in the
 * reference's operand order with correctly rounded division: bit-identical. */
int ebg_marker_strain_rate_and_pressure(ebg_ctx *ctx, const ebg_marker_gather_in *in,
                                        ebg_marker_gather_out *out);

/* ---- Stokes-continuity system (assembled-operator conventions) ----
 * Replaces the inputs StokesBuildData gathers (solver/build_system/utils/
 * BuildStructs.jl:37-177).  Fields are DEVICE pointers with the reference
 * shapes: etas0 (ynum,xnum), etan0 (ynum-1,xnum-1), rho1_vy (ynum,xnum+1; may be
 * NULL without interface stabilization), RX1 (ynum+1,xnum), RY1 (ynum,xnum+1),
 * RC1 (ynum-1,xnum-1).  The b* arrays are HOST pointers, column-major
 * (xnum+1)x2 for the top and bottom arrays and (ynum+1)x2 for left and right
 * (v_ghost = b[:,1] + b[:,2]*v_inner); bintern_zone (Int64[8]) and
 * bintern_velocity (Float64[8]) are HOST pointers or NULL.  Fields are copied. */
int ebg_stokes_set_system(ebg_ctx *ctx, const double *etas0, const double *etan0,
                          const double *rho1_vy, const double *RX1, const double *RY1,
                          const double *RC1, const double *btopx, const double *btopy,
                          const double *bbottomx, const double *bbottomy, const double *bleftx,
                          const double *blefty, const double *brightx, const double *brighty,
                          const int64_t *bintern_zone, const double *bintern_velocity,
                          int pressure_bc_mode, double pressure_bc, double pscale,
                          double timestep, double gravity_y, int iuse_interface_stabilization);

/* L*S for S in the reference's unknown order (per cell [vx, vy, p'], cells
 * column-major): bit-identical to summing the reference's COO triplets of
 * StokesBuildManager.build_system_of_equations (StokesBuildManager.jl:661-706). */
int ebg_stokes_apply(ebg_ctx *ctx, const double *S, double *LS);
/* The right-hand side R the reference's build produces (same order). */
int ebg_stokes_rhs(ebg_ctx *ctx, double *R);
/* K11: |L*S - R| (StokesResiduals.stokes_calc_nonlinear_system_residual_optimized!,
 * solver/residuals/StokesResiduals.jl:63-82). */
int ebg_stokes_residual(ebg_ctx *ctx, const double *S, double *resnl);
/* K10: SystemSolver.reload_solu_vel! + apply_vel_bc! (SystemSolver.jl:188-271):
 * vx1 (ynum+1,xnum), vy1 (ynum,xnum+1), pr1 = S[p']*pscale (ynum-1,xnum-1). */
int ebg_stokes_unpack(ebg_ctx *ctx, const double *S, double *vx1, double *vy1, double *pr1);
/* StokesResiduals.compute_stokes_residuals! (solver/residuals/StokesResiduals.jl:94-137)
 * with XStokes/YStokes/Continuity.calculate_residual!: per-equation residuals
 * of (vx1, vy1, pr1) on the vx, vy and pressure grids for the current system
 * (zone nodes and unwritten entries are 0). */
int ebg_compute_stokes_residuals(ebg_ctx *ctx, const double *vx1, const double *vy1, const double *pr1,
                                 double *resx1, double *resy1, double *resc1);
/* MathTools.get_L2_norms / get_inf_norms (src/utils/MathTools.jl:64-84) as used by
 * StokesNorms.calculate_stokes_norms! (solver/norms/StokesNorms.jl:8-160) on the
 * concatenation of nseg (1..4) device segments: out = {||v_old - v||_2,
 * relative L2 (1e38 if ||v|| = 0), ||v_old - v||_inf, relative inf}; v_old NULL
 * (or a NULL entry) means zeros, so out[0] is then ||v||_2.  n, v, v_old are host
 * arrays of nseg entries. */
int ebg_change_norms(ebg_ctx *ctx, int nseg, const int64_t *n, const double *const *v,
                     const double *const *v_old, double *out);
/* TimeStep update (time_step/update/Update.jl:43-98, 198-218): the step
 * limited by marker_cell_displ_max * x/ystpavg / max|vx1|, max|vy1|. */
int ebg_calculate_new_timestep(ebg_ctx *ctx, const double *vx1, const double *vy1, double timestep_old,
                               double marker_cell_displ_max, double *timestep_new, int *updated);

typedef struct {
    int precision;        /* 1: FP32 V-cycle preconditioner (mixed), 0: FP64 (plain) */
    int restart;          /* FGMRES restart length (2-D 1..254, 3-D 1..240), default 30; the 2-D
                             host wrappers use 100: the Krylov basis (2 m + 1 vectors, 7.7 GB at
                             2001 x 801) holds the plastic systems without a restart */
    int max_iterations;   /* total FGMRES iterations */
    double rtol;          /* stop at ||R - L*S||_2 / ||R||_2 <= rtol (FP64) */
    int nlevels;          /* multigrid levels, 0 = reference rule (LevelManager.jl:441-457) */
    int finest_sweeps;    /* smoothing sweeps on the finest level */
    int sweep_multiplier; /* sweeps multiply by this per coarser level ... */
    int sweep_cap;        /* ... capped at this (MultigridDataManager.jl:223-241) */
    double theta_s;       /* velocity relaxation (relax_stokes) */
    double theta_c;       /* pressure relaxation (relax_continuity) */
    int use_initial_guess;/* 1: start from S, 0: start from zero */
    int learned;          /* learned correction (needs ebg_cnn_create): 0 off; 1 use z + N where it
                             lowers the FP64 residual, else the V-cycle z; 2 always z + N;
                             3 z + N as an extra flexible Krylov direction after each V-cycle */
    double bc_d[4];       /* preconditioner ghost mirror factors top, bottom (vx), left, right (vy):
                             +1 free slip, -1 no slip */
    int smoother;         /* 1: LDS-tiled smoother (default), 0: global phased sweeps in the
                             reference multigrid's order */
    int local_sweeps;     /* tiled smoother: sweeps per tile launch between halo refreshes */
    int tile;             /* tiled smoother: tile edge (nodes) on levels that exceed one LDS tile */
    int flexible;         /* 1: flexible GMRES, storing every preconditioned vector (default);
                             0 (3-D solve only): right-preconditioned GMRES storing only the
                             basis, the V-cycle applied once more per restart (half the memory,
                             so twice the restart length fits) */
    int coarse_unknowns;  /* 2-D solve: > 0 (default 6000) ends the preconditioner hierarchy at
                             the first level below the finest with at most this many unknowns
                             and solves that level exactly (dense inverse of its operator, built
                             once per system with rocSOLVER); 0: smooth every level (no
                             reference counterpart) */
    int additive_tiles;   /* 2-D tiled smoother on the levels below the finest: 1 (default)
                             updates every tile in one launch from the previous iterate (block
                             Jacobi between tiles, double-buffered, a quarter of the launches);
                             0 launches the four tile colours in turn (block Gauss-Seidel) */
    int learned_patience; /* learned 1: after this many consecutive rejections by the residual
                             judge the solve falls back to the plain V-cycle for its remaining
                             iterations (default 4; 0: judge every iteration).  Learned 1 and 2
                             use the network only in the first restart cycle; a solve that
                             restarts finishes on the plain V-cycle */
} ebg_solve_opts;

typedef struct {
    int iterations;
    int restarts;
    double relres;         /* final ||R - L*S|| / ||R|| (FP64, recomputed) */
    double relres_initial;
    double seconds;
    int status;
    int learned_accepted; /* learned 1: corrections kept by the residual judge;
                             learned 3: learned Krylov directions */
    int learned_evaluated;/* learned 1: corrections the judge compared (the network stops being
                             evaluated after opts.learned_patience rejections in a row) */
} ebg_solve_stats;

int ebg_stokes_solve_opts_default(ebg_solve_opts *opts);
/* on = 1 (default): the tiled V-cycle of the 2-D preconditioner is captured
 * once into a HIP graph (per hierarchy, precision and smoother settings) and
 * replayed for each preconditioner application; on = 0 launches its kernels
 * one by one.  Identical results.  No counterpart in the reference (a launch
 * strategy of this GPU solver). */
int ebg_stokes_set_graphs(ebg_ctx *ctx, int on);

/* Replaces SystemSolver.solve_system_optimized / ParallelSolver.parallel_direct_solver
 * (solver/system_solver/SystemSolver.jl:67-91, parallel_solver/ParallelSolver.jl:52-126).
 * S (device, reference order, length 3*(xnum-1)*(ynum-1)) receives the solution.
 * Returns EBG_OK, EBG_NOT_CONVERGED (S holds the last iterate) or an error. */
int ebg_stokes_solve(ebg_ctx *ctx, double *S, const ebg_solve_opts *opts, ebg_solve_stats *stats);
/* The same solve with a caller right-hand side R (reference order, device)
 * instead of the system's; no reference counterpart (used to build training
 * samples e = L^-1 v for the learned correction). */
int ebg_stokes_solve_rhs(ebg_ctx *ctx, const double *R, double *S, const ebg_solve_opts *opts,
                         ebg_solve_stats *stats);

/* One application of the solver's multigrid preconditioner z = M r (reference
 * order, device).  No reference counterpart: exposes the preconditioner the
 * solve uses, for testing and for generating training samples. */
int ebg_stokes_precondition(ebg_ctx *ctx, const double *r, double *z, const ebg_solve_opts *opts);

/* ---- learned preconditioner correction (no reference counterpart) ----
 * A fully convolutional network on the cell grid (BF16 weights and
 * activations, FP32 accumulation) maps (log10 viscosity, r, z = M r) to a
 * correction of z; with opts.learned the solve uses z + correction wherever
 * it lowers the FP64 residual ||r - L z||, else the plain V-cycle output.
 * Images are [batch][H][W][channels] FP32 on the device with H = xnum-1 and
 * W = ynum-1 (pixel h*W + w = reference cell index). */
int ebg_cnn_create(ebg_ctx *ctx, uint64_t seed);
int ebg_cnn_num_params(ebg_ctx *ctx, int64_t *n);
int ebg_cnn_get_params(ebg_ctx *ctx, float *host);
int ebg_cnn_set_params(ebg_ctx *ctx, const float *host);
/* out [B][H][W][3] = network(in [B][H][W][7]) */
int ebg_cnn_forward(ebg_ctx *ctx, int B, int H, int W, const float *in, float *out);
/* mean squared error against target [B][H][W][3]; loss_and_grad also leaves
 * the parameter gradients for ebg_cnn_adamw_step and copies them to grad_host
 * when it is not NULL */
int ebg_cnn_loss(ebg_ctx *ctx, int B, int H, int W, const float *in, const float *target, double *loss);
int ebg_cnn_loss_and_grad(ebg_ctx *ctx, int B, int H, int W, const float *in, const float *target,
                          double *loss, float *grad_host);
int ebg_cnn_adamw_step(ebg_ctx *ctx, float lr, float beta1, float beta2, float eps, float weight_decay);
/* Training sample of the current Stokes system from an error vector e
 * (reference order, device): r = L e, z = M r, in_img = (log10 etan, r/s_r,
 * z/s_z), tgt_img = (e - z)/s_z with s the per-component RMS of r and z. */
int ebg_cnn_make_sample(ebg_ctx *ctx, const double *e, const ebg_solve_opts *opts, float *in_img,
                        float *tgt_img);

/* ---- the reference's experimental 2-D multigrid (FP64) ----
 * Replaces multigrid/multigrid_solver/MultigridSolver.jl:113-135 and the
 * level, smoother, restriction and prolongation modules it drives. */
enum {
    EBG_MG_ETAS = 0, EBG_MG_ETAN = 1, EBG_MG_VX = 2, EBG_MG_VY = 3, EBG_MG_PR = 4,
    EBG_MG_RX = 5, EBG_MG_RY = 6, EBG_MG_RC = 7, EBG_MG_RESX = 8, EBG_MG_RESY = 9,
    EBG_MG_RESC = 10
};
/* Level hierarchy with uniform grids (LevelManager.jl:441-519); nlevels 0 =
 * reference rule.  levels_out may be NULL. */
int ebg_mg2d_create(ebg_ctx *ctx, int64_t xnum, int64_t ynum, double xsize, double ysize,
                    int nlevels, int *levels_out);
int ebg_mg2d_level_shape(ebg_ctx *ctx, int level, int64_t *xnum, int64_t *ynum);
/* Copy one level field (EBG_MG_*) to dst or from src (device pointers, one NULL). */
int ebg_mg2d_copy_field(ebg_ctx *ctx, int level, int which, double *dst, const double *src);
/* LevelInterpolation.interpolate_viscosity_to_coarser_levels!
 * (multigrid/interpolation/ViscosityRestriction.jl:68-298). */
int ebg_mg2d_restrict_viscosity(ebg_ctx *ctx);
/* Smoother.stokes_continuity2d_viscous_smoother! (multigrid/smoother/Smoother.jl:66-99). */
int ebg_mg2d_smoother(ebg_ctx *ctx, int level, int nsweeps, double theta_s, double theta_c,
                      double prnorm, const double *bc_d4);
/* The same smoother on LDS tiles (the preconditioner's smoother).  tile 0:
 * the whole level is one tile and the result equals ebg_mg2d_smoother bit for
 * bit (error if the level exceeds the LDS limit); tile 4..32: tiles of that
 * edge in four colours, local_sweeps sweeps per launch with the halo held at
 * its launch-start values (a block Gauss-Seidel / Schwarz smoother). */
int ebg_mg2d_smoother_tiled(ebg_ctx *ctx, int level, int nsweeps, int local_sweeps, int tile,
                            double theta_s, double theta_c, double prnorm, const double *bc_d4);
/* ncycles V-cycles (vcycle/MultigridVCycle.jl:197-267 without viscosity
 * continuation): SmoothAndRestrict.jl:44-71 then SmoothAndProlongate.jl:60-115. */
int ebg_mg2d_vcycles(ebg_ctx *ctx, int ncycles, int finest_sweeps, int multiplier, int cap,
                     double theta_s, double theta_c, double relax_v, double relax_p,
                     double prnorm, const double *bc_d4);
/* Exact solve of one level's equations for its right-hand sides RX, RY, RC:
 * the level's operator (the smoother's residual arithmetic and ghost rules
 * with mirror factors bc_d4, host, NULL = free slip) with the continuity row
 * of cell (1,1) replaced by p(1,1) = 0, assembled densely and inverted; vx,
 * vy, pr receive the solution and the ghost pass is applied.  Levels up to
 * 20000 unknowns.  No reference counterpart: it is the coarsest-level solve of
 * the Stokes preconditioner (ebg_solve_opts.coarse_unknowns), exposed on the
 * reference-order hierarchy so it can be checked against the reference's
 * residuals. */
int ebg_mg2d_direct_solve(ebg_ctx *ctx, int level, const double *bc_d4);
/* Parameters of the reference's stand-alone V-cycle driver: VcycleParameters,
 * ViscosityScalingParameters, RelaxationParameters and ResidualScalingFactors
 * (multigrid_data/structures/MultigridStructures.jl:8-90). */
typedef struct ebg_mg_run_params {
    int nvcycles;                /* vcycle.nvcycles */
    int nvcycles_viscosity_jump; /* V-cycles between viscosity jumps (15) */
    int finest_sweeps;           /* smoothing iterations on the finest level */
    int sweep_multiplier;        /* per-level growth of the sweep count */
    int sweep_cap;               /* largest sweep count on any level */
    double nviscosity_jumps;     /* number of viscosity jumps (12.0) */
    double convergence_criterion;
    double resmax;               /* residual that triggers an early jump (1e-6) */
    double stokesscale;          /* ScalingFactors.jl:3-18 */
    double continscale;
    double theta_s, theta_c;     /* relax_stokes, relax_continuity */
    double relax_v, relax_p;     /* relax_velocity, relax_pressure */
    double pressure_bc;          /* pressure of cell (1,1,1) on the finest level */
} ebg_mg_run_params;
/* MultigridSolver.run_multigrid_solver for MultigridData2d with
 * use_multimulti = false (MultigridSolver.jl:113-135): the finest-level
 * viscosity range (MultigridDataManager.jl:321-373), restriction to every
 * level, saved originals, then V-cycles while the viscosity contrast rises
 * from 1 to the full range, one jump every nvcycles_viscosity_jump cycles
 * (ViscosityScaling.jl:26-39, 62-79, 112-131; resmax is not used in 2-D),
 * recording the mean principal residuals (CalculateMeanResiduals.jl:130-172)
 * and stopping on the reference's test (MultigridVCycle.jl:197-267, 414-434,
 * 454-468).  bc_d4 as for ebg_mg2d_vcycles.  history (host, may be NULL)
 * receives 3 doubles per cycle run: log10 RMS of the x-, y-Stokes and
 * continuity residuals. */
int ebg_mg2d_run(ebg_ctx *ctx, const ebg_mg_run_params *params, const double *bc_d4, int *cycles_done,
                 int *converged, double *history);

/* MultigridSolver.run_multigrid_solver for MultigridData2d with
 * use_multimulti = true (MultigridSolver.jl:113-135), the driver of the
 * reference's test/multigrid/run_stokes_sinker_2d_multimulti.jl.  The caller
 * fills level 1 as for ebg_mg2d_run; those solution and right-hand-side
 * arrays become the global level 0 (MultigridDataManager.jl:376-421), level 1
 * then solves for corrections whose right-hand side is level 0's residual
 * (Smoother.jl:117-150).  After each V-cycle the maximum viscosity grows by
 * its factor every nvcycles_viscosity_jump cycles; at its second overstep of
 * the full maximum level 0 takes level 1's correction, level 1's right-hand
 * side becomes level 0's new residual, whose mean scaled residuals are the
 * global residuals, level 1's solution is zeroed and the ramp restarts from
 * constant viscosity (ViscosityScaling.jl:170-205, 207-337).  The run stops
 * when the last global residuals fall below convergence_criterion, two
 * global updates or more existing (MultigridVCycle.jl:424-431, 488-504).  On
 * return level 1's vx, vy, pr hold level 0's (global) solution.  history as
 * for ebg_mg2d_run; global_history (host, may be NULL) receives 3 doubles
 * per global update, at most global_capacity updates; global_updates the
 * number of updates. */
int ebg_mg2d_run_multimulti(ebg_ctx *ctx, const ebg_mg_run_params *params, const double *bc_d4, int *cycles_done,
                            int *converged, double *history, int *global_updates, double *global_history,
                            int global_capacity);

/* ---- the reference's experimental 3-D multigrid (FP64) ----
 * Replaces multigrid/multigrid_solver/MultigridSolver.jl:113-135 for
 * MultigridData3d and the level, smoother, restriction and prolongation
 * modules it drives (free-slip walls, the only boundary type the 3-D solver
 * uses).  Arrays are column-major (y fastest, then x, then z) with the
 * reference's LevelData shapes: vx (ynum+1, xnum, znum+1), vy (ynum, xnum+1,
 * znum+1), vz (ynum+1, xnum+1, znum), pr/etan (ynum-1, xnum-1, znum-1),
 * etaxy (ynum, xnum, znum-1), etaxz (ynum-1, xnum, znum), etayz (ynum,
 * xnum-1, znum). */
enum {
    EBG_MG3_ETAN = 0, EBG_MG3_ETAXY = 1, EBG_MG3_ETAXZ = 2, EBG_MG3_ETAYZ = 3,
    EBG_MG3_VX = 4, EBG_MG3_VY = 5, EBG_MG3_VZ = 6, EBG_MG3_PR = 7,
    EBG_MG3_RX = 8, EBG_MG3_RY = 9, EBG_MG3_RZ = 10, EBG_MG3_RC = 11,
    EBG_MG3_RESX = 12, EBG_MG3_RESY = 13, EBG_MG3_RESZ = 14, EBG_MG3_RESC = 15
};
/* Level hierarchy with uniform grids (LevelManager.make_level_vector_3d,
 * LevelManager.jl:411-435, 459-482); nlevels 0 = reference rule. */
int ebg_mg3d_create(ebg_ctx *ctx, int64_t xnum, int64_t ynum, int64_t znum, double xsize, double ysize,
                    double zsize, int nlevels, int *levels_out);
int ebg_mg3d_level_shape(ebg_ctx *ctx, int level, int64_t *xnum, int64_t *ynum, int64_t *znum);
/* Levels of at most `nodes` nodes ((ynum+1)(xnum+1)(znum+1), default 2000)
 * run each smoother call as one single-workgroup launch; results are
 * identical either way (no reference counterpart: a launch-overhead knob). */
int ebg_mg3d_set_fused_limit(ebg_ctx *ctx, int64_t nodes);
/* Copy one level field (EBG_MG3_*) to dst or from src (device pointers, one NULL). */
int ebg_mg3d_copy_field(ebg_ctx *ctx, int level, int which, double *dst, const double *src);
/* ViscosityRestriction.viscosity_restriction3d! on every level
 * (multigrid/interpolation/ViscosityRestriction.jl:55-230). */
int ebg_mg3d_restrict_viscosity(ebg_ctx *ctx);
/* BoundaryConditions.set_boundary_conditions3d! (BoundaryConditions.jl:92-174). */
int ebg_mg3d_set_bc(ebg_ctx *ctx, int level);
/* Smoother.stokes_continuity3d_viscous_smoother! (multigrid/smoother/Smoother.jl:50-64):
 * nsweeps x [ghost pass; red-black sweep], pressure normalisation, residuals. */
int ebg_mg3d_smoother(ebg_ctx *ctx, int level, int nsweeps, double theta_s, double theta_c, double prnorm);
/* Restriction.restrict_stokes3d_residuals! (restriction/Restriction.jl:64-376):
 * residuals of `level` to the right-hand sides of level+1. */
int ebg_mg3d_restrict(ebg_ctx *ctx, int level);
/* Prolongation.prolongate_stokes3d_solution plus the correction
 * (Prolongation.jl:60-140, SmoothAndProlongate.jl:39-56): level's solution
 * added to level-1. */
int ebg_mg3d_prolong(ebg_ctx *ctx, int level, double relax_v, double relax_p);
/* ncycles V-cycles (vcycle/MultigridVCycle.jl:79-135 without viscosity
 * continuation): SmoothAndRestrict.jl:15-46 then SmoothAndProlongate.jl:14-60. */
int ebg_mg3d_vcycles(ebg_ctx *ctx, int ncycles, int finest_sweeps, int multiplier, int cap,
                     double theta_s, double theta_c, double relax_v, double relax_p, double prnorm);
/* MultigridSolver.run_multigrid_solver for MultigridData3d with
 * use_multimulti = false (MultigridSolver.jl:113-135): takes the range of the
 * finest-level viscosity (MultigridDataManager.jl:305-373), restricts it to
 * every level, keeps the originals, then runs V-cycles while raising the
 * viscosity contrast from 1 to the full range in nviscosity_jumps steps
 * (ViscosityScaling.jl:11-24, 41-60, 81-110), recording the mean principal residuals
 * (CalculateMeanResiduals.jl:14-70) and stopping on the reference's test
 * (MultigridVCycle.jl:79-135, 390-412, 436-452).  The level-0 viscosity arrays hold the
 * last scaled viscosity on return, as the reference's do.  history (host, may
 * be NULL) receives 4 doubles per cycle run: log10 RMS of the x-, y-, z-Stokes
 * and continuity residuals. */
int ebg_mg3d_run(ebg_ctx *ctx, const ebg_mg_run_params *params, int *cycles_done, int *converged,
                 double *history);

/* MultigridSolver.run_multigrid_solver for MultigridData3d with
 * use_multimulti = true: the 3-D form of ebg_mg2d_run_multimulti
 * (ViscosityScaling.jl:133-168, 217-265, Smoother.jl:101-141, the
 * global-residual test of MultigridVCycle.jl:400-411, 470-486); resmax is not
 * used.  On return level 1's vx, vy, vz, pr hold level 0's (global) solution;
 * global_history receives 4 doubles per global update, at most
 * global_capacity updates. */
int ebg_mg3d_run_multimulti(ebg_ctx *ctx, const ebg_mg_run_params *params, int *cycles_done, int *converged,
                            double *history, int *global_updates, double *global_history, int global_capacity);

/* ---- 3-D Stokes solve on the 3-D multigrid's system ----
 * No assembled 3-D system exists in the reference: its 3-D Stokes path is the
 * multigrid of MultigridSolver.run_multigrid_solver (MultigridSolver.jl:113-135)
 * iterating V-cycles on the residual equations of Residuals.jl:10-560.  These
 * entry points solve that same system with FGMRES in FP64, preconditioned by
 * one V-cycle in FP32 (opts.precision 1) or FP64 (0).
 * Vectors are [vx | vy | vz | pr] with the level-0 shapes above
 * (ebg_stokes3d_vector_length).  Unknowns are the interior velocity nodes and
 * all pressure cells; the continuity equation of cell (1,1,1) is replaced by
 * the gauge p(1,1,1) = prnorm (PressureUpdate.jl:5-21) scaled by
 * 1/etan(1,1,1). */
/* Viscosity arrays are device FP64 with the shapes above; the grid is the
 * uniform level grid of LevelManager.initialize_grid! (LevelManager.jl:354-366). */
int ebg_stokes3d_set_system(ebg_ctx *ctx, int64_t xnum, int64_t ynum, int64_t znum, double xsize,
                            double ysize, double zsize, const double *etan, const double *etaxy,
                            const double *etaxz, const double *etayz);
int ebg_stokes3d_vector_length(ebg_ctx *ctx, int64_t *n);
/* y = A x; the ghost slots of x are overwritten with its free-slip ghost
 * values first, the ghost slots of y are zero. */
int ebg_stokes3d_apply(ebg_ctx *ctx, double *x, double *y);
/* z = M r: one V-cycle with right-hand side r (ghost slots of r ignored). */
int ebg_stokes3d_precondition(ebg_ctx *ctx, const double *r, double *z, const ebg_solve_opts *opts);
/* Solve A x = b where b is rhs = [RX | RY | RZ | RC] on the unknowns (ghost
 * entries ignored) with the gauge entry prnorm/etan(1,1,1).  Uses opts.precision,
 * restart (1..240), max_iterations, rtol, nlevels, finest_sweeps, sweep_multiplier,
 * sweep_cap, theta_s, theta_c, use_initial_guess; the 2-D-only fields are
 * ignored.  x (device, vector layout) receives the solution with its ghost
 * nodes set. */
int ebg_stokes3d_solve(ebg_ctx *ctx, const double *rhs, double prnorm, double *x,
                       const ebg_solve_opts *opts, ebg_solve_stats *stats);

/* ---- heat equation (basic grid) ----
 * Thermal boundary conditions, DEVICE arrays in the reference's (n, 2)
 * column-major layout (bcs.jl:8-27): btopt, bbottomt (xnum, 2),
 * bleftt, brightt (ynum, 2); column 1 the value, column 2 the factor of the
 * inner neighbour: tk[1,j] = btopt[j,1] + btopt[j,2]*tk[2,j]. */
typedef struct {
    const double *btopt, *bbottomt, *bleftt, *brightt;
} ebg_thermal_bc;

/* Replaces ApplyThermalBC.apply_thermal_bc_along_boundaries!
 * (interpolation/utils/heat_funcs/ApplyThermalBC.jl:20-44): the top and
 * bottom rows of the inner columns, then both side columns over every row,
 * on tk1 (ynum,xnum) in place.  Bit-identical. */
int ebg_apply_thermal_bc(ebg_ctx *ctx, const ebg_thermal_bc *bc, double *tk1);

/* Inputs of the heat right-hand side, DEVICE arrays: hr1, ha1, rho1, sxy2,
 * etas1 (ynum,xnum); vx1 (ynum+1,xnum); vy1 (ynum,xnum+1); sxx2, etan1
 * (ynum-1,xnum-1).  Only hr1 is needed with both options off. */
typedef struct {
    const double *hr1, *ha1, *rho1, *vx1, *vy1, *sxy2, *etas1, *sxx2, *etan1;
    double gravity_x, gravity_y;
    int32_t iuse_adiabatic_heating, iuse_shear_heating;
} ebg_heat_rhs_in;

/* Replaces HeatSolver.calculate_heat_rhs_terms_grid! (heat_solver/HeatSolver.jl:107-118;
 * HeatRhs.initialize_rhs_terms_array_for_heat!, add_adiabatic_terms_to_rhs_grid!,
 * add_friction_terms_to_rhs_grid!, rhs/HeatRhs.jl:5-79): RT1 (ynum,xnum).
 * Bit-identical, the squares evaluated as products. */
int ebg_heat_rhs(ebg_ctx *ctx, const ebg_heat_rhs_in *in, double *RT1);

/* One heat system (HeatBuildStructs.HeatBuildData, build_system/utils/BuildStructs.jl:70-96):
 * DEVICE RT1, tk0 (old temperature), rhocp1, kt1 (ynum,xnum), the boundary
 * conditions and the heat-loop time step (s). */
typedef struct {
    const double *RT1, *tk0, *rhocp1, *kt1;
    ebg_thermal_bc bc;
    double timestep;
} ebg_heat_system;

/* The rows of HeatBuildManager.build_system_of_equations
 * (build_system/HeatBuildManager.jl:47-83, InternalNodes.jl:19-197,
 * BoundaryNodes.jl:16-251) without assembling them: LT = L*T (optional) and
 * the right-hand side R (optional), both (ynum,xnum) in the reference's
 * unknown order itk = (j-1)*ynum + i.  Each row's triplets are accumulated
 * from 0.0 in the order the reference emits them, so L*T equals the COO
 * product bit for bit. */
int ebg_heat_system_apply(ebg_ctx *ctx, const ebg_heat_system *sys, const double *T, double *LT, double *R);

typedef struct {
    int32_t restart;        /* FGMRES restart length, default 40 (at most 254) */
    int32_t max_iterations; /* default 400 */
    double rtol;            /* ||R - L*T|| / ||R|| target, default 1e-12 */
    int32_t pre_sweeps, post_sweeps;  /* red-black Gauss-Seidel sweeps per level, default 2 and 2 */
} ebg_heat_solve_opts;

/* Replaces SystemSolver.discretize_and_solve_conductive_heat_equation
 * (heat_solver/solver/SystemSolver.jl:32-66): build, parallel_direct_solver
 * (the MUMPS seam) and reload_solu_temp.  tk2 (ynum,xnum) receives the
 * solution of the system above by flexible GMRES (FP64) preconditioned with a
 * geometric multigrid V-cycle, to ||R - L*T||/||R|| <= rtol.  opts and stats
 * may be NULL.  Returns EBG_OK, EBG_NOT_CONVERGED (tk2 holds the last
 * iterate) or an error. */
int ebg_heat_solve(ebg_ctx *ctx, const ebg_heat_system *sys, double *tk2, const ebg_heat_solve_opts *opts,
                   ebg_solve_stats *stats);

/* Replaces HeatResiduals.compute_residuals_temp! (solver/residuals/HeatResiduals.jl:12-162):
 * rest (ynum,xnum), 0 on the boundary.  Bit-identical. */
int ebg_heat_residuals(ebg_ctx *ctx, const ebg_heat_system *sys, const double *tk2, double *rest);

typedef struct {
    const double *RT1, *rhocp1, *kt1;
    ebg_thermal_bc bc;
    double timestep;          /* model time step (s) */
    double max_temp_change;   /* K */
    int32_t max_substeps;     /* guard against a loop that cannot finish, 0: 100000 */
    const double *tk1;        /* optional: the marker-interpolated temperature, for dtk1 */
} ebg_heat_loop_in;

typedef struct {
    double *tk0;   /* in: the starting temperature (tk1), out: the final solution */
    double *tk2;   /* out: the final solution */
    double *rest;  /* out: residuals of the last solve */
    double *dtk1;  /* optional out (needs tk1): tk2 - tk1, HeatSolver.calculate_grid_temp_change!
                      (HeatSolver.jl:158-163) */
} ebg_heat_loop_out;

typedef struct {
    int32_t substeps, solves, iterations, not_converged;
    double last_relres, timestep_last, timestep_sum;
} ebg_heat_loop_stats;

/* Replaces HeatLoop.run_conductive_heat_solver_loop! (heat_solver/heat_loop/HeatLoop.jl:12-78):
 * solves with the whole time step, and when the largest temperature change
 * exceeds max_temp_change solves again with the step scaled by
 * max_temp_change / change; repeats until the sub-steps add up to the time
 * step, each starting from the previous solution.  Returns EBG_OK,
 * EBG_NOT_CONVERGED (some solve missed rtol) or an error. */
int ebg_run_conductive_heat_solver_loop(ebg_ctx *ctx, const ebg_heat_loop_in *in, ebg_heat_loop_out *out,
                                        const ebg_heat_solve_opts *opts, ebg_heat_loop_stats *stats);

/* ---- marker temperature and sub-grid diffusion ----
 * DEVICE marker arrays (length marknum): x, y, the stored 1-based upper-left
 * basic cells marker_xn/yn (Int32) and inside_flags (optional: NULL computes
 * them from the grid limits). */
typedef struct {
    int64_t marknum;
    const double *x, *y;
    const int8_t *inside_flags;
    const int32_t *marker_xn, *marker_yn;
} ebg_marker_cells;

/* mode 0 replaces GridTemperatureToMarkers.interpolate!
 * (interpolation/utils/GridTemperatureToMarkers.jl:15-35): marker_TK = grid at
 * the marker; mode 1 replaces MarkerTemperature.add_remaining_grid_change_to_marker_temperature!
 * (heat_solver/marker_temperature/MarkerTemperature.jl:18-45): marker_TK +=
 * grid at the marker (grid = dtk1).  Inside markers only; bit-identical. */
int ebg_grid_to_marker_temperature(ebg_ctx *ctx, const ebg_marker_cells *m, const double *grid, int mode,
                                   double *marker_TK);

/* Replaces MarkerTemperature.apply_sticky_temperature_correction!
 * (MarkerTemperature.jl:52-72): inside markers of the sticky-air or
 * sticky-water material take temperature_top (an id of -1 matches nothing). */
int ebg_apply_sticky_temperature_correction(ebg_ctx *ctx, const ebg_marker_cells *m, const int16_t *matid,
                                            int matid_sticky_air, int matid_sticky_water, double temperature_top,
                                            double *marker_TK);

typedef struct {
    ebg_marker_cells markers;
    const double *tk1, *kt1, *rhocp1;   /* (ynum,xnum) */
    double timestep, subgrid_diff_coef_temp;
} ebg_subgrid_heat_in;

/* Replaces SubgridDiffusion.subgrid_heat_diffusion_update! followed by
 * remove_subgrid_effects_from_grid_changes! (heat_solver/subgrid_heat/SubgridDiffusion.jl:52-185):
 * marker_TK += dT_sg (inside markers), dtkn = the inclusive average of dT_sg
 * (0 where no weight), dtk1 -= dtkn; wtnodes (optional out) receives the
 * weight sums.  Bit-identical except exp, from the device math library.
 * The per-marker dT_sg buffer is the library's own (one per context). */
int ebg_subgrid_heat_diffusion_update(ebg_ctx *ctx, const ebg_subgrid_heat_in *in, double *marker_TK,
                                      double *dtkn, double *dtk1, double *wtnodes);

typedef struct {
    ebg_marker_cells markers;
    const int16_t *matid;
    const double *mat_mu;      /* [nmat], indexed by matid - 1 */
    int64_t nmat;
    const double *eta;         /* marker viscoplastic viscosity */
    const double *sxy1;        /* (ynum,xnum) pre-solve grid stresses */
    const double *sxx1;        /* (ynum-1,xnum-1) */
    double timestep, subgrid_diff_coef_stress;
} ebg_subgrid_stress_in;

typedef struct {
    double *marker_sxy, *marker_sxx;   /* in/out */
    double *dsxy, *dsxx;               /* in: grid stress changes; out: the remaining changes */
    double *dsxyn, *dsxxn;             /* out: sub-grid changes on the grid, 0 where no weight */
    double *wtetas, *wtetan;           /* optional out: weight sums */
} ebg_subgrid_stress_out;

/* Replaces SubGridStress.Update.update_subgrid_stress!
 * (stokes_continuity_solver/subgrid_stress/update/Update.jl:55-71;
 * SubgridShearStress.jl:25-96, SubgridNormalStress.jl:24-89,
 * DiffusionTerm.jl:13-29, RemainingStress.jl:16-23), which runs before the
 * marker stress correction when subgrid_diff_coef_stress > 0
 * (StokesContinuitySolver.jl:365-381).  Bit-identical except exp. */
int ebg_update_subgrid_stress(ebg_ctx *ctx, const ebg_subgrid_stress_in *in, ebg_subgrid_stress_out *out);

/* ---- marker plasticity (the Picard loop over markers) ----
 * DEVICE marker arrays (length marknum): inside_flags, y, matid (1-based) with
 * mat_mu[nmat]; marker cohesion, sine of the friction angle (fric), strain
 * rates exx, exy, strain-rate ratio sr_ratio and pressure pr.  forecast: 0
 * ViscoelasticStressForecast, 1 PureElasticStressForecast (itype_plasticity).
 * Sticky-air, mobile-wall and plate-extension ids (-1: none) matter only with
 * no_yielding_in_mobile_wall. */
typedef struct {
    int64_t marknum;
    const int8_t *inside_flags;
    const double *y;
    const int16_t *matid;
    const double *mat_mu;
    int64_t nmat;
    const double *cohesion, *fric, *exx, *exy, *sr_ratio, *pr;
    int32_t forecast;
    int32_t ntimestep;
    double timestep, y_sealevel;
    double viscosity_min, viscosity_max, yield_stress_min, yield_stress_max;
    int32_t iuse_fluid_pressure_for_yield, no_yielding_in_mobile_wall;
    int32_t matid_sticky_air, matid_mobile_wall, matid_plate_extension;
} ebg_marker_plastic_in;

typedef struct {
    double *sxx, *sxy;   /* in/out (the viscoelastic forecast corrects them) */
    double *eta;         /* in/out: marker viscoplastic viscosity */
    float *pfailure;     /* in/out: 1 stays 1, else the yield flag */
} ebg_marker_plastic_out;

/* Replaces MarkerPlasticity.update_plastic_yielding!
 * (rheology/plastic_failure/markers/MarkerPlasticity.jl:34-43; ViscoelasticForecast.jl:14-127,
 * ElasticForecast.jl:14-110, YieldCheck.jl, Limits.jl), the marker step of the
 * Picard loop over markers (PlasticityLoopMarkers.jl:48-53).  Bit-identical
 * except asin/cos of the yield stress and the square roots. */
int ebg_update_marker_plastic_yielding(ebg_ctx *ctx, const ebg_marker_plastic_in *in, ebg_marker_plastic_out *out);

#ifdef __cplusplus
}
#endif
#endif
