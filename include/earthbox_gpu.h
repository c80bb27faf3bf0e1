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
/* Use an external hipStream_t (e.g. the caller's current stream); NULL restores
 * the context's own stream. */
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
 * gridy_b[1], gridy_b[ynum] (staggered_grid/utils/Dimensions.jl:6-19). */
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
 * Arithmetic is FP64 with the reference's evaluation order (no contraction), so
 * velocities, spins and positions are bit-identical to the reference loops. */
int ebg_rk4_advect(ebg_ctx *ctx, int64_t marknum, int order, double timestep,
                   double *marker_x, double *marker_y,
                   const int32_t *marker_xn, const int32_t *marker_yn,
                   const int8_t *inside_flags,
                   const double *vx1, const double *vy1, const double *esp,
                   double *marker_vx, double *marker_vy, double *marker_spin,
                   double *marker_sxx, double *marker_sxy, int update_location);

/* Replaces GridStrainRate.calculate_spin!
 * (stokes_continuity_solver/utils/strain_rate/GridStrainRate.jl:161-182). */
int ebg_calculate_spin(ebg_ctx *ctx, const double *vx1, const double *vy1, double *esp);

#ifdef __cplusplus
}
#endif
#endif
