/*
 * CPU oracle for the EarthBox.jl per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference Julia loops.  It exists so that tests, the smoke check and the CPU
 * baseline leg of the benchmark have something independent to compare the HIP
 * kernels against.  Nothing in the product library links, loads or calls it.
 *
 * Conventions (mirroring the reference, which is Julia):
 *   - 2-D arrays are column-major with the y (row) index i fastest, exactly as
 *     Julia stores Matrix{Float64}; an array with `ld` rows holds element
 *     (i, j) (1-based) at a[(j-1)*ld + (i-1)].
 *   - Marker cell indices (xn, yn) are 1-based Int32, as in the reference.
 *   - All floating point is IEEE binary64 evaluated left to right; the library
 *     is compiled with -ffp-contract=off so no fused multiply-adds appear
 *     (Julia does not contract a*b+c either).
 */
#ifndef EB_ORACLE_H
#define EB_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 1-based accessors */
#define EBO_M(a, ld, i, j) ((a)[((size_t)(j) - 1) * (size_t)(ld) + (size_t)((i) - 1)])
#define EBO_V(a, i) ((a)[(size_t)(i) - 1])

/* Uniform 2-D staggered grid, restating
 * staggered_grid/types_manager/uniform/core/Initialization.jl:27-43 (basic),
 * staggered_grid/utils/Spacing.jl:6-18,124-133 (spacings),
 * staggered_grid/utils/VxGridTools.jl:40-99, VyGridTools.jl:41-100,
 * PressureGridTools.jl:34-86 (staggered coordinates and spacings).
 * Output buffers must hold: gridx_b[xnum], gridy_b[ynum], xstp_b[xnum-1],
 * ystp_b[ynum-1], gridx_vy[xnum+1], xstp_vy[xnum], gridy_vx[ynum+1],
 * ystp_vx[ynum], gridx_pr[xnum-1], gridy_pr[ynum-1], xstp_pr[xnum-2],
 * ystp_pr[ynum-2]. */
void ebo_grid2d_uniform(int64_t xnum, int64_t ynum, double xsize, double ysize,
                        double *xstpavg, double *ystpavg,
                        double *gridx_b, double *gridy_b,
                        double *xstp_b, double *ystp_b,
                        double *gridx_vy, double *xstp_vy,
                        double *gridy_vx, double *ystp_vx,
                        double *gridx_pr, double *gridy_pr,
                        double *xstp_pr, double *ystp_pr);

/* searchsortedlast(grid, x) clamped to [1, n-1]
 * (interpolation/mapping/MarkerGridMapping.jl:78-86). */
int32_t ebo_index_of_left_node(double x, const double *grid, int64_t n);

/* Inside flags, src/utils/GridFuncs.jl:54-80: 1 if xmin<=x<=xmax and
 * ymin<=y<=ymax else -1. */
void ebo_inside_flags(int64_t marknum, const double *x, const double *y,
                      double xmin, double xmax, double ymin, double ymax,
                      int8_t *flags);

/* Upper-left basic-grid cell indices of markers,
 * MarkerGridMapping.jl:99-129 (basic) and :142-173 (vy grid). */
void ebo_cell_indices_basic(int64_t marknum, const double *x, const double *y,
                            const int8_t *flags, const double *gridx_b, int64_t xnum,
                            const double *gridy_b, int64_t ynum,
                            int32_t *xn, int32_t *yn);

/* RK4 velocity and spin at markers:
 * advection/velocity_and_spin/VelocityAndSpin.jl:45-129 with rk_update
 * (:131-154), interpolate_vx/vy/spin (:343-428), advect_point (:430-444),
 * RungeKutta.jl:9-64, GridToMarker.jl:190-226, MarkerGridMapping.jl:354-441.
 * vx1 is (ynum+1) x xnum, vy1 is ynum x (xnum+1), esp is ynum x xnum.
 * Outputs marker_vx/vy/spin; markers with flags != 1 get zeros. */
void ebo_rk4_velocity_and_spin(int64_t marknum, int order, double timestep,
                               int64_t xnum, int64_t ynum,
                               const double *x, const double *y,
                               const int32_t *xn, const int32_t *yn,
                               const int8_t *flags,
                               const double *gridx_b, const double *gridy_b,
                               const double *gridx_vy, const double *gridy_vx,
                               const double *xstp_b, const double *ystp_b,
                               const double *xstp_vy, const double *ystp_vx,
                               const double *vx1, const double *vy1, const double *esp,
                               double *marker_vx, double *marker_vy, double *marker_spin);

/* x += dt*vx, y += dt*vy for inside markers (location/MarkersLocation.jl:11-41). */
void ebo_set_threads(int n);
void ebo_update_marker_location(int64_t marknum, double timestep, const int8_t *flags,
                                const double *mvx, const double *mvy, double *x, double *y);

/* Stress rotation by spin*dt (stress_rotation/MarkersStressRotation.jl:15-74). */
void ebo_rotate_marker_stress(int64_t marknum, double timestep, const int8_t *flags,
                              const double *spin, double *sxx, double *sxy);

/* esp = 0.5*(dvy/dx - dvx/dy) on basic nodes
 * (stokes_continuity_solver/utils/strain_rate/GridStrainRate.jl:161-182). */
void ebo_calculate_spin(int64_t xnum, int64_t ynum, const double *xstp_vy,
                        const double *ystp_vx, const double *vx1, const double *vy1,
                        double *esp);

#ifdef __cplusplus
}
#endif
#endif
