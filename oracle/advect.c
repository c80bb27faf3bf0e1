/*
 * Oracle (test infrastructure only): grid coordinates, marker->cell indexing,
 * RK4 velocity/spin gather, marker location update, stress rotation and grid
 * spin, restated from the reference Julia.  Every function cites the reference
 * file:line it follows.  Build with -ffp-contract=off.
 */
#include "eb_oracle.h"
#include "eb_mg3d_oracle.h"
#include <math.h>
#include <omp.h>

/* Initialization.jl:27-43, Spacing.jl:6-18 and :124-133,
 * VxGridTools.jl:40-99 (vxgrid_y_or_z_coordinates_loop!, _spacing_loop!),
 * VyGridTools.jl:41-100, PressureGridTools.jl:34-86 (spacing_loop!,
 * coordinate_loop!). */
static void basic_axis(int64_t n, double step, double *g, double *stp)
{
    EBO_V(g, 1) = 0.0;
    for (int64_t i = 2; i <= n; ++i) EBO_V(g, i) = EBO_V(g, i - 1) + step;
    for (int64_t i = 1; i <= n - 1; ++i) EBO_V(stp, i) = EBO_V(g, i + 1) - EBO_V(g, i);
}

static void staggered_axis(int64_t n, const double *g, const double *stp,
                           double *gs, double *stps)
{
    /* coordinates: VxGridTools.jl vxgrid_y_or_z_coordinates_loop! */
    EBO_V(gs, 1) = EBO_V(g, 1) - EBO_V(stp, 1) / 2.0;
    EBO_V(gs, n + 1) = EBO_V(g, n) + EBO_V(stp, n - 1) / 2.0;
    for (int64_t i = 2; i <= n; ++i) EBO_V(gs, i) = (EBO_V(g, i) + EBO_V(g, i - 1)) / 2.0;
    /* spacing: vxgrid_y_or_z_spacing_loop! */
    EBO_V(stps, 1) = EBO_V(stp, 1);
    EBO_V(stps, n) = EBO_V(stp, n - 1);
    for (int64_t i = 2; i <= n - 1; ++i) EBO_V(stps, i) = (EBO_V(g, i + 1) - EBO_V(g, i - 1)) / 2.0;
}

static void pressure_axis(int64_t n, const double *g, double *gp, double *stpp)
{
    for (int64_t i = 1; i <= n - 2; ++i) EBO_V(stpp, i) = (EBO_V(g, i + 2) - EBO_V(g, i)) / 2.0;
    EBO_V(gp, 1) = (EBO_V(g, 1) + EBO_V(g, 2)) / 2.0;
    for (int64_t i = 2; i <= n - 1; ++i) EBO_V(gp, i) = EBO_V(gp, i - 1) + EBO_V(stpp, i - 1);
}

void ebo_grid2d_uniform(int64_t xnum, int64_t ynum, double xsize, double ysize,
                        double *xstpavg, double *ystpavg,
                        double *gridx_b, double *gridy_b,
                        double *xstp_b, double *ystp_b,
                        double *gridx_vy, double *xstp_vy,
                        double *gridy_vx, double *ystp_vx,
                        double *gridx_pr, double *gridy_pr,
                        double *xstp_pr, double *ystp_pr)
{
    *xstpavg = xsize / (double)(xnum - 1);
    *ystpavg = ysize / (double)(ynum - 1);
    basic_axis(xnum, *xstpavg, gridx_b, xstp_b);
    basic_axis(ynum, *ystpavg, gridy_b, ystp_b);
    staggered_axis(xnum, gridx_b, xstp_b, gridx_vy, xstp_vy);
    staggered_axis(ynum, gridy_b, ystp_b, gridy_vx, ystp_vx);
    pressure_axis(xnum, gridx_b, gridx_pr, xstp_pr);
    pressure_axis(ynum, gridy_b, gridy_pr, ystp_pr);
}

/* MarkerGridMapping.jl:78-86: clamp(searchsortedlast(grid, x), 1, n-1).
 * searchsortedlast returns the largest i with grid[i] <= x (0 if none); a NaN
 * compares greater than every grid value and lands on n, which the clamp maps
 * to n-1 exactly as the reference does. */
int32_t ebo_index_of_left_node(double x, const double *g, int64_t n)
{
    int64_t lo = 0, hi = n + 1;
    while (lo < hi - 1) {
        int64_t m = lo + (hi - lo) / 2;
        if (x < EBO_V(g, m)) hi = m;
        else lo = m;
    }
    if (lo < 1) lo = 1;
    if (lo > n - 1) lo = n - 1;
    return (int32_t)lo;
}

/* GridFuncs.jl:54-80 (get_marker_inside_flags / check_in_domain_optimized) */
void ebo_inside_flags(int64_t marknum, const double *x, const double *y,
                      double xmin, double xmax, double ymin, double ymax, int8_t *flags)
{
    for (int64_t m = 0; m < marknum; ++m) {
        double xm = x[m], ym = y[m];
        flags[m] = (xmin <= xm && xm <= xmax && ymin <= ym && ym <= ymax) ? 1 : -1;
    }
}

/* MarkerGridMapping.jl:99-129 */
void ebo_cell_indices_basic(int64_t marknum, const double *x, const double *y,
                            const int8_t *flags, const double *gridx_b, int64_t xnum,
                            const double *gridy_b, int64_t ynum, int32_t *xn, int32_t *yn)
{
    for (int64_t m = 0; m < marknum; ++m) {
        if (flags[m] == 1) {
            xn[m] = ebo_index_of_left_node(x[m], gridx_b, xnum);
            yn[m] = ebo_index_of_left_node(y[m], gridy_b, ynum);
        }
    }
}

/* GridToMarker.jl:190-226: weights then summation in the order UL, LL, UR, LR
 * starting from 0.0. */
static double marker_value(int32_t yn, int32_t xn, double dy, double dx,
                           const double *a, int64_t ld)
{
    double wul = (1.0 - dx) * (1.0 - dy);
    double wll = (1.0 - dx) * dy;
    double wur = dx * (1.0 - dy);
    double wlr = dx * dy;
    double v = 0.0;
    v += wul * EBO_M(a, ld, yn, xn);
    v += wll * EBO_M(a, ld, yn + 1, xn);
    v += wur * EBO_M(a, ld, yn, xn + 1);
    v += wlr * EBO_M(a, ld, yn + 1, xn + 1);
    return v;
}

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
                               double *marker_vx, double *marker_vy, double *marker_spin)
{
    /* markers are independent; threaded over markers as the reference's
     * Threads.@threads loop (VelocityAndSpin.jl:77), ebo_set_threads sets the count */
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < marknum; ++m) {
        if (flags[m] != 1) {          /* VelocityAndSpin.jl:120-126 */
            marker_vx[m] = 0.0;
            marker_vy[m] = 0.0;
            marker_spin[m] = 0.0;
            continue;
        }
        const double x0 = x[m], y0 = y[m];
        double xc = x0, yc = y0;
        double vxs[4], vys[4], sps[4];
        for (int k = 1; k <= 4; ++k) {
            int32_t xi, yi;
            if (k == 1) {             /* VelocityAndSpin.jl:164-168 */
                xi = xn[m];
                yi = yn[m];
            } else {
                xi = ebo_index_of_left_node(xc, gridx_b, xnum);
                yi = ebo_index_of_left_node(yc, gridy_b, ynum);
            }
            /* interpolate_vx (:343-371): upr_left_x_mapping_vx_grid and
             * upr_left_y_mapping_vx_grid (MarkerGridMapping.jl:378-409) */
            double dxv = (xc - EBO_V(gridx_b, xi)) / EBO_V(xstp_b, xi);
            int32_t iyv = yi;
            if (yc > EBO_V(gridy_vx, iyv + 1)) iyv += 1;
            if (iyv > ynum) iyv -= 1;
            double dyv = (yc - EBO_V(gridy_vx, iyv)) / EBO_V(ystp_vx, iyv);
            double vxm = marker_value(iyv, xi, dyv, dxv, vx1, ynum + 1);
            /* interpolate_vy (:373-400): MarkerGridMapping.jl:413-441 */
            int32_t jxv = xi;
            if (xc > EBO_V(gridx_vy, jxv + 1)) jxv += 1;
            if (jxv > xnum) jxv = (int32_t)xnum;
            double dxy = (xc - EBO_V(gridx_vy, jxv)) / EBO_V(xstp_vy, jxv);
            double dyy = (yc - EBO_V(gridy_b, yi)) / EBO_V(ystp_b, yi);
            double vym = marker_value(yi, jxv, dyy, dxy, vy1, ynum);
            /* interpolate_spin (:402-428) */
            double dxs = (xc - EBO_V(gridx_b, xi)) / EBO_V(xstp_b, xi);
            double dys = (yc - EBO_V(gridy_b, yi)) / EBO_V(ystp_b, yi);
            double spm = marker_value(yi, xi, dys, dxs, esp, ynum);
            vxs[k - 1] = vxm;
            vys[k - 1] = vym;
            sps[k - 1] = spm;
            /* advect_point (:430-444) -> RungeKutta.jl:9-47 */
            if (k < 3) {
                xc = x0 + timestep / 2.0 * vxm;
                yc = y0 + timestep / 2.0 * vym;
            } else if (k == 3) {
                xc = x0 + timestep * vxm;
                yc = y0 + timestep * vym;
            } else {
                xc = x0;
                yc = y0;
            }
        }
        if (order == 4) {             /* RungeKutta.jl:62-64 */
            marker_vx[m] = (vxs[0] + 2.0 * vxs[1] + 2.0 * vxs[2] + vxs[3]) / 6.0;
            marker_vy[m] = (vys[0] + 2.0 * vys[1] + 2.0 * vys[2] + vys[3]) / 6.0;
            marker_spin[m] = (sps[0] + 2.0 * sps[1] + 2.0 * sps[2] + sps[3]) / 6.0;
        } else {
            marker_vx[m] = vxs[0];
            marker_vy[m] = vys[0];
            marker_spin[m] = sps[0];
        }
    }
}

/* OpenMP thread count of the threaded marker loops (the CPU baseline runs
 * them on one core and on all cores). */
void ebo_set_threads(int n)
{
    omp_set_num_threads(n > 0 ? n : 1);
}

/* MarkersLocation.jl:11-41 */
void ebo_update_marker_location(int64_t marknum, double timestep, const int8_t *flags,
                                const double *mvx, const double *mvy, double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < marknum; ++m) {
        if (flags[m] == 1) {
            x[m] = x[m] + timestep * mvx[m];
            y[m] = y[m] + timestep * mvy[m];
        }
    }
}

/* MarkersStressRotation.jl:15-74 */
void ebo_rotate_marker_stress(int64_t marknum, double timestep, const int8_t *flags,
                              const double *spin, double *sxx, double *sxy)
{
    for (int64_t m = 0; m < marknum; ++m) {
        if (flags[m] != 1) continue;
        double s = spin[m];
        double t1 = 2.0 * s * timestep;
        double t2 = s * timestep;
        double s1 = sin(t1), c1 = cos(t1), s2 = sin(t2), c2 = cos(t2);
        double sxy_old = sxy[m], sxx_old = sxx[m];
        double sxy_new = sxx_old * s1 + sxy_old * c1;
        double sxx_new = sxx_old * (c2 * c2 - s2 * s2) - sxy_old * s1;
        sxy[m] = sxy_new;
        sxx[m] = sxx_new;
    }
}

/* GridStrainRate.jl:161-182 */
void ebo_calculate_spin(int64_t xnum, int64_t ynum, const double *xstp_vy,
                        const double *ystp_vx, const double *vx1, const double *vy1,
                        double *esp)
{
    for (int64_t j = 1; j <= xnum; ++j)
        for (int64_t i = 1; i <= ynum; ++i)
            EBO_M(esp, ynum, i, j) = 0.5 * (
                (EBO_M(vy1, ynum, i, j + 1) - EBO_M(vy1, ynum, i, j)) / EBO_V(xstp_vy, j) -
                (EBO_M(vx1, ynum + 1, i + 1, j) - EBO_M(vx1, ynum + 1, i, j)) / EBO_V(ystp_vx, i));
}

/* One axis of a uniform level grid (Initialization.jl:47-74 for the basic
 * nodes, VxGridTools.jl:40-99 for the staggered nodes with ghosts,
 * PressureGridTools.jl:34-86 for the pressure nodes, SxyGridTools.jl:30-40 for
 * the shear-centre nodes).  Buffers: g_b[n], stp_b[n-1], g_v[n+1], stp_v[n],
 * g_p[n-1], stp_p[n-2], g_s[n-1]. */
void ebo_axis_uniform(int64_t n, double size, double *g_b, double *stp_b, double *g_v,
                      double *stp_v, double *g_p, double *stp_p, double *g_s)
{
    basic_axis(n, size / (double)(n - 1), g_b, stp_b);
    ebo_axis_from_basic(n, g_b, stp_b, g_v, stp_v, g_p, stp_p, g_s);
}

/* The same for given basic node coordinates g_b (Spacing.jl:124-133 for the
 * basic steps). */
void ebo_axis_from_basic(int64_t n, const double *g_b, double *stp_b, double *g_v, double *stp_v,
                         double *g_p, double *stp_p, double *g_s)
{
    for (int64_t i = 1; i <= n - 1; ++i) EBO_V(stp_b, i) = EBO_V(g_b, i + 1) - EBO_V(g_b, i);
    staggered_axis(n, g_b, stp_b, g_v, stp_v);
    pressure_axis(n, g_b, g_p, stp_p);
    EBO_V(g_s, 1) = EBO_V(g_b, 1) + EBO_V(stp_b, 1) / 2.0;
    EBO_V(g_s, n - 1) = EBO_V(g_b, n) - EBO_V(stp_b, n - 1) / 2.0;
    for (int64_t i = 2; i <= n - 2; ++i) EBO_V(g_s, i) = EBO_V(g_b, i) + EBO_V(stp_b, i) / 2.0;
}
