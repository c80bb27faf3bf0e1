/* TEST INFRASTRUCTURE ONLY: grid -> marker gathers after the Stokes solve,
 * restated loop by loop from
 *   MarkerGridMapping.calc_normalized_distances_from_upper_left_basic_node!
 *     (interpolation/mapping/MarkerGridMapping.jl:185-218),
 *   MarkerStrainRate.calculate_marker_strain_rate!
 *     (stokes_continuity_solver/utils/strain_rate/MarkerStrainRate.jl:23-69),
 *   MarkerStrainRateRatio.calculate_marker_strain_rate_ratio!
 *     (stokes_continuity_solver/utils/strain_rate/MarkerStrainRateRatio.jl:71-188),
 *   MarkerPressure.calculate_marker_pressure!
 *     (stokes_continuity_solver/utils/MarkerPressure.jl:17-49),
 *   StressCorrection.apply_marker_stress_correction_for_remaining_grid_stress!
 *     (stokes_continuity_solver/subgrid_stress/stress_correction/StressCorrection.jl:27-83),
 * with MarkerGridMapping.upr_left_index_pressure / upr_left_dist_pressure
 * (MarkerGridMapping.jl:311-350) and GridToMarker.get_marker_value /
 * calc_marker_weights (interpolation/utils/GridToMarker.jl:190-226).  Arrays
 * are 1-based through EBO_M/EBO_V as in the reference. */
#include <math.h>

#include "eb_stokes_oracle.h"

/* MarkerGridMapping.jl:311-331 */
static int32_t upr_left_index_pressure(int64_t nnodes, int32_t basic, double loc, const double *stag)
{
    int32_t p = (loc < EBO_V(stag, basic + 1)) ? basic - 1 : basic;
    if (p < 1) p = 1;
    if (p > (int32_t)(nnodes - 2)) p = (int32_t)(nnodes - 2);
    return p;
}

/* MarkerGridMapping.jl:339-350 */
static double upr_left_dist_pressure(int32_t p, double loc, const double *stag, const double *stp)
{
    return (loc - EBO_V(stag, p + 1)) / EBO_V(stp, p + 1);
}

/* GridToMarker.jl:190-226 */
static double get_marker_value(int32_t yn, int32_t xn, double dy, double dx, const double *a, int64_t ld)
{
    const double wul = (1.0 - dx) * (1.0 - dy), wll = (1.0 - dx) * dy;
    const double wur = dx * (1.0 - dy), wlr = dx * dy;
    double v = 0.0;
    v += wul * EBO_M(a, ld, yn, xn);
    v += wll * EBO_M(a, ld, yn + 1, xn);
    v += wur * EBO_M(a, ld, yn, xn + 1);
    v += wlr * EBO_M(a, ld, yn + 1, xn + 1);
    return v;
}

/* MarkerStrainRateRatio.jl:146-186 (the warnings aside) */
static double marker_strain_rate_ratio(double dt, double sxx, double sxy, double dsxx, double dsxy,
                                       double mu, double eta, double eii_v)
{
    if (eii_v > 0) {
        const double tf = 1.0 / (dt * mu);
        const double a = 0.5 * (sxx / eta + dsxx * tf);
        const double b = 0.5 * (sxy / eta + dsxy * tf);
        return sqrt(a * a + b * b) / eii_v;
    }
    return 1.0;
}

void ebo_marker_strain_rate_and_pressure(const ebo_gather_in *s, double *marker_dx, double *marker_dy,
                                         double *marker_exx, double *marker_exy, double *marker_sr_ratio,
                                         double *marker_pr, double *marker_sxx, double *marker_sxy)
{
    const int64_t xnum = s->xnum, ynum = s->ynum;
    for (int64_t m = 0; m < s->marknum; ++m)  /* calc_normalized_distances_from_upper_left_basic_node! */
        if (s->inside[m] == 1) {
            const int32_t xn = s->xn[m], yn = s->yn[m];
            marker_dx[m] = (s->x[m] - EBO_V(s->gridx_b, xn)) / EBO_V(s->xstp_b, xn);
            marker_dy[m] = (s->y[m] - EBO_V(s->gridy_b, yn)) / EBO_V(s->ystp_b, yn);
        }
    for (int64_t m = 0; m < s->marknum; ++m)  /* calculate_marker_strain_rate! */
        if (s->inside[m] == 1) {
            const int32_t ypr = upr_left_index_pressure(ynum, s->yn[m], s->y[m], s->gridy_vx);
            const int32_t xpr = upr_left_index_pressure(xnum, s->xn[m], s->x[m], s->gridx_vy);
            const double dypr = upr_left_dist_pressure(ypr, s->y[m], s->gridy_vx, s->ystp_vx);
            const double dxpr = upr_left_dist_pressure(xpr, s->x[m], s->gridx_vy, s->xstp_vy);
            marker_exx[m] = get_marker_value(ypr, xpr, dypr, dxpr, s->exx, ynum - 1);
            marker_exy[m] = get_marker_value(s->yn[m], s->xn[m], marker_dy[m], marker_dx[m], s->exy, ynum);
        }
    if (s->sxx2)
        for (int64_t m = 0; m < s->marknum; ++m)  /* calculate_marker_strain_rate_ratio! */
            if (s->inside[m] == 1) {
                const int32_t yn = s->yn[m], xn = s->xn[m];
                const double dy = marker_dy[m], dx = marker_dx[m];
                const int32_t ypr = upr_left_index_pressure(ynum, yn, s->y[m], s->gridy_vx);
                const int32_t xpr = upr_left_index_pressure(xnum, xn, s->x[m], s->gridx_vy);
                const double dypr = upr_left_dist_pressure(ypr, s->y[m], s->gridy_vx, s->ystp_vx);
                const double dxpr = upr_left_dist_pressure(xpr, s->x[m], s->gridx_vy, s->xstp_vy);
                const double sxx = get_marker_value(ypr, xpr, dypr, dxpr, s->sxx2, ynum - 1);
                const double sxy = get_marker_value(yn, xn, dy, dx, s->sxy2, ynum);
                const double dsxx = get_marker_value(ypr, xpr, dypr, dxpr, s->dsxx, ynum - 1);
                const double dsxy = get_marker_value(yn, xn, dy, dx, s->dsxy, ynum);
                const double mu = EBO_V(s->mat_mu, s->matid[m]);
                const double exx = marker_exx[m], exy = marker_exy[m];
                const double eii_v = sqrt(exx * exx + exy * exy);
                marker_sr_ratio[m] = marker_strain_rate_ratio(s->timestep, sxx, sxy, dsxx, dsxy, mu,
                                                              s->eta[m], eii_v);
            }
    for (int64_t m = 0; m < s->marknum; ++m)  /* calculate_marker_pressure! */
        if (s->inside[m] == 1) {
            const int32_t ypr = upr_left_index_pressure(ynum, s->yn[m], s->y[m], s->gridy_vx);
            const int32_t xpr = upr_left_index_pressure(xnum, s->xn[m], s->x[m], s->gridx_vy);
            const double dypr = upr_left_dist_pressure(ypr, s->y[m], s->gridy_vx, s->ystp_vx);
            const double dxpr = upr_left_dist_pressure(xpr, s->x[m], s->gridx_vy, s->xstp_vy);
            marker_pr[m] = get_marker_value(ypr, xpr, dypr, dxpr, s->pr1, ynum - 1);
        }
    if (marker_sxx && marker_sxy)
        for (int64_t m = 0; m < s->marknum; ++m)  /* apply_marker_stress_correction_for_remaining_grid_stress! */
            if (s->inside[m] == 1) {
                const double sxy = marker_sxy[m], sxx = marker_sxx[m];
                marker_sxy[m] = sxy + get_marker_value(s->yn[m], s->xn[m], marker_dy[m], marker_dx[m], s->dsxy, ynum);
                const int32_t ypr = upr_left_index_pressure(ynum, s->yn[m], s->y[m], s->gridy_vx);
                const int32_t xpr = upr_left_index_pressure(xnum, s->xn[m], s->x[m], s->gridx_vy);
                const double dypr = upr_left_dist_pressure(ypr, s->y[m], s->gridy_vx, s->ystp_vx);
                const double dxpr = upr_left_dist_pressure(xpr, s->x[m], s->gridx_vy, s->xstp_vy);
                marker_sxx[m] = sxx + get_marker_value(ypr, xpr, dypr, dxpr, s->dsxx, ynum - 1);
            }
}
