/*
 * Oracle (test infrastructure only): the heat-equation path of one time step
 * and the sub-grid diffusion of temperature and stress, restated from the
 * reference:
 *   heat scatter   interpolation/utils/MarkerToHeatArrays.jl:29-35,
 *                  BilinearNumerator.jl:31-141, BilinearAverage.jl:10-50,276-294,
 *                  backup/BackupTransportArrays.jl:144-178,
 *                  utils/heat_funcs/ApplyThermalBC.jl:20-44
 *   right side     heat_solver/HeatSolver.jl:107-118, rhs/HeatRhs.jl:5-79
 *   system         solver/build_system/HeatBuildManager.jl:47-83,
 *                  internal/InternalNodes.jl:19-197, boundary/BoundaryNodes.jl:16-251,
 *                  utils/BuildStructs.jl:97-124 (stencil values)
 *   residuals      solver/residuals/HeatResiduals.jl:12-162
 *   marker updates heat_solver/subgrid_heat/SubgridDiffusion.jl:66-246,
 *                  marker_temperature/MarkerTemperature.jl:18-72,
 *                  interpolation/utils/GridTemperatureToMarkers.jl:15-35
 *   stress         stokes_continuity_solver/subgrid_stress/update/stress_change/
 *                  SubgridShearStress.jl:25-96, SubgridNormalStress.jl:24-89,
 *                  DiffusionTerm.jl:13-29
 * The reference's x^2.0 is evaluated as the correctly rounded square x*x.
 */
#include "eb_heat_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* GridToMarker.get_marker_value (GridToMarker.jl:190-213) */
static double basic_value(int32_t yn, int32_t xn, double dy, double dx, const double *a, int64_t ld)
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

/* MarkerGridMapping.jl:211-214 (basic-cell distances) */
static void basic_dist(const ebo_heat_grid *g, double x, double y, int32_t xn, int32_t yn, double *dx,
                       double *dy)
{
    *dx = (x - EBO_V(g->gridx_b, xn)) / EBO_V(g->xstp_b, xn);
    *dy = (y - EBO_V(g->gridy_b, yn)) / EBO_V(g->ystp_b, yn);
}

void ebo_scatter_basic_inclusive(const ebo_heat_grid *g, int64_t M, const double *x, const double *y,
                                 const int8_t *flags, const int32_t *xn, const int32_t *yn, int nfield,
                                 const double *const *vals, double *const *out, double *wtnodes)
{
    const int64_t ny = g->ynum, nb = g->xnum * g->ynum;
    double **bak = malloc(sizeof(double *) * (nfield ? nfield : 1));
    for (int f = 0; f < nfield; ++f) {  /* backup, then clear (BackupTransportArrays.jl:144-178) */
        bak[f] = malloc(sizeof(double) * nb);
        memcpy(bak[f], out[f], sizeof(double) * nb);
        memset(out[f], 0, sizeof(double) * nb);
    }
    memset(wtnodes, 0, sizeof(double) * nb);
    for (int64_t m = 0; m < M; ++m) {
        if (flags[m] != 1) continue;
        double dx, dy;
        basic_dist(g, x[m], y[m], xn[m], yn[m], &dx, &dy);
        /* MarkerGridNodeWeights.jl:68-91 */
        const double ul = (1.0 - dx) * (1.0 - dy), ll = (1.0 - dx) * dy;
        const double ur = dx * (1.0 - dy), lr = dx * dy;
        const int32_t ix = xn[m], iy = yn[m];
        /* WeightFuncs.update_weight_inclusive! (WeightFuncs.jl:67-89) */
        EBO_M(wtnodes, ny, iy, ix) += ul * 1.0 * 1.0;
        EBO_M(wtnodes, ny, iy + 1, ix) += ll * 1.0 * 1.0;
        EBO_M(wtnodes, ny, iy, ix + 1) += ur * 1.0 * 1.0;
        EBO_M(wtnodes, ny, iy + 1, ix + 1) += lr * 1.0 * 1.0;
        for (int f = 0; f < nfield; ++f) {
            const double v = vals[f][m];
            EBO_M(out[f], ny, iy, ix) += ul * v * 1.0;
            EBO_M(out[f], ny, iy + 1, ix) += ll * v * 1.0;
            EBO_M(out[f], ny, iy, ix + 1) += ur * v * 1.0;
            EBO_M(out[f], ny, iy + 1, ix + 1) += lr * v * 1.0;
        }
    }
    for (int f = 0; f < nfield; ++f) {  /* BilinearAverage.jl:276-294 */
        for (int64_t k = 0; k < nb; ++k) out[f][k] = (wtnodes[k] != 0) ? out[f][k] / wtnodes[k] : bak[f][k];
        free(bak[f]);
    }
    free(bak);
}

void ebo_marker_heat_production(int64_t M, const int8_t *flags, const int16_t *matid, const double *mat_hr,
                                const double *serp, double *hr)
{
    for (int64_t m = 0; m < M; ++m) {
        if (flags[m] != 1) {
            hr[m] = 0.0;
            continue;
        }
        const double r = EBO_V(mat_hr, matid[m]);
        hr[m] = serp ? r + serp[m] : r;
    }
}

void ebo_apply_thermal_bc(int64_t xnum, int64_t ynum, const double *btopt, const double *bbottomt,
                          const double *bleftt, const double *brightt, double *tk1)
{
    for (int64_t j = 2; j <= xnum - 1; ++j) {
        EBO_M(tk1, ynum, 1, j) = EBO_M(btopt, xnum, j, 1) + EBO_M(btopt, xnum, j, 2) * EBO_M(tk1, ynum, 2, j);
        EBO_M(tk1, ynum, ynum, j) =
            EBO_M(bbottomt, xnum, j, 1) + EBO_M(bbottomt, xnum, j, 2) * EBO_M(tk1, ynum, ynum - 1, j);
    }
    for (int64_t i = 1; i <= ynum; ++i) {
        EBO_M(tk1, ynum, i, 1) = EBO_M(bleftt, ynum, i, 1) + EBO_M(bleftt, ynum, i, 2) * EBO_M(tk1, ynum, i, 2);
        EBO_M(tk1, ynum, i, xnum) =
            EBO_M(brightt, ynum, i, 1) + EBO_M(brightt, ynum, i, 2) * EBO_M(tk1, ynum, i, xnum - 1);
    }
}

void ebo_heat_rhs(int64_t xnum, int64_t ynum, const double *hr1, const double *ha1, const double *rho1,
                  const double *vx1, const double *vy1, const double *sxy2, const double *etas1,
                  const double *sxx2, const double *etan1, double gx, double gy, int iuse_adiabatic,
                  int iuse_shear, double *RT1)
{
    const int64_t ny = ynum, nb = xnum * ynum, nvx = ynum + 1, npr = ynum - 1;
    memcpy(RT1, hr1, sizeof(double) * nb);  /* HeatRhs.jl:5-8 */
    if (iuse_adiabatic)                        /* HeatRhs.jl:25-45 */
        for (int64_t j = 2; j <= xnum - 1; ++j)
            for (int64_t i = 2; i <= ynum - 1; ++i)
                EBO_M(RT1, ny, i, j) += EBO_M(ha1, ny, i, j) * EBO_M(rho1, ny, i, j) *
                                        (gx * (EBO_M(vx1, nvx, i, j) + EBO_M(vx1, nvx, i + 1, j)) +
                                         gy * (EBO_M(vy1, ny, i, j) + EBO_M(vy1, ny, i, j + 1))) /
                                        2.0;
    if (iuse_shear)  /* HeatRhs.jl:58-79 */
        for (int64_t j = 2; j <= xnum - 1; ++j)
            for (int64_t i = 2; i <= ynum - 1; ++i) {
                const double s = EBO_M(sxy2, ny, i, j);
                EBO_M(RT1, ny, i, j) += s * s / EBO_M(etas1, ny, i, j);
                const double a = EBO_M(sxx2, npr, i - 1, j - 1), b = EBO_M(sxx2, npr, i, j - 1);
                const double c = EBO_M(sxx2, npr, i - 1, j), d = EBO_M(sxx2, npr, i, j);
                EBO_M(RT1, ny, i, j) += (a * a / EBO_M(etan1, npr, i - 1, j - 1) +
                                         b * b / EBO_M(etan1, npr, i, j - 1) +
                                         c * c / EBO_M(etan1, npr, i - 1, j) + d * d / EBO_M(etan1, npr, i, j)) /
                                        4.0;
            }
}

#define PUT(row, col, val)        \
    do {                          \
        Li[nz] = (row);           \
        Lj[nz] = (col);           \
        Lv[nz] = (val);           \
        ++nz;                     \
    } while (0)

int64_t ebo_heat_build(const ebo_heat_grid *g, const double *btopt, const double *bbottomt,
                       const double *bleftt, const double *brightt, const double *RT1, const double *tk0,
                       const double *rhocp1, const double *kt1, double dt, int64_t *Li, int64_t *Lj,
                       double *Lv, double *R)
{
    const int64_t nx = g->xnum, ny = g->ynum;
    int64_t nz = 0;
    for (int64_t j = 1; j <= nx; ++j)
        for (int64_t i = 1; i <= ny; ++i) {
            const int64_t itk = (j - 1) * ny + i;
            if (j == 1 || j == nx || i == 1 || i == ny) {  /* BoundaryNodes.jl:16-39 */
                if (i == 1 && 1 < j && j < nx) {
                    R[itk - 1] = EBO_M(btopt, nx, j, 1);
                    PUT(itk, itk, 1.0);
                    PUT(itk, itk + 1, -EBO_M(btopt, nx, j, 2));
                }
                if (i == ny && 1 < j && j < nx) {
                    R[itk - 1] = EBO_M(bbottomt, nx, j, 1);
                    PUT(itk, itk, 1.0);
                    PUT(itk, itk - 1, -EBO_M(bbottomt, nx, j, 2));
                }
                if (j == 1) {
                    R[itk - 1] = EBO_M(bleftt, ny, i, 1);
                    PUT(itk, itk, 1.0);
                    PUT(itk, itk + ny, -EBO_M(bleftt, ny, i, 2));
                }
                if (j == nx) {
                    R[itk - 1] = EBO_M(brightt, ny, i, 1);
                    PUT(itk, itk, 1.0);
                    PUT(itk, itk - ny, -EBO_M(brightt, ny, i, 2));
                }
                continue;
            }
            /* BuildStructs.HeatStencilData (BuildStructs.jl:97-124) */
            const double tkC = EBO_M(tk0, ny, i, j), rhocpC = EBO_M(rhocp1, ny, i, j);
            const double ktU = EBO_M(kt1, ny, i - 1, j), ktC = EBO_M(kt1, ny, i, j);
            const double ktD = EBO_M(kt1, ny, i + 1, j), ktL = EBO_M(kt1, ny, i, j - 1);
            const double ktR = EBO_M(kt1, ny, i, j + 1);
            const double dxL = EBO_V(g->xstp_b, j - 1), dxR = EBO_V(g->xstp_b, j);
            const double dxC = EBO_V(g->xstp_pr, j - 1), dyU = EBO_V(g->ystp_b, i - 1);
            const double dyD = EBO_V(g->ystp_b, i), dyC = EBO_V(g->ystp_pr, i - 1);
            R[itk - 1] = EBO_M(RT1, ny, i, j) + tkC * rhocpC / dt;  /* InternalNodes.jl:37-53 */
            PUT(itk, itk,
                rhocpC / dt + ((ktL + ktC) / dxL + (ktC + ktR) / dxR) / 2.0 / dxC +
                    ((ktU + ktC) / dyU + (ktC + ktD) / dyD) / 2.0 / dyC);
            PUT(itk, itk - ny, -(ktL + ktC) / 2.0 / dxL / dxC);
            PUT(itk, itk + ny, -(ktC + ktR) / 2.0 / dxR / dxC);
            PUT(itk, itk - 1, -(ktU + ktC) / 2.0 / dyU / dyC);
            PUT(itk, itk + 1, -(ktC + ktD) / 2.0 / dyD / dyC);
        }
    return nz;
}

void ebo_heat_residuals(const ebo_heat_grid *g, const double *RT1, const double *rhocp1, const double *tk0,
                        const double *tk2, const double *kt1, double dt, double *rest)
{
    const int64_t nx = g->xnum, ny = g->ynum;
    for (int64_t j = 1; j <= nx; ++j)
        for (int64_t i = 1; i <= ny; ++i) {
            if (j == 1 || j == nx || i == 1 || i == ny) {
                EBO_M(rest, ny, i, j) = 0.0;
                continue;
            }
            double r = EBO_M(RT1, ny, i, j) +
                       -EBO_M(rhocp1, ny, i, j) * (EBO_M(tk2, ny, i, j) - EBO_M(tk0, ny, i, j)) / dt;
            r += ((EBO_M(kt1, ny, i, j) + EBO_M(kt1, ny, i, j + 1)) *
                      (EBO_M(tk2, ny, i, j + 1) - EBO_M(tk2, ny, i, j)) / EBO_V(g->xstp_b, j) -
                  (EBO_M(kt1, ny, i, j - 1) + EBO_M(kt1, ny, i, j)) *
                      (EBO_M(tk2, ny, i, j) - EBO_M(tk2, ny, i, j - 1)) / EBO_V(g->xstp_b, j - 1)) /
                 2.0 / EBO_V(g->xstp_pr, j - 1);
            r += ((EBO_M(kt1, ny, i, j) + EBO_M(kt1, ny, i + 1, j)) *
                      (EBO_M(tk2, ny, i + 1, j) - EBO_M(tk2, ny, i, j)) / EBO_V(g->ystp_b, i) -
                  (EBO_M(kt1, ny, i - 1, j) + EBO_M(kt1, ny, i, j)) *
                      (EBO_M(tk2, ny, i, j) - EBO_M(tk2, ny, i - 1, j)) / EBO_V(g->ystp_b, i - 1)) /
                 2.0 / EBO_V(g->ystp_pr, i - 1);
            EBO_M(rest, ny, i, j) = r;
        }
}

void ebo_grid_to_marker_temperature(const ebo_heat_grid *g, int64_t M, const double *x, const double *y,
                                    const int8_t *flags, const int32_t *xn, const int32_t *yn,
                                    const double *grid, int mode, double *marker_TK)
{
    for (int64_t m = 0; m < M; ++m) {
        if (flags[m] != 1) continue;
        double dx, dy;
        basic_dist(g, x[m], y[m], xn[m], yn[m], &dx, &dy);
        const double v = basic_value(yn[m], xn[m], dy, dx, grid, g->ynum);
        marker_TK[m] = mode ? marker_TK[m] + v : v;
    }
}

void ebo_subgrid_temperature_markers(const ebo_heat_grid *g, int64_t M, const double *x, const double *y,
                                     const int8_t *flags, const int32_t *xn, const int32_t *yn,
                                     const double *tk1, const double *kt1, const double *rhocp1, double dt,
                                     double coef, double xstpavg, double ystpavg, double *marker_TK,
                                     double *dT)
{
    for (int64_t m = 0; m < M; ++m) {
        if (flags[m] != 1) {
            dT[m] = 0.0;
            continue;
        }
        double dx, dy;
        basic_dist(g, x[m], y[m], xn[m], yn[m], &dx, &dy);
        const double tn = basic_value(yn[m], xn[m], dy, dx, tk1, g->ynum);
        const double total = tn - marker_TK[m];
        const double k = basic_value(yn[m], xn[m], dy, dx, kt1, g->ynum);
        const double rc = basic_value(yn[m], xn[m], dy, dx, rhocp1, g->ynum);
        /* SubgridDiffusion.jl:187-197 and :218-246 */
        const double tsg = rc / k / (2.0 / (xstpavg * xstpavg) + 2.0 / (ystpavg * ystpavg));
        double term = -coef * dt / tsg;
        term = term > -30.0 ? term : -30.0;
        const double d = total * (1.0 - exp(term));
        dT[m] = d;
        marker_TK[m] += d;
    }
}

void ebo_sticky_temperature_correction(int64_t M, const int8_t *flags, const int16_t *matid, int16_t air,
                                       int16_t water, double ttop, double *marker_TK)
{
    for (int64_t m = 0; m < M; ++m)
        if (flags[m] == 1 && (matid[m] == air || matid[m] == water) && marker_TK[m] != ttop)
            marker_TK[m] = ttop;
}

/* MarkerGridMapping.jl:311-350 */
static int32_t pr_index(int64_t nnodes, int32_t basic, double loc, const double *stag)
{
    int32_t p = (loc < EBO_V(stag, basic + 1)) ? basic - 1 : basic;
    if (p < 1) p = 1;
    if (p > (int32_t)(nnodes - 2)) p = (int32_t)(nnodes - 2);
    return p;
}

/* DiffusionTerm.calculate_subgrid_stress_diffusion_term (DiffusionTerm.jl:13-29) */
static double stress_diffusion(double eta, double mu, double dt, double coef)
{
    const double maxwell = eta / mu;
    double s = -coef * dt / maxwell;
    s = s > -30.0 ? s : -30.0;
    return 1.0 - exp(s);
}

void ebo_subgrid_stress_markers(const ebo_heat_grid *g, int64_t M, const double *x, const double *y,
                                const int8_t *flags, const int32_t *xn, const int32_t *yn,
                                const int16_t *matid, const double *mat_mu, const double *eta,
                                const double *sxy1, const double *sxx1, double dt, double coef,
                                double *marker_sxy, double *marker_sxx, double *dsxy_m, double *dsxx_m)
{
    const int64_t nx = g->xnum, ny = g->ynum;
    for (int64_t m = 0; m < M; ++m) {  /* SubgridShearStress.jl:60-95 */
        if (flags[m] != 1) {
            dsxy_m[m] = 0.0;
            continue;
        }
        const double term = stress_diffusion(eta[m], EBO_V(mat_mu, matid[m]), dt, coef);
        double dx, dy;
        basic_dist(g, x[m], y[m], xn[m], yn[m], &dx, &dy);
        const double from_grid = basic_value(yn[m], xn[m], dy, dx, sxy1, ny);
        const double d = (from_grid - marker_sxy[m]) * term;
        marker_sxy[m] += d;
        dsxy_m[m] = d;
    }
    for (int64_t m = 0; m < M; ++m) {  /* SubgridNormalStress.jl:58-88 */
        if (flags[m] != 1) {
            dsxx_m[m] = 0.0;
            continue;
        }
        const double term = stress_diffusion(eta[m], EBO_V(mat_mu, matid[m]), dt, coef);
        /* GridToMarker.get_marker_value_from_pressure_grid_array (GridToMarker.jl:104-160) */
        const int32_t yp = pr_index(ny, yn[m], y[m], g->gridy_vx);
        const int32_t xp = pr_index(nx, xn[m], x[m], g->gridx_vy);
        const double dyp = (y[m] - EBO_V(g->gridy_vx, yp + 1)) / EBO_V(g->ystp_vx, yp + 1);
        const double dxp = (x[m] - EBO_V(g->gridx_vy, xp + 1)) / EBO_V(g->xstp_vy, xp + 1);
        const double from_grid = basic_value(yp, xp, dyp, dxp, sxx1, ny - 1);
        const double d = (from_grid - marker_sxx[m]) * term;
        marker_sxx[m] += d;
        dsxx_m[m] = d;
    }
}
