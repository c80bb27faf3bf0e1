/* Oracle (test infrastructure only): reference 2-D multigrid restatement. */
#ifndef EB_MG_ORACLE_H
#define EB_MG_ORACLE_H
#include "eb_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One level (multigrid/level_manager/LevelManager.jl:78-352, LevelData2d).
 * Shapes: vx/RX/rx (ynum+1, xnum), vy/RY/ry (ynum, xnum+1),
 * pr/RC/rc/etan (ynum-1, xnum-1), etas (ynum, xnum). */
typedef struct {
    int64_t xnum, ynum;
    const double *xstp_b, *ystp_b, *xstp_vy, *ystp_vx;
    double *etas, *etan;
    double *vx, *vy, *pr;
    double *RX, *RY, *RC;
    double *rx, *ry, *rc;
} ebo_mglevel;

/* Separable fine->coarse (or coarse->fine) maps (GridMappingManager.jl):
 * for each grid type, left-node index and normalised distance per axis,
 * indexed by the source-axis node (1-based, stored 0-based). */
typedef struct {
    const int32_t *vx_iy, *vx_jx, *vy_iy, *vy_jx, *pr_iy, *pr_jx, *b_iy, *b_jx;
    const double *vx_dy, *vx_dx, *vy_dy, *vy_dx, *pr_dy, *pr_dx, *b_dy, *b_dx;
} ebo_mgmap;

double ebo_mg_vx_residual(const ebo_mglevel *l, int64_t i, int64_t j, double *coef);
double ebo_mg_vy_residual(const ebo_mglevel *l, int64_t i, int64_t j, double *coef);
double ebo_mg_pr_residual(const ebo_mglevel *l, int64_t i, int64_t j);
void ebo_mg_set_bc(ebo_mglevel *l, double d_top, double d_bottom, double d_left, double d_right);
void ebo_mg_sweep(ebo_mglevel *l, double theta_s, double theta_c);
void ebo_mg_normalize_pressure(ebo_mglevel *l, double prnorm);
void ebo_mg_residuals(ebo_mglevel *l);
void ebo_mg_smoother(ebo_mglevel *l, int64_t nsweeps, double theta_s, double theta_c,
                     double prnorm, const double *bc_d);
void ebo_mg_axis_map(int64_t na, const double *ga, int64_t nb, const double *gb,
                     int32_t *idx, double *d);
void ebo_mg_restrict(const ebo_mglevel *f, ebo_mglevel *c, const ebo_mgmap *m,
                     double *wtx, double *wty, double *wtc);
void ebo_mg_prolong(const ebo_mglevel *c, ebo_mglevel *f, const ebo_mgmap *m,
                    double relax_v, double relax_p);
void ebo_mg_viscosity_restrict(const ebo_mglevel *f, ebo_mglevel *c, const ebo_mgmap *m,
                               double *wts, double *wtn);
void ebo_mg_scale_viscosity(int64_t n, const double *eta_orig, double *eta, double eta_min_o,
                            double eta_max_o, double eta_min_cur, double eta_max_cur);
void ebo_mg_mean_residuals(const ebo_mglevel *l, double stokesscale, double continscale, double *out);

#ifdef __cplusplus
}
#endif
#endif
