/* Oracle (test infrastructure only): reference 3-D multigrid restatement. */
#ifndef EB_MG3D_ORACLE_H
#define EB_MG3D_ORACLE_H
#include "eb_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 1-based column-major 3-D accessor: element (i, j, k) of an array with n1
 * rows (y) and n2 columns (x) per z-plane, as Julia stores Array{Float64,3}. */
#define EBO_A3(a, n1, n2, i, j, k) \
    ((a)[((size_t)(k) - 1) * (size_t)(n1) * (size_t)(n2) + ((size_t)(j) - 1) * (size_t)(n1) + (size_t)((i) - 1)])

/* One level (multigrid/level_manager/LevelManager.jl:78-134, LevelData).
 * Shapes (y, x, z): vx/RX/rx (ynum+1, xnum, znum+1); vy/RY/ry (ynum, xnum+1,
 * znum+1); vz/RZ/rz (ynum+1, xnum+1, znum); pr/RC/rc/etan (ynum-1, xnum-1,
 * znum-1); etaxy (ynum, xnum, znum-1); etaxz (ynum-1, xnum, znum); etayz
 * (ynum, xnum-1, znum).  Steps (StencilArrays3d, LevelManager.jl:18-76):
 * xstp_b/ystp_b/zstp_b basic (n-1), ystp_vx/zstp_vx/xstp_vy staggered (n). */
typedef struct {
    int64_t xnum, ynum, znum;
    const double *xstp_b, *ystp_b, *zstp_b, *ystp_vx, *zstp_vx, *xstp_vy;
    double *etan, *etaxy, *etaxz, *etayz;
    double *vx, *vy, *vz, *pr;
    double *RX, *RY, *RZ, *RC;
    double *rx, *ry, *rz, *rc;
} ebo_mg3level;

/* Separable fine->coarse maps (GridMappingManager.jl:187-249).  Axis grid
 * kinds: 0 basic nodes, 1 staggered nodes with ghosts (vx y/z, vy x/z, vz
 * x/y), 2 pressure nodes, 3 shear-centre nodes (etaxy z, etaxz y, etayz x).
 * For each kind and axis: left coarse node index and normalised distance per
 * fine node (1-based values, 0-based storage). */
typedef struct {
    const int32_t *iy[4], *jx[4], *kz[4];
    const double *dy[4], *dx[4], *dz[4];
} ebo_mg3map;

double ebo_mg3_vx_residual(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k, double *coef);
double ebo_mg3_vy_residual(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k, double *coef);
double ebo_mg3_vz_residual(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k, double *coef);
double ebo_mg3_pr_residual(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k);
void ebo_mg3_set_bc(ebo_mg3level *l);
void ebo_mg3_sweep(ebo_mg3level *l, double theta_s, double theta_c);
void ebo_mg3_sweep_serial(ebo_mg3level *l, double theta_s, double theta_c);
void ebo_mg3_normalize_pressure(ebo_mg3level *l, double prnorm);
void ebo_mg3_residuals(ebo_mg3level *l);
void ebo_mg3_smoother(ebo_mg3level *l, int64_t nsweeps, double theta_s, double theta_c,
                      double prnorm, int serial);
void ebo_mg3_smoother_velocity(ebo_mg3level *l, int64_t nsweeps, double theta_s);
void ebo_mg3_restrict(const ebo_mg3level *f, ebo_mg3level *c, const ebo_mg3map *m,
                      double *wtx, double *wty, double *wtz, double *wtc);
void ebo_mg3_prolong(const ebo_mg3level *c, ebo_mg3level *f, const ebo_mg3map *m,
                     double relax_v, double relax_p);
void ebo_mg3_viscosity_restrict(const ebo_mg3level *f, ebo_mg3level *c, const ebo_mg3map *m,
                                double *wxy, double *wxz, double *wyz, double *wn);
void ebo_mg3_apply(ebo_mg3level *l);
void ebo_mg3_mean_residuals(const ebo_mg3level *l, double stokesscale, double continscale, double *out);
void ebo_axis_uniform(int64_t n, double size, double *g_b, double *stp_b, double *g_v,
                      double *stp_v, double *g_p, double *stp_p, double *g_s);
void ebo_axis_from_basic(int64_t n, const double *g_b, double *stp_b, double *g_v, double *stp_v,
                         double *g_p, double *stp_p, double *g_s);

#ifdef __cplusplus
}
#endif
#endif
