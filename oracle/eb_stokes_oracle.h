/* Oracle (test infrastructure only): assembled Stokes-continuity system. */
#ifndef EB_STOKES_ORACLE_H
#define EB_STOKES_ORACLE_H
#include "eb_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Inputs of the K8 build, mirroring StokesBuildData
 * (solver/build_system/utils/BuildStructs.jl:37-177). Matrices are
 * column-major with the reference's shapes:
 *   etan (ynum-1, xnum-1), etas (ynum, xnum), rho1_vy (ynum, xnum+1),
 *   RX (ynum+1, xnum), RY (ynum, xnum+1), RC (ynum-1, xnum-1),
 *   btopx/btopy/bbottomx/bbottomy (xnum+1, 2), bleftx/blefty/brightx/brighty (ynum+1, 2).
 * bintern_zone[8] and bintern_velocity[8] are 0-based copies of the 1-based arrays. */
typedef struct {
    int64_t xnum, ynum;
    const double *xstp, *ystp, *xstpc, *ystpc;
    double xstpavg, ystpavg;
    const double *etan, *etas, *rho1_vy;
    const double *RX, *RY, *RC;
    const double *btopx, *btopy, *bbottomx, *bbottomy;
    const double *bleftx, *blefty, *brightx, *brighty;
    const int64_t *bintern_zone;
    const double *bintern_velocity;
    int64_t pressure_bc_mode;
    double pressure_bc;
    double pscale, timestep, gravity_y;
    int64_t iuse_interface_stabilization;
} ebo_stokes_in;

/* Inputs of K7 (viscoelastic terms and right-hand sides). */
typedef struct {
    int64_t xnum, ynum;
    const double *xstp_b, *ystp_b, *xstp_vy, *ystp_vx;
    const double *etas1, *mus1, *sxy1;      /* (ynum, xnum) */
    const double *etan1, *mun1, *sxx1;      /* (ynum-1, xnum-1) */
    const double *rho1, *rho1_vy;           /* (ynum, xnum), (ynum, xnum+1) */
    const double *dilatation, *eii_plastic_pressure; /* (ynum-1, xnum-1), may be NULL */
    double timestep, gravity_x, gravity_y;
    int64_t iuse_interface_stabilization;
} ebo_rhs_in;

/* Inputs of the grid strain rate and forecast stress (GridStrainRate.jl,
 * GridStress.jl). */
typedef struct {
    int64_t xnum, ynum;
    const double *xstp_b, *ystp_b, *xstp_vy, *ystp_vx;
    const double *vx1, *vy1;                /* (ynum+1, xnum), (ynum, xnum+1) */
    const double *etas1, *mus1, *sxy1;      /* (ynum, xnum) */
    const double *etan1, *mun1, *sxx1;      /* (ynum-1, xnum-1) */
    double timestep;
} ebo_strain_in;

void ebo_grid_strain_rate_and_stress(const ebo_strain_in *s, double *exx, double *exy, double *eii,
                                     double *sxy2, double *sxx2, double *dsxy, double *dsxx);

/* Inputs of the grid -> marker gathers after the solve (MarkerStrainRate.jl,
 * MarkerStrainRateRatio.jl, MarkerPressure.jl); sxx2 == NULL skips the ratio. */
typedef struct {
    int64_t xnum, ynum, marknum;
    const double *gridx_b, *gridy_b, *xstp_b, *ystp_b;
    const double *gridx_vy, *xstp_vy, *gridy_vx, *ystp_vx;
    const double *x, *y;
    const int32_t *xn, *yn;
    const int8_t *inside;
    const double *exx, *exy, *pr1;          /* (ynum-1,xnum-1), (ynum,xnum), (ynum-1,xnum-1) */
    const double *sxx2, *sxy2, *dsxx, *dsxy;
    const double *eta;
    const int16_t *matid;
    const double *mat_mu;
    double timestep;
} ebo_gather_in;

void ebo_marker_strain_rate_and_pressure(const ebo_gather_in *s, double *marker_dx, double *marker_dy,
                                         double *marker_exx, double *marker_exy, double *marker_sr_ratio,
                                         double *marker_pr, double *marker_sxx, double *marker_sxy);

int64_t ebo_stokes_build(const ebo_stokes_in *s, int64_t *Li, int64_t *Lj, double *Lv, double *R);
double ebo_pressure_scaling_coeff(double etan11, double xstpavg, double ystpavg);
void ebo_stokes_rhs(const ebo_rhs_in *r, double *etas0, double *sxy0, double *etan0,
                    double *sxx0, double *RX1, double *RY1, double *RC1);
void ebo_stokes_unpack(const ebo_stokes_in *s, const double *S, double *vx1, double *vy1, double *pr1);
void ebo_stokes_grid_residuals(const ebo_stokes_in *s, const double *vx1, const double *vy1,
                               const double *pr1, double *resx, double *resy, double *resc);
void ebo_coo_residual(int64_t N, int64_t nnz, const int64_t *Li, const int64_t *Lj,
                      const double *Lv, const double *x, const double *R, double *res);
void ebo_coo_spmv(int64_t N, int64_t nnz, const int64_t *Li, const int64_t *Lj,
                  const double *Lv, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
