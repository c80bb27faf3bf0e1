/* Oracle (test infrastructure only): marker -> grid Stokes property scatter. */
#ifndef EB_SCATTER_ORACLE_H
#define EB_SCATTER_ORACLE_H
#include "eb_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t marknum, xnum, ynum;
    const double *x, *y;
    const int8_t *flags;
    const double *gridx_b, *gridy_b, *xstp_b, *ystp_b, *gridx_vy, *xstp_vy;
    const double *rho, *cohesion, *fric, *eta_flow, *eta, *sxy, *sxx, *meltfrac;
    const float *pfailure, *dilatation;
    const int16_t *matid;
    const double *mat_mu; /* indexed by matid (1-based) */
} ebo_scatter_in;

/* Grid fields are in/out: on entry the previous values (the backup used where
 * a node receives no marker weight), on exit the new averages.  Shapes:
 * basic (ynum,xnum): rho1 etas1 eta_flow mus1 sxy1 plastics cohesion fric_deg;
 * pressure (ynum-1,xnum-1): etan1 mun1 sxx1 plasticn dilatation meltfrac;
 * vy (ynum,xnum+1): rho1_vy. */
typedef struct {
    double *rho1, *rho1_vy, *etas1, *etan1, *eta_flow, *mus1, *mun1, *sxy1, *sxx1;
    double *plastics, *plasticn, *cohesion, *fric_deg, *dilatation, *meltfrac;
    double *wtnodes, *wtetas, *wtetan, *wtnodes_vy;
    int32_t *xn, *yn;
} ebo_scatter_out;

/* returns 1 if a zero shear modulus was met (the reference raises) */
int ebo_scatter_stokes(const ebo_scatter_in *in, ebo_scatter_out *out);

#ifdef __cplusplus
}
#endif
#endif
