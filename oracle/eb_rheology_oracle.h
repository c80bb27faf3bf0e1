/*
 * TEST INFRASTRUCTURE ONLY: plain-C restatement of the reference's marker
 * rheology (failure properties, composite-creep flow viscosity, viscoplastic
 * viscosity), the grid -> marker viscoplastic viscosity and the marker strain
 * update.  Conventions as eb_oracle.h; material tables column-major with
 * leading dimension nmat, 1-based matid.
 */
#ifndef EB_RHEOLOGY_ORACLE_H
#define EB_RHEOLOGY_ORACLE_H
#include "eb_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t marknum;
    const int8_t *flags;
    const int16_t *matid;
    const double *strain_plastic, *fric_ini, *melt_damage;
    const double *sxx, *sxy, *exx, *exy, *TK, *pr, *GII, *sr_ratio;
    int64_t nmat;
    const int64_t *mat_flow_type;
    const double *mat_flow, *mat_plastic, *mat_mu;
    double timestep, temperature_top, viscosity_min, viscosity_max, powerlaw_stress_min, gas_constant;
    int64_t iuse_viscous_strain_soft;
    double vsoftfac;
    int64_t itype_global, ntimestep;
} ebo_rheology_in;

void ebo_failure_properties(const ebo_rheology_in *in, double *cohesion, double *fric);
void ebo_flow_viscosity(const ebo_rheology_in *in, double *eta_flow, double *preexp);
void ebo_viscoplastic_viscosity(const ebo_rheology_in *in, const double *eta_flow, double *eta);
void ebo_marker_viscoplastic_from_grid(int64_t marknum, const int8_t *flags, const int32_t *xn,
                                       const int32_t *yn, const double *dx, const double *dy, int64_t ynum,
                                       const double *plastic_yield, const double *etas1,
                                       const double *eta_flow, double *eta, float *pfailure);
void ebo_update_marker_strain(int64_t marknum, const int8_t *flags, double timestep, double healing,
                              const double *exx, const double *exy, const double *sxx, const double *sxy,
                              const double *eta, const double *eta_flow, const float *pfailure, double *GII,
                              double *strain_plastic, double *strain_rate_plastic);

#ifdef __cplusplus
}
#endif
#endif
