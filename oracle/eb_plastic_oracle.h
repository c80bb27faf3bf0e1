/*
 * TEST INFRASTRUCTURE ONLY: plain-C restatement of the nodal plastic-yield
 * update of the reference (rheology/plastic_failure/nodes/PlasticFailureNodes.jl
 * :24-319 with plastic_failure/utils/{CheckYield,YieldStress,FluidPressure}.jl
 * and utils/MobileWall.jl:106-122, 279-329).  Conventions as eb_oracle.h.
 */
#ifndef EB_PLASTIC_ORACLE_H
#define EB_PLASTIC_ORACLE_H
#include "eb_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t xnum, ynum;
    const double *gridx_b, *gridy_b;
    const double *plastic_yield, *etas1, *eta_flow, *mus1, *cohesion_grid, *fric_degrees_grid, *sxy2; /* (ynum, xnum) */
    const double *sxx2, *pr1;                                                                       /* (ynum-1, xnum-1) */
    double timestep, timesum;
    int64_t ntimestep;
    double viscosity_min, viscosity_max;
    int64_t iuse_fluid_pressure_for_yield;
    double y_sealevel;
    int64_t no_yielding_in_mobile_wall, no_yielding_in_plate_extension;
    double velocity_internal_x, x_left_mobile_wall, x_right_mobile_wall, y_top_mobile_wall,
        y_bottom_mobile_wall, plate_extension_thickness, plate_extension_width;
} ebo_plastic_in;

/* etavp_new, plastic_strain_rate, plastic_yield_new (ynum, xnum) out;
 * yield_error (ynum, xnum) in/out; returns nnode_yield and global error. */
void ebo_update_nodes_for_plastic_yielding(const ebo_plastic_in *in, double *etavp_new,
                                           double *plastic_strain_rate, double *plastic_yield_new,
                                           double *yield_error, int64_t *nnode_yield,
                                           double *global_yield_error);

void ebo_harmonic_normal_viscosity(int64_t xnum, int64_t ynum, const double *etas1, const double *plastics,
                                   const double *plasticn, const double *plastic_yield, double *etan1);
void ebo_plastic_strain_rate_at_pressure_nodes(int64_t xnum, int64_t ynum, const double *eii_basic,
                                               double *eii_pressure);

/* Marker plastic yielding of the Picard loop over markers. */
typedef struct {
    int64_t marknum;
    const int8_t *flags;
    const double *y;
    const int16_t *matid;
    const double *mat_mu; /* 1-based by matid */
    const double *cohesion, *fric, *exx, *exy, *sr_ratio, *pr;
    int64_t forecast, ntimestep;
    double timestep, y_sealevel, viscosity_min, viscosity_max, yield_stress_min, yield_stress_max;
    int64_t iuse_fluid_pressure_for_yield, no_yielding_in_mobile_wall;
    int64_t matid_sticky_air, matid_mobile_wall, matid_plate_extension;
} ebo_marker_plastic_in;

void ebo_update_marker_plastic_yielding(const ebo_marker_plastic_in *in, double *sxx, double *sxy, double *eta,
                                        float *pfailure);

#ifdef __cplusplus
}
#endif
#endif
