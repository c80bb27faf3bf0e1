/* Oracle (test infrastructure only): heat-equation path and sub-grid
 * diffusion (temperature and stress) of the reference, restated in C. */
#ifndef EB_HEAT_ORACLE_H
#define EB_HEAT_ORACLE_H
#include "eb_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Basic-grid geometry used by the heat path: basic coordinates and steps,
 * pressure-grid steps (the control-volume widths of the heat stencil) and the
 * staggered coordinates of the pressure-cell mapping. */
typedef struct {
    int64_t xnum, ynum;
    const double *gridx_b, *gridy_b, *xstp_b, *ystp_b, *xstp_pr, *ystp_pr;
    const double *gridx_vy, *xstp_vy, *gridy_vx, *ystp_vx;
} ebo_heat_grid;

/* Inclusive basic-grid scatter of nfield marker values with the serial marker
 * order of BilinearNumerator.calc_marker_weight_sums_for_basic_grid_heat!
 * (interpolation/utils/BilinearNumerator.jl:31-141) and the arithmetic
 * average of BilinearAverage.jl:276-294.  Each out[f] holds the backup on entry
 * (the value kept where a node receives no weight) and the average on exit;
 * wtnodes receives the weight sums.  Markers use the upper-left cells xn, yn
 * (1-based) and distances from the basic grid. */
void ebo_scatter_basic_inclusive(const ebo_heat_grid *g, int64_t marknum, const double *x, const double *y,
                                 const int8_t *flags, const int32_t *xn, const int32_t *yn, int nfield,
                                 const double *const *vals, double *const *out, double *wtnodes);

/* Heat-production value of each marker (BilinearNumerator.jl:107-113):
 * mat_hr[matid] plus the serpentinization heat production when serp != NULL. */
void ebo_marker_heat_production(int64_t marknum, const int8_t *flags, const int16_t *matid,
                                const double *mat_hr, const double *serp, double *hr);

/* ApplyThermalBC.apply_thermal_bc_along_boundaries!
 * (interpolation/utils/heat_funcs/ApplyThermalBC.jl:20-44); btopt, bbottomt
 * are (xnum, 2) and bleftt, brightt (ynum, 2), column-major. */
void ebo_apply_thermal_bc(int64_t xnum, int64_t ynum, const double *btopt, const double *bbottomt,
                          const double *bleftt, const double *brightt, double *tk1);

/* HeatSolver.calculate_heat_rhs_terms_grid! (heat_solver/HeatSolver.jl:107-118,
 * rhs/HeatRhs.jl:5-79): RT1 = hr1, plus the adiabatic and shear-heating terms. */
void ebo_heat_rhs(int64_t xnum, int64_t ynum, const double *hr1, const double *ha1, const double *rho1,
                  const double *vx1, const double *vy1, const double *sxy2, const double *etas1,
                  const double *sxx2, const double *etan1, double gravity_x, double gravity_y,
                  int iuse_adiabatic, int iuse_shear, double *RT1);

/* HeatBuildManager.build_system_of_equations
 * (heat_solver/solver/build_system/HeatBuildManager.jl:47-83 with
 * InternalNodes.jl and BoundaryNodes.jl): COO triplets (1-based) in the
 * reference's order and the right-hand side.  Li, Lj, Lv hold 5*xnum*ynum.
 * Returns the number of non-zeros. */
int64_t ebo_heat_build(const ebo_heat_grid *g, const double *btopt, const double *bbottomt,
                       const double *bleftt, const double *brightt, const double *RT1, const double *tk0,
                       const double *rhocp1, const double *kt1, double timestep, int64_t *Li, int64_t *Lj,
                       double *Lv, double *R);

/* HeatResiduals.compute_residuals_temp! (solver/residuals/HeatResiduals.jl:12-162). */
void ebo_heat_residuals(const ebo_heat_grid *g, const double *RT1, const double *rhocp1, const double *tk0,
                        const double *tk2, const double *kt1, double timestep, double *rest);

/* GridTemperatureToMarkers.interpolate! (mode 0, utils/GridTemperatureToMarkers.jl:15-35)
 * and MarkerTemperature.add_remaining_grid_change_to_marker_temperature!
 * (mode 1, heat_solver/marker_temperature/MarkerTemperature.jl:18-45). */
void ebo_grid_to_marker_temperature(const ebo_heat_grid *g, int64_t marknum, const double *x, const double *y,
                                    const int8_t *flags, const int32_t *xn, const int32_t *yn,
                                    const double *grid, int mode, double *marker_TK);

/* SubgridDiffusion.calculate_subgrid_temperature_change_and_correct_marker_temperature!
 * (heat_solver/subgrid_heat/SubgridDiffusion.jl:66-115): per-marker sub-grid
 * temperature change dT (0 outside) and the corrected marker temperature. */
void ebo_subgrid_temperature_markers(const ebo_heat_grid *g, int64_t marknum, const double *x,
                                     const double *y, const int8_t *flags, const int32_t *xn,
                                     const int32_t *yn, const double *tk1, const double *kt1,
                                     const double *rhocp1, double timestep, double coef, double xstpavg,
                                     double ystpavg, double *marker_TK, double *dT);

/* MarkerTemperature.apply_sticky_temperature_correction!
 * (MarkerTemperature.jl:52-72). */
void ebo_sticky_temperature_correction(int64_t marknum, const int8_t *flags, const int16_t *matid,
                                       int16_t matid_air, int16_t matid_water, double temperature_top,
                                       double *marker_TK);

/* SubgridShearStress / SubgridNormalStress marker parts
 * (stokes_continuity_solver/subgrid_stress/update/stress_change/SubgridShearStress.jl:25-96,
 * SubgridNormalStress.jl:24-89, DiffusionTerm.jl:13-29): the relaxed
 * nodal-marker differences (0 outside) and the corrected marker stresses. */
void ebo_subgrid_stress_markers(const ebo_heat_grid *g, int64_t marknum, const double *x, const double *y,
                                const int8_t *flags, const int32_t *xn, const int32_t *yn,
                                const int16_t *matid, const double *mat_mu, const double *eta,
                                const double *sxy1, const double *sxx1, double timestep, double coef,
                                double *marker_sxy, double *marker_sxx, double *dsxy_m, double *dsxx_m);

#ifdef __cplusplus
}
#endif
#endif
