/* TEST INFRASTRUCTURE ONLY: nodal plastic-yield update, restated from
 * rheology/plastic_failure/nodes/PlasticFailureNodes.jl:24-319. */
#include <math.h>

#include "eb_plastic_oracle.h"

/* Julia's max / min on Float64: NaN wins, and +0.0 > -0.0 */
static double jl_max(double a, double b)
{
    if (a != a || b != b) return a + b;
    if (a == b) return signbit(a) ? b : a;
    return a > b ? a : b;
}
static double jl_min(double a, double b)
{
    if (a != a || b != b) return a + b;
    if (a == b) return signbit(a) ? a : b;
    return a < b ? a : b;
}

/* PlasticFailureNodes.jl:222-248 (interpolate_from_pressure_grid_to_basic_grid) */
static double p2b(int64_t i, int64_t j, int64_t ynum, int64_t xnum, const double *a)
{
    const int64_t ld = ynum - 1;
    if (1 < i && i < ynum && 1 < j && j < xnum)
        return (EBO_M(a, ld, i - 1, j - 1) + EBO_M(a, ld, i, j - 1) + EBO_M(a, ld, i - 1, j) + EBO_M(a, ld, i, j)) / 4.0;
    if (i == 1 && j == 1) return EBO_M(a, ld, i, j);
    if (i == ynum && j == 1) return EBO_M(a, ld, i - 1, j);
    if (i == ynum && j == xnum) return EBO_M(a, ld, i - 1, j - 1);
    if (i == 1 && j == xnum) return EBO_M(a, ld, i, j - 1);
    if (j == 1) return (EBO_M(a, ld, i, j) + EBO_M(a, ld, i - 1, j)) / 2.0;
    if (j == xnum) return (EBO_M(a, ld, i - 1, j - 1) + EBO_M(a, ld, i, j - 1)) / 2.0;
    if (i == 1) return (EBO_M(a, ld, i, j) + EBO_M(a, ld, i, j - 1)) / 2.0;
    if (i == ynum) return (EBO_M(a, ld, i - 1, j) + EBO_M(a, ld, i - 1, j - 1)) / 2.0;
    return 0.0;
}

/* YieldStress.jl calc_yield_stress with FluidPressure.calculate_fluid_pressure
 * (density 1000, gravity 9.81) and stress_yield_tension_method2 */
static double yield_stress(double cohesion, double sine_fric, double pressure, int64_t use_fluid, double y,
                           double y_sealevel)
{
    const double phi = asin(sine_fric);
    const double cosphi = cos(phi);
    double s;
    if (use_fluid == 0) {
        s = cohesion * cosphi + sine_fric * pressure;
    } else {
        const double depth = jl_max(0.0, y - y_sealevel);
        const double pf = 1000.0 * 9.81 * depth;
        if (pressure < pf) s = cohesion * cosphi;
        else s = cohesion * cosphi + sine_fric * (pressure - pf);
    }
    return jl_max(s, 0.0);
}

void ebo_update_nodes_for_plastic_yielding(const ebo_plastic_in *in, double *etavp_new,
                                           double *plastic_strain_rate, double *plastic_yield_new,
                                           double *yield_error, int64_t *nnode_yield_out,
                                           double *global_yield_error_out)
{
    const int64_t xnum = in->xnum, ynum = in->ynum, ld = ynum;
    const double dt = in->timestep;
    const double weight = 0.0;
    const double pi_term = M_PI / 180.0;
    int64_t nnode_yield = 0;
    double err_sum = 0.0;
    for (int64_t k = 0; k < xnum * ynum; ++k) {
        etavp_new[k] = in->eta_flow[k];
        plastic_yield_new[k] = 0.0;
        plastic_strain_rate[k] = 0.0;
    }
    for (int64_t j = 1; j <= xnum; ++j)
        for (int64_t i = 1; i <= ynum; ++i) {
            const double yield_marker = EBO_M(in->plastic_yield, ld, i, j);
            const double eta_old = EBO_M(in->etas1, ld, i, j);
            const double eta_flow = EBO_M(in->eta_flow, ld, i, j);
            const double mu = EBO_M(in->mus1, ld, i, j);
            const double cohesion = EBO_M(in->cohesion_grid, ld, i, j);
            const double fric = sin(EBO_M(in->fric_degrees_grid, ld, i, j) * pi_term);
            int applicable = in->ntimestep > 0 && (cohesion > 0.0 || fric > 0.0);
            const double xb = EBO_V(in->gridx_b, j), yb = EBO_V(in->gridy_b, i);
            if (in->no_yielding_in_mobile_wall) {
                const double xl = in->x_left_mobile_wall + in->velocity_internal_x * in->timesum;
                const double xr = in->x_right_mobile_wall + in->velocity_internal_x * in->timesum;
                if (xl <= xb && xb <= xr && in->y_top_mobile_wall <= yb && yb <= in->y_bottom_mobile_wall)
                    applicable = 0;
            }
            if (in->no_yielding_in_plate_extension) {
                const double xl = in->x_left_mobile_wall + in->velocity_internal_x * in->timesum;
                const double xmin = xl - in->plate_extension_width, xmax = xl;
                const double ymin = in->y_bottom_mobile_wall - in->plate_extension_thickness;
                const double ymax = in->y_bottom_mobile_wall;
                if (xmin <= xb && xb <= xmax && ymin <= yb && yb <= ymax) applicable = 0;
            }
            double *ev = &EBO_M(etavp_new, ld, i, j), *py = &EBO_M(plastic_yield_new, ld, i, j);
            if (!applicable) {
                *ev = eta_flow;
                *py = 0.0;
                continue;
            }
            const double pressure = p2b(i, j, ynum, xnum, in->pr1);
            const double sxx = p2b(i, j, ynum, xnum, in->sxx2);
            const double sxy = EBO_M(in->sxy2, ld, i, j);
            const double sii = sqrt(sxy * sxy + sxx * sxx);
            const double sii_e = sii * (mu * dt + eta_old) / eta_old;
            const double sy = yield_stress(cohesion, fric, pressure, in->iuse_fluid_pressure_for_yield, yb,
                                           in->y_sealevel);
            const int at_yield = yield_marker > 0;
            if (at_yield) {
                const double e = sii - sy;
                err_sum = err_sum + e * e;
                EBO_M(yield_error, ld, i, j) = e;
                ++nnode_yield;
            }
            if (sy < sii_e) {
                double eta_new = dt * mu * sy / (sii_e - sy);
                if (eta_new < eta_flow) {
                    *py = 1.0;
                    eta_new = pow(eta_new, 1.0 - weight) * pow(eta_old, weight);
                    eta_new = jl_max(eta_new, in->viscosity_min);
                    eta_new = jl_min(eta_new, in->viscosity_max);
                    *ev = eta_new;
                    EBO_M(plastic_strain_rate, ld, i, j) = 0.5 * sii / eta_new - 0.5 * sii / eta_flow;
                    if (!at_yield) {
                        const double e = sii - sy;
                        err_sum = err_sum + e * e;
                        EBO_M(yield_error, ld, i, j) = e;
                        ++nnode_yield;
                    }
                } else {
                    *ev = eta_flow;
                    *py = 0.0;
                }
            } else {
                *ev = eta_flow;
                *py = 0.0;
            }
        }
    *nnode_yield_out = nnode_yield;
    *global_yield_error_out = nnode_yield > 0 ? sqrt(err_sum / (double)nnode_yield) : 0.0;
}

/* HarmonicAverage.normal_viscosity_from_harmonic_avg_shear_viscosity!
 * (interpolation/utils/stokes_funcs/HarmonicAverage.jl) with
 * CheckPlasticYield.plastic_yielding (tolerance 0) and MathTools.harmonic4
 * (utils/MathTools.jl:42-62, factors 0) */
void ebo_harmonic_normal_viscosity(int64_t xnum, int64_t ynum, const double *etas1, const double *plastics,
                                   const double *plasticn, const double *plastic_yield, double *etan1)
{
    const int64_t ny = ynum;
    for (int64_t i = 1; i <= ynum - 1; ++i)
        for (int64_t j = 1; j <= xnum - 1; ++j) {
            int y = EBO_M(plasticn, ny - 1, i, j) > 0.0;
            if (EBO_M(plastics, ny, i, j) > 0.0 || EBO_M(plastics, ny, i, j + 1) > 0.0 ||
                EBO_M(plastics, ny, i + 1, j) > 0.0 || EBO_M(plastics, ny, i + 1, j + 1) > 0.0)
                y = 1;
            if (EBO_M(plastic_yield, ny, i, j) > 0.0 || EBO_M(plastic_yield, ny, i, j + 1) > 0.0 ||
                EBO_M(plastic_yield, ny, i + 1, j) > 0.0 || EBO_M(plastic_yield, ny, i + 1, j + 1) > 0.0)
                y = 1;
            if (!y) continue;
            const double vinv = 1.0 / EBO_M(etas1, ny, i, j) + 1.0 / EBO_M(etas1, ny, i, j + 1) +
                                1.0 / EBO_M(etas1, ny, i + 1, j) + 1.0 / EBO_M(etas1, ny, i + 1, j + 1) + 0.0 + 0.0 +
                                0.0 + 0.0;
            const double den = 4.0 + 0.0 + 0.0 + 0.0 + 0.0;
            if (vinv != 0.0 && den != 0.0) EBO_M(etan1, ny - 1, i, j) = 1.0 / (vinv / den);
        }
}

/* PlasticStrainRateInterpolation.interpolate_at_pressure_nodes!
 * (interpolation/utils/stokes_funcs/PlasticStrainRateInterpolation.jl:18-41) */
void ebo_plastic_strain_rate_at_pressure_nodes(int64_t xnum, int64_t ynum, const double *eii_basic,
                                               double *eii_pressure)
{
    const int64_t ny = ynum;
    for (int64_t j = 1; j <= xnum - 1; ++j)
        for (int64_t i = 1; i <= ynum - 1; ++i)
            EBO_M(eii_pressure, ny - 1, i, j) = (EBO_M(eii_basic, ny, i, j) + EBO_M(eii_basic, ny, i, j + 1) +
                                                 EBO_M(eii_basic, ny, i + 1, j) + EBO_M(eii_basic, ny, i + 1, j + 1)) /
                                                4.0;
}

/* MarkerPlasticity.update_plastic_yielding! (rheology/plastic_failure/markers/):
 * ViscoelasticForecast.jl:14-127 (forecast 0), ElasticForecast.jl:14-110
 * (forecast 1), YieldCheck.jl:2-64, Limits.jl:2-21, CheckYield.jl:3-12,
 * ViscoelasticFactor.jl:19-30; x^2.0 as x*x. */
void ebo_update_marker_plastic_yielding(const ebo_marker_plastic_in *in, double *sxx, double *sxy, double *eta,
                                        float *pfailure)
{
    for (int64_t m = 0; m < in->marknum; ++m) {
        if (in->flags[m] != 1) continue;
        const int16_t id = in->matid[m];
        const double cohesion = in->cohesion[m], sfric = in->fric[m];
        int flag = 0;
        if (in->ntimestep > 0 && (cohesion > 0.0 || sfric > 0.0)) {
            const double visc = eta[m], mu = EBO_V(in->mat_mu, id);
            const double s_xx = sxx[m], s_xy = sxy[m];
            const double exx = in->exx[m], exy = in->exy[m], ratio = in->sr_ratio[m], dt = in->timestep;
            double sy = yield_stress(cohesion, sfric, in->pr[m], in->iuse_fluid_pressure_for_yield, in->y[m],
                                     in->y_sealevel);
            sy = jl_max(sy, in->yield_stress_min);
            sy = jl_min(sy, in->yield_stress_max);
            double fxx, fxy;
            if (in->forecast == 0) {
                const double z = visc / (visc + dt * mu);
                fxx = s_xx * z + 2.0 * visc * exx * ratio * (1.0 - z);
                fxy = s_xy * z + 2.0 * visc * exy * ratio * (1.0 - z);
            } else {
                fxx = s_xx + 2.0 * mu * dt * exx * ratio;
                fxy = s_xy + 2.0 * mu * dt * exy * ratio;
            }
            const double sinv = sqrt(fxx * fxx + fxy * fxy);
            int yielding;
            if (in->no_yielding_in_mobile_wall)
                yielding = sinv > sy && id != in->matid_sticky_air && id != in->matid_mobile_wall &&
                           id != in->matid_plate_extension;
            else
                yielding = sinv > sy;
            if (yielding) {
                flag = 1;
                double v;
                if (in->forecast == 0) {
                    const double sold = sqrt(s_xx * s_xx + s_xy * s_xy);
                    if (sold != 0.0) {
                        sxx[m] = s_xx * sy / sold;
                        sxy[m] = s_xy * sy / sold;
                    } else {
                        sxx[m] = sy;
                        sxy[m] = sy;
                    }
                    v = sy / 2.0 / (ratio * sqrt(exx * exx + exy * exy));
                } else {
                    v = mu * dt * sy / (sinv - sy);
                }
                v = jl_max(v, in->viscosity_min);
                eta[m] = jl_min(v, in->viscosity_max);
            }
        }
        if (pfailure[m] != 1.0f) pfailure[m] = (float)flag;
    }
}
