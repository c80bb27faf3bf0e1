/*
 * Oracle (test infrastructure only): the reference's marker rheology restated
 * loop for loop, serially over markers:
 *   failure properties  rheology/plastic_failure/failure_properties/FailureProperties.jl
 *                       (calculate_failure_properties_loop!, update_final_friction_angle,
 *                        strain_weakening_or_hardening, calc_cohesion, calc_friction)
 *   flow viscosity      rheology/flow_viscosity/update/types/composite_viscosity/CompositeViscosity.jl
 *                       (update_marker_flow_viscosity!, forecast_viscoelastic_stress,
 *                        update_bisection_limits), rheology/composite_creep/CompositeCreep.jl,
 *                       composite_creep/core/{CreepStrainRate,ExponentialTerm,PeierlsCreep}.jl,
 *                       rheology/viscous_strain_softening/ViscousStrainSoftening.jl
 *   viscoplastic        rheology/viscoplastic_viscosity/ViscoplasticViscosity.jl:21-61
 *   grid -> marker      rheology/plastic_failure/nodes/GridViscoplasticityToMarkers.jl
 *   marker strain       advection/strain/MarkersStrain.jl
 * Build with -ffp-contract=off.
 */
#include "eb_rheology_oracle.h"
#include <math.h>

#define MAT(a, ld, id, col) ((a)[((int64_t)(col) - 1) * (ld) + ((int64_t)(id) - 1)])

static double jmax(double a, double b) { return (b > a) ? b : a; }

void ebo_failure_properties(const ebo_rheology_in *in, double *cohesion_out, double *fric_out)
{
    const int64_t ld = in->nmat;
    for (int64_t m = 0; m < in->marknum; ++m) {
        if (in->flags[m] != 1) continue;
        const int16_t id = in->matid[m];
        const double strain = in->strain_plastic[m];
        const double sfi_marker = in->fric_ini[m];
        const double ci = MAT(in->mat_plastic, ld, id, 1), cf = MAT(in->mat_plastic, ld, id, 2);
        const double sfi_mat = MAT(in->mat_plastic, ld, id, 3), sff_mat = MAT(in->mat_plastic, ld, id, 4);
        const double ei = MAT(in->mat_plastic, ld, id, 5), ef = MAT(in->mat_plastic, ld, id, 6);
        double sff;
        if (sfi_mat > 0.0 && sff_mat > 0.0) {
            double fac = sff_mat / sfi_mat;
            sff = sfi_marker * fac;
        } else {
            sff = sff_mat;
        }
        double coh = ci, sf = sfi_marker;
        if (strain >= ef) {
            coh = cf;
            sf = sff;
        } else if (ei <= strain && strain < ef) {
            coh = ci + (cf - ci) / (ef - ei) * (strain - ei);
            sf = sfi_marker + (sff - sfi_marker) / (ef - ei) * (strain - ei);
        }
        if (in->melt_damage) {
            double md = in->melt_damage[m];
            sf = sf / md;
            coh = coh / md;
        }
        cohesion_out[m] = coh;
        fric_out[m] = sf;
    }
}

/* ExponentialTerm.calculate_activation_energy_term */
static double activation(double Ea, double Va, double R, double T, double P)
{
    double fac1 = 1.0 / (R * T);
    return (Ea * 1000.0 + Va * P) * fac1;
}

typedef struct {
    double Ad, Ead, Vad, nd, Ap, mp, np, sp, Af, Eaf, Vaf;
} law_t;

/* CompositeCreep.calculate_composite_creep_viscosity_and_strain_rate (viscosity
 * only; the stress-independent terms in and out through t[3]) */
static double composite(const law_t *c, double P, double T, double R, double s, int use_old, double t[3])
{
    if (!use_old) t[0] = exp(-activation(c->Ead, c->Vad, R, T, P));
    double st = (c->nd == 1.0) ? s : pow(s, c->nd);
    double e_dis = c->Ad * t[0] * st;
    double e_p = 0.0;
    if (c->Ap > 0.0 && T < 1200.0) {
        if (!use_old) t[1] = activation(c->Ead, c->Vad, R, T, P);
        double ratio = s / c->sp;
        double tm = (c->mp == 1.0) ? 1.0 - ratio : 1.0 - pow(ratio, c->mp);
        double inner = (c->np == 2.0) ? tm * tm : pow(tm, c->np);
        double et = exp(-t[1] * inner);
        e_p = c->Ap * et * (s * s);
    }
    double e_dif = 0.0;
    if (c->Af > 0.0) {
        if (!use_old) t[2] = exp(-activation(c->Eaf, c->Vaf, R, T, P));
        e_dif = c->Af * t[2] * s;
    }
    double e = jmax(e_dis, e_p) + e_dif;
    return 0.5 * (s * 1e6) / e;
}

static double forecast(double dt, double mu, double ratio, double sxx, double sxy, double exx, double exy,
                       double smin, double visc)
{
    double vef = visc / (mu * dt + visc);
    double fac1 = 2.0 * visc * ratio;
    double omv = 1.0 - vef;
    double sxx_n = fac1 * exx * omv + sxx * vef;
    double sxy_n = fac1 * exy * omv + sxy * vef;
    return jmax(sqrt(sxx_n * sxx_n + sxy_n * sxy_n), smin);
}

void ebo_flow_viscosity(const ebo_rheology_in *in, double *eta_flow, double *preexp)
{
    const int64_t ld = in->nmat;
    const double PaToMPa = 1.0 / 1e6;
    for (int64_t m = 0; m < in->marknum; ++m) {
        if (in->flags[m] != 1) continue;
        const int16_t id = in->matid[m];
        const int64_t flow_type = in->mat_flow_type[id - 1];
        double viscosity = 0.0;
        double Ad = MAT(in->mat_flow, ld, id, 2);
        if (flow_type == 0) {
            viscosity = MAT(in->mat_flow, ld, id, 1);
        } else if (flow_type == 1 || flow_type == 2) {
            double sxx = in->sxx[m], sxy = in->sxy[m], exx = in->exx[m], exy = in->exy[m];
            double strain = in->GII[m];
            double P = in->pr[m] * PaToMPa;
            double ratio = in->sr_ratio[m];
            double T = jmax(in->TK[m], in->temperature_top);
            double ei = MAT(in->mat_plastic, ld, id, 5), ef = MAT(in->mat_plastic, ld, id, 6);
            law_t c;
            c.Ead = MAT(in->mat_flow, ld, id, 4);
            c.Vad = MAT(in->mat_flow, ld, id, 5);
            c.nd = MAT(in->mat_flow, ld, id, 3);
            c.Ap = MAT(in->mat_flow, ld, id, 6);
            c.mp = MAT(in->mat_flow, ld, id, 7);
            c.np = MAT(in->mat_flow, ld, id, 8);
            c.sp = MAT(in->mat_flow, ld, id, 9);
            c.Af = MAT(in->mat_flow, ld, id, 10);
            c.Eaf = MAT(in->mat_flow, ld, id, 11);
            c.Vaf = MAT(in->mat_flow, ld, id, 12);
            double mu = in->mat_mu[id - 1];
            if (in->iuse_viscous_strain_soft == 1) {
                double amin = Ad, amax = Ad * in->vsoftfac;
                if (ei < strain && strain < ef) Ad = amin + (amax - amin) / (ef - ei) * (strain - ei);
                if (strain >= ef) Ad = amax;
            }
            c.Ad = Ad;
            double R = in->gas_constant, dt = in->timestep, smin = in->powerlaw_stress_min;
            double si = jmax(sqrt(sxx * sxx + sxy * sxy), smin);
            double t[3] = {0.0, 0.0, 0.0};
            viscosity = composite(&c, P, T, R, si * PaToMPa, 0, t);
            double sf = forecast(dt, mu, ratio, sxx, sxy, exx, exy, smin, viscosity);
            int icount = 0;
            double ds = 1e39;
            while (icount < 20 && ds > 1.0) {
                icount += 1;
                double savg = (si + sf) / 2.0;
                viscosity = composite(&c, P, T, R, savg * PaToMPa, 1, t);
                double sfu = forecast(dt, mu, ratio, sxx, sxy, exx, exy, smin, viscosity);
                if ((sf > si && sfu > savg) || (sf < si && sfu < savg)) si = savg;
                else sf = savg;
                ds = fabs(sf - si);
            }
            /* clamp(viscosity, min, max) */
            viscosity = viscosity > in->viscosity_max ? in->viscosity_max
                                                      : (viscosity < in->viscosity_min ? in->viscosity_min : viscosity);
        }
        eta_flow[m] = viscosity;
        preexp[m] = Ad;
    }
}

void ebo_viscoplastic_viscosity(const ebo_rheology_in *in, const double *eta_flow, double *eta)
{
    for (int64_t m = 0; m < in->marknum; ++m) {
        if (in->flags[m] != 1) continue;
        double vf = eta_flow[m], vp = eta[m];
        if (in->itype_global == 0) {
            vp = vf;
        } else if (in->itype_global == 1) {
            if (in->ntimestep == 0 || vp == 0.0) vp = vf;
        }
        eta[m] = vp;
    }
}

void ebo_marker_viscoplastic_from_grid(int64_t marknum, const int8_t *flags, const int32_t *xn,
                                       const int32_t *yn, const double *dx, const double *dy, int64_t ynum,
                                       const double *plastic_yield, const double *etas1,
                                       const double *eta_flow, double *eta, float *pfailure)
{
    for (int64_t m = 0; m < marknum; ++m) {
        if (flags[m] != 1) continue;
        int32_t x = xn[m], y = yn[m];
        double pul = EBO_M(plastic_yield, ynum, y, x), pll = EBO_M(plastic_yield, ynum, y + 1, x);
        double pur = EBO_M(plastic_yield, ynum, y, x + 1), plr = EBO_M(plastic_yield, ynum, y + 1, x + 1);
        int yielding = pul > 0.0 || pll > 0.0 || pur > 0.0 || plr > 0.0;
        double vflow = eta_flow[m];
        if (yielding) {
            double ddx = dx[m], ddy = dy[m];
            double wul = (1.0 - ddx) * (1.0 - ddy), wll = (1.0 - ddx) * ddy;
            double wur = ddx * (1.0 - ddy), wlr = ddx * ddy;
            double eul = EBO_M(etas1, ynum, y, x), ell = EBO_M(etas1, ynum, y + 1, x);
            double eur = EBO_M(etas1, ynum, y, x + 1), elr = EBO_M(etas1, ynum, y + 1, x + 1);
            double denom = pul / eul * wul + pll / ell * wll + pur / eur * wur + plr / elr * wlr;
            double numer = pul * wul + pll * wll + pur * wur + plr * wlr;
            double vvp = numer / denom;
            if (vvp >= vflow) {
                eta[m] = vflow;
            } else {
                eta[m] = vvp;
                pfailure[m] = 1.0f;
            }
        } else {
            eta[m] = vflow;
        }
    }
}

void ebo_update_marker_strain(int64_t marknum, const int8_t *flags, double timestep, double healing,
                              const double *exx, const double *exy, const double *sxx, const double *sxy,
                              const double *eta, const double *eta_flow, const float *pfailure, double *GII,
                              double *strain_plastic, double *strain_rate_plastic)
{
    for (int64_t m = 0; m < marknum; ++m) { /* update_total_strain! */
        if (flags[m] != 1) continue;
        GII[m] = GII[m] + timestep * sqrt(exx[m] * exx[m] + exy[m] * exy[m]);
    }
    for (int64_t m = 0; m < marknum; ++m) { /* update_plastic_strain! */
        if (flags[m] != 1 || pfailure[m] != 1.0f) continue;
        double sii = sqrt(sxx[m] * sxx[m] + sxy[m] * sxy[m]);
        double psr = 0.5 * sii * (1.0 / eta[m] - 1.0 / eta_flow[m]);
        psr = jmax(0.0, psr);
        double ps = strain_plastic[m] + timestep * psr - healing * timestep;
        ps = jmax(0.0, ps);
        strain_plastic[m] = ps;
        strain_rate_plastic[m] = psr;
    }
}
