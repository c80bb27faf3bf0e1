// Marker rheology on the device: failure properties with strain weakening,
// the composite-creep flow viscosity with the visco-elastic stress forecast,
// the viscoplastic viscosity, the grid -> marker viscoplastic viscosity after
// the Picard loop over nodes, and the marker strain update after advection.
//
// One thread owns one marker and keeps every intermediate in registers: the
// reference's three marker loops of the rheology update (failure properties,
// flow viscosity, viscoplastic viscosity) become one read of the marker
// arrays and one write of each output.  Expressions keep the reference's
// operand order (the library is compiled without contraction); exp and pow
// are the device math library's.
#include "ebg_common.hpp"
#include "ebg_yield.hpp"

#include <cmath>

namespace ebg {

namespace {

constexpr double PA_TO_MPA = 1.0 / 1e6;   // CompositeViscosity.jl: PaToMPa
constexpr double KJ_TO_J = 1000.0;        // ExponentialTerm.jl
constexpr int BISECTION_STEPS = 20;       // CompositeViscosity.jl: max_iterations

// Julia's max/min/clamp on non-NaN values, NaN passed through as Julia does
__device__ __forceinline__ double jmax(double a, double b) { return (b > a) ? b : a; }
__device__ __forceinline__ double jclamp(double x, double lo, double hi)
{
    return x > hi ? hi : (x < lo ? lo : x);
}

struct RheoArgs {
    ebg_marker_rheology_in in;
    ebg_marker_rheology_out out;
};

// ---- FailureProperties.jl ------------------------------------------------
// update_final_friction_angle, strain_weakening_or_hardening, calc_cohesion,
// calc_friction
__device__ __forceinline__ void failure_properties(const RheoArgs& a, int mt, int64_t m, double& cohesion,
                                                   double& sine_fric)
{
    const int64_t ld = a.in.nmat;
    const double* P = a.in.mat_plastic;
    const double strain = a.in.strain_plastic[m];
    const double sf_ini_marker = a.in.fric_ini[m];
    const double c_ini = P[0 * ld + mt], c_fin = P[1 * ld + mt];
    const double sf_ini_mat = P[2 * ld + mt], sf_fin_mat = P[3 * ld + mt];
    const double e_ini = P[4 * ld + mt], e_fin = P[5 * ld + mt];
    double sf_final;
    if (sf_ini_mat > 0.0 && sf_fin_mat > 0.0) {
        const double fac = sf_fin_mat / sf_ini_mat;
        sf_final = sf_ini_marker * fac;
    } else {
        sf_final = sf_fin_mat;
    }
    cohesion = c_ini;
    sine_fric = sf_ini_marker;
    if (strain >= e_fin) {
        cohesion = c_fin;
        sine_fric = sf_final;
    } else if (e_ini <= strain && strain < e_fin) {
        cohesion = c_ini + (c_fin - c_ini) / (e_fin - e_ini) * (strain - e_ini);
        sine_fric = sf_ini_marker + (sf_final - sf_ini_marker) / (e_fin - e_ini) * (strain - e_ini);
    }
    if (a.in.melt_damage) {
        const double md = a.in.melt_damage[m];
        sine_fric = sine_fric / md;
        cohesion = cohesion / md;
    }
}

// ---- CompositeCreep.jl and its core files ---------------------------------
struct CreepLaw {
    double A_dis, Ea_dis, Va_dis, n_dis;
    double A_p, m_p, n_p, s_p;
    double A_dif, Ea_dif, Va_dif;
};

struct CreepTerms {   // the stress-independent terms the bisection reuses
    double exp_dis, act_p, exp_dif;
};

// ExponentialTerm.calculate_activation_energy_term
__device__ __forceinline__ double activation_term(double Ea, double Va, double R, double T, double P_MPa)
{
    const double fac1 = 1.0 / (R * T);
    return (Ea * KJ_TO_J + Va * P_MPa) * fac1;
}

// calculate_composite_creep_viscosity_and_strain_rate: viscosity for the
// stress invariant s (MPa); with use_old the stress-independent terms come
// from t, otherwise they are computed and stored in t.
__device__ double composite_viscosity(const CreepLaw& c, double P_MPa, double T, double R, double s, bool use_old,
                                      CreepTerms& t)
{
    // dislocation creep (CreepStrainRate.calculate_creep_strain_rate)
    if (!use_old) t.exp_dis = exp(-activation_term(c.Ea_dis, c.Va_dis, R, T, P_MPa));
    const double st = (c.n_dis == 1.0) ? s : pow(s, c.n_dis);
    const double e_dis = c.A_dis * t.exp_dis * st;
    // Peierls creep (PeierlsCreep.calculate_peierls_creep_strain_rate), with
    // the dislocation activation energy and volume as the reference passes them
    double e_p = 0.0;
    if (c.A_p > 0.0 && T < 1200.0) {
        if (!use_old) t.act_p = activation_term(c.Ea_dis, c.Va_dis, R, T, P_MPa);
        const double ratio = s / c.s_p;
        const double tm = (c.m_p == 1.0) ? 1.0 - ratio : 1.0 - pow(ratio, c.m_p);
        const double inner = (c.n_p == 2.0) ? tm * tm : pow(tm, c.n_p);
        const double et = exp(-t.act_p * inner);
        e_p = c.A_p * et * (s * s);
    }
    // diffusion creep (stress exponent 1)
    double e_dif = 0.0;
    if (c.A_dif > 0.0) {
        if (!use_old) t.exp_dif = exp(-activation_term(c.Ea_dif, c.Va_dif, R, T, P_MPa));
        e_dif = c.A_dif * t.exp_dif * s;
    }
    const double e = jmax(e_dis, e_p) + e_dif;
    return 0.5 * (s * 1e6) / e;
}

// CompositeViscosity.forecast_viscoelastic_stress
__device__ __forceinline__ double forecast_stress(double dt, double mu, double ratio, double sxx, double sxy,
                                                  double exx, double exy, double smin, double visc)
{
    const double vef = visc / (mu * dt + visc);
    const double fac1 = 2.0 * visc * ratio;
    const double omv = 1.0 - vef;
    const double sxx_n = fac1 * exx * omv + sxx * vef;
    const double sxy_n = fac1 * exy * omv + sxy * vef;
    return jmax(sqrt(sxx_n * sxx_n + sxy_n * sxy_n), smin);
}

// CompositeViscosity.update_marker_flow_viscosity! for one marker
__device__ void flow_viscosity(const RheoArgs& a, int mt, int64_t m, double& visc_out, double& preexp_out)
{
    const ebg_marker_rheology_in& in = a.in;
    const int64_t ld = in.nmat;
    const double* F = in.mat_flow;
    const long long flow_type = in.mat_flow_type[mt];
    double viscosity = 0.0;
    double A_dis = F[1 * ld + mt];
    if (flow_type == 0) {
        viscosity = F[0 * ld + mt];
    } else if (flow_type == 1 || flow_type == 2) {
        const double sxx = in.sxx[m], sxy = in.sxy[m];
        const double exx = in.exx[m], exy = in.exy[m];
        const double strain = in.GII[m];
        const double P_MPa = in.pr[m] * PA_TO_MPA;
        const double ratio = in.sr_ratio[m];
        const double T = jmax(in.TK[m], in.temperature_top);
        const double e_ini = in.mat_plastic[4 * ld + mt], e_fin = in.mat_plastic[5 * ld + mt];
        CreepLaw c;
        c.Ea_dis = F[3 * ld + mt];
        c.Va_dis = F[4 * ld + mt];
        c.n_dis = F[2 * ld + mt];
        c.A_p = F[5 * ld + mt];
        c.m_p = F[6 * ld + mt];
        c.n_p = F[7 * ld + mt];
        c.s_p = F[8 * ld + mt];
        c.A_dif = F[9 * ld + mt];
        c.Ea_dif = F[10 * ld + mt];
        c.Va_dif = F[11 * ld + mt];
        const double mu = in.mat_mu[mt];
        if (in.iuse_viscous_strain_soft == 1) {
            // ViscousStrainSoftening.update_pre_exponential_for_viscous_strain_softening
            const double a_min = A_dis, a_max = A_dis * in.vsoftfac;
            if (e_ini < strain && strain < e_fin)
                A_dis = a_min + (a_max - a_min) / (e_fin - e_ini) * (strain - e_ini);
            if (strain >= e_fin) A_dis = a_max;
        }
        c.A_dis = A_dis;
        const double R = in.gas_constant, dt = in.timestep, smin = in.powerlaw_stress_min;
        double s_ini = jmax(sqrt(sxx * sxx + sxy * sxy), smin);
        CreepTerms t{0.0, 0.0, 0.0};
        viscosity = composite_viscosity(c, P_MPa, T, R, s_ini * PA_TO_MPA, false, t);
        double s_fc = forecast_stress(dt, mu, ratio, sxx, sxy, exx, exy, smin, viscosity);
        int icount = 0;
        double dstress = 1e39;
        while (icount < BISECTION_STEPS && dstress > 1.0) {
            ++icount;
            const double s_avg = (s_ini + s_fc) / 2.0;
            viscosity = composite_viscosity(c, P_MPa, T, R, s_avg * PA_TO_MPA, true, t);
            const double s_fu = forecast_stress(dt, mu, ratio, sxx, sxy, exx, exy, smin, viscosity);
            // update_bisection_limits
            if ((s_fc > s_ini && s_fu > s_avg) || (s_fc < s_ini && s_fu < s_avg)) s_ini = s_avg;
            else s_fc = s_avg;
            dstress = fabs(s_fc - s_ini);
        }
        viscosity = jclamp(viscosity, in.viscosity_min, in.viscosity_max);
    }
    visc_out = viscosity;
    preexp_out = A_dis;
}

__global__ void __launch_bounds__(256) marker_rheology_kernel(RheoArgs a)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.in.marknum) return;
    if (a.in.inside_flags && a.in.inside_flags[m] != 1) return;
    const int mt = (int)a.in.matid[m] - 1;
    if (mt < 0 || mt >= a.in.nmat) return;   // no material row: the marker is left as it is
    if (a.in.stages & EBG_RHEO_FAILURE) {
        double coh, sf;
        failure_properties(a, mt, m, coh, sf);
        a.out.cohesion[m] = coh;
        a.out.fric[m] = sf;
    }
    double eta_flow = 0.0;
    if (a.in.stages & EBG_RHEO_FLOW) {
        double pre;
        flow_viscosity(a, mt, m, eta_flow, pre);
        a.out.eta_flow[m] = eta_flow;
        a.out.preexp[m] = pre;
    }
    if (a.in.stages & EBG_RHEO_VISCOPLASTIC) {
        // ViscoplasticViscosity.get_updated_marker_viscosity
        if (!(a.in.stages & EBG_RHEO_FLOW)) eta_flow = a.out.eta_flow[m];
        const double eta = a.out.eta[m];
        if (a.in.itype_global == 0) {
            a.out.eta[m] = eta_flow;
        } else if (a.in.itype_global == 1) {
            if (a.in.ntimestep == 0 || eta == 0.0) a.out.eta[m] = eta_flow;
        }
    }
}

// ---- GridViscoplasticityToMarkers.jl ----------------------------------------
struct GridVpArgs {
    int64_t marknum;
    int xnum, ynum;
    double xmin, xmax, ymin, ymax;
    const double *gridx_b, *gridy_b, *xstp_b, *ystp_b;
    const double *x, *y;
    const int32_t *xn, *yn;
    const int8_t* flags;
    const double *plastic_yield, *etas1, *eta_flow;
    double* eta;
    float* pfailure;
};

__global__ void __launch_bounds__(256) marker_grid_viscoplasticity_kernel(GridVpArgs a)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.marknum) return;
    const double x = a.x[m], y = a.y[m];
    bool inside;
    if (a.flags) inside = a.flags[m] == 1;
    else inside = (a.xmin <= x && x <= a.xmax && a.ymin <= y && y <= a.ymax);
    if (!inside) return;
    // stored cells are 1..n-1 for inside markers; clamping keeps any other
    // value inside the arrays
    const int xn = min(max((int)a.xn[m], 1), a.xnum - 1);
    const int yn = min(max((int)a.yn[m], 1), a.ynum - 1);
    const int ny = a.ynum;
    const size_t c0 = (size_t)(xn - 1) * ny + (yn - 1), c1 = c0 + ny;
    const double pul = a.plastic_yield[c0], pll = a.plastic_yield[c0 + 1];
    const double pur = a.plastic_yield[c1], plr = a.plastic_yield[c1 + 1];
    const double vflow = a.eta_flow[m];
    if (!(pul > 0.0 || pll > 0.0 || pur > 0.0 || plr > 0.0)) {
        a.eta[m] = vflow;
        return;
    }
    // MarkerGridMapping.calc_normalized_distances_from_upper_left_basic_node!
    const double dx = (x - a.gridx_b[xn - 1]) / a.xstp_b[xn - 1];
    const double dy = (y - a.gridy_b[yn - 1]) / a.ystp_b[yn - 1];
    // GridToMarker.calc_marker_weights
    const double wul = (1.0 - dx) * (1.0 - dy);
    const double wll = (1.0 - dx) * dy;
    const double wur = dx * (1.0 - dy);
    const double wlr = dx * dy;
    // interpolate_grid_viscosity_to_marker
    const double eul = a.etas1[c0], ell = a.etas1[c0 + 1], eur = a.etas1[c1], elr = a.etas1[c1 + 1];
    const double denom = pul / eul * wul + pll / ell * wll + pur / eur * wur + plr / elr * wlr;
    const double numer = pul * wul + pll * wll + pur * wur + plr * wlr;
    const double vvp = numer / denom;
    if (vvp >= vflow) {
        a.eta[m] = vflow;
    } else {
        a.eta[m] = vvp;
        a.pfailure[m] = 1.0f;
    }
}

// ---- MarkersStrain.jl ----------------------------------------------------------
struct StrainArgs {
    int64_t marknum;
    const int8_t* flags;
    double dt, heal;
    const double *exx, *exy, *sxx, *sxy, *eta, *eta_flow;
    const float* pfailure;
    double *GII, *strain_plastic, *strain_rate_plastic;
};

__global__ void __launch_bounds__(256) marker_strain_kernel(StrainArgs a)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.marknum) return;
    if (a.flags && a.flags[m] != 1) return;
    // update_total_strain! / calculate_strain
    const double exx = a.exx[m], exy = a.exy[m];
    a.GII[m] = a.GII[m] + a.dt * sqrt(exx * exx + exy * exy);
    // update_plastic_strain! / calculate_plastic_strain
    if (a.pfailure[m] == 1.0f) {
        const double sxx = a.sxx[m], sxy = a.sxy[m];
        const double sii = sqrt(sxx * sxx + sxy * sxy);
        double psr = 0.5 * sii * (1.0 / a.eta[m] - 1.0 / a.eta_flow[m]);
        psr = jmax(0.0, psr);
        double ps = a.strain_plastic[m] + a.dt * psr - a.heal * a.dt;
        ps = jmax(0.0, ps);
        a.strain_plastic[m] = ps;
        a.strain_rate_plastic[m] = psr;
    }
}

unsigned blocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

// MarkerPlasticity.update_plastic_yielding! (rheology/plastic_failure/markers/):
// ViscoelasticForecast.jl:14-127 (forecast 0) and ElasticForecast.jl:14-110
// (forecast 1) with YieldCheck.jl, Limits.jl, CheckYield.jl:3-12 and
// ViscoelasticFactor.jl:19-30.  The squares of the reference's x^2.0 are
// products.
__global__ void __launch_bounds__(256) marker_plastic_yield_kernel(ebg_marker_plastic_in in,
                                                                   ebg_marker_plastic_out out)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= in.marknum || in.inside_flags[m] != 1) return;
    const int id = in.matid[m];
    const double cohesion = in.cohesion[m], sfric = in.fric[m];
    int flag = 0;
    if (in.ntimestep > 0 && (cohesion > 0.0 || sfric > 0.0)) {
        const double eta = out.eta[m];
        const double mu = (id >= 1 && id <= in.nmat) ? in.mat_mu[id - 1] : __builtin_nan("");
        const double sxx = out.sxx[m], sxy = out.sxy[m];
        const double exx = in.exx[m], exy = in.exy[m], ratio = in.sr_ratio[m];
        const double dt = in.timestep;
        double sy = yield_stress(cohesion, sfric, in.pr[m], in.iuse_fluid_pressure_for_yield, in.y[m], in.y_sealevel);
        sy = jl_min(jl_max(sy, in.yield_stress_min), in.yield_stress_max);
        double fxx, fxy;
        if (in.forecast == 0) {
            const double z = eta / (eta + dt * mu);
            fxx = sxx * z + 2.0 * eta * exx * ratio * (1.0 - z);
            fxy = sxy * z + 2.0 * eta * exy * ratio * (1.0 - z);
        } else {
            fxx = sxx + 2.0 * mu * dt * exx * ratio;
            fxy = sxy + 2.0 * mu * dt * exy * ratio;
        }
        const double sinv = sqrt(fxx * fxx + fxy * fxy);
        bool yielding = sinv > sy;
        if (yielding && in.no_yielding_in_mobile_wall &&
            (id == in.matid_sticky_air || id == in.matid_mobile_wall || id == in.matid_plate_extension))
            yielding = false;
        if (yielding) {
            flag = 1;
            double eta_new;
            if (in.forecast == 0) {
                const double sold = sqrt(sxx * sxx + sxy * sxy);
                if (sold != 0.0) {
                    out.sxx[m] = sxx * sy / sold;
                    out.sxy[m] = sxy * sy / sold;
                } else {
                    out.sxx[m] = sy;
                    out.sxy[m] = sy;
                }
                const double eii = ratio * sqrt(exx * exx + exy * exy);
                eta_new = sy / 2.0 / eii;
            } else {
                eta_new = mu * dt * sy / (sinv - sy);
            }
            out.eta[m] = jl_min(jl_max(eta_new, in.viscosity_min), in.viscosity_max);
        }
    }
    if (out.pfailure[m] != 1.0f) out.pfailure[m] = (float)flag;
}

}  // namespace
}  // namespace ebg

using namespace ebg;

extern "C" int ebg_update_marker_rheology(ebg_ctx* ctx, const ebg_marker_rheology_in* in,
                                          ebg_marker_rheology_out* out)
{
    EBG_CHECK(ctx != nullptr, EBG_ERR_ARG, "ebg_update_marker_rheology: ctx is NULL");
    EBG_CHECK(in && out, EBG_ERR_ARG, "ebg_update_marker_rheology: NULL arguments");
    EBG_CHECK(in->marknum >= 0, EBG_ERR_ARG, "ebg_update_marker_rheology: marknum < 0");
    EBG_CHECK((in->stages & ~(EBG_RHEO_FAILURE | EBG_RHEO_FLOW | EBG_RHEO_VISCOPLASTIC)) == 0 && in->stages != 0,
              EBG_ERR_ARG, "ebg_update_marker_rheology: stages must combine EBG_RHEO_* flags");
    if (in->marknum == 0) return EBG_OK;
    EBG_CHECK(in->matid && in->nmat >= 1, EBG_ERR_ARG, "ebg_update_marker_rheology: matid and nmat >= 1 required");
    if (in->stages & EBG_RHEO_FAILURE)
        EBG_CHECK(in->strain_plastic && in->fric_ini && in->mat_plastic && out->cohesion && out->fric, EBG_ERR_ARG,
                  "ebg_update_marker_rheology: failure stage needs strain_plastic, fric_ini, mat_plastic, "
                  "cohesion, fric");
    if (in->stages & EBG_RHEO_FLOW)
        EBG_CHECK(in->sxx && in->sxy && in->exx && in->exy && in->TK && in->pr && in->GII && in->sr_ratio &&
                      in->mat_flow_type && in->mat_flow && in->mat_plastic && in->mat_mu && out->eta_flow &&
                      out->preexp,
                  EBG_ERR_ARG, "ebg_update_marker_rheology: flow stage needs the marker stress, strain rate, "
                               "TK, pr, GII, sr_ratio arrays, the material tables, eta_flow and preexp");
    if (in->stages & EBG_RHEO_VISCOPLASTIC)
        EBG_CHECK(out->eta && out->eta_flow, EBG_ERR_ARG,
                  "ebg_update_marker_rheology: viscoplastic stage needs eta and eta_flow");
    EBG_CHECK(std::isfinite(in->timestep), EBG_ERR_ARG, "ebg_update_marker_rheology: timestep not finite");
    EBG_HIP(hipSetDevice(ctx->device));
    RheoArgs a{*in, *out};
    hipLaunchKernelGGL(marker_rheology_kernel, dim3(blocks256(in->marknum)), dim3(256), 0, ctx->stream, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

extern "C" int ebg_calc_marker_viscoplastic_viscosity(ebg_ctx* ctx, int64_t marknum, const double* marker_x,
                                                      const double* marker_y, const int32_t* marker_xn,
                                                      const int32_t* marker_yn, const int8_t* inside_flags,
                                                      const double* plastic_yield, const double* etas1,
                                                      const double* marker_eta_flow, double* marker_eta,
                                                      float* marker_pfailure)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_calc_marker_viscoplastic_viscosity: no grid set");
    EBG_CHECK(marknum >= 0, EBG_ERR_ARG, "ebg_calc_marker_viscoplastic_viscosity: marknum < 0");
    if (marknum == 0) return EBG_OK;
    EBG_CHECK(marker_x && marker_y && marker_xn && marker_yn && plastic_yield && etas1 && marker_eta_flow &&
                  marker_eta && marker_pfailure,
              EBG_ERR_ARG, "ebg_calc_marker_viscoplastic_viscosity: NULL array");
    const Grid2D& g = ctx->grid;
    GridVpArgs a;
    a.marknum = marknum;
    a.xnum = (int)g.xnum;
    a.ynum = (int)g.ynum;
    a.xmin = g.xmin;
    a.xmax = g.xmax;
    a.ymin = g.ymin;
    a.ymax = g.ymax;
    a.gridx_b = g.gridx_b;
    a.gridy_b = g.gridy_b;
    a.xstp_b = g.xstp_b;
    a.ystp_b = g.ystp_b;
    a.x = marker_x;
    a.y = marker_y;
    a.xn = marker_xn;
    a.yn = marker_yn;
    a.flags = inside_flags;
    a.plastic_yield = plastic_yield;
    a.etas1 = etas1;
    a.eta_flow = marker_eta_flow;
    a.eta = marker_eta;
    a.pfailure = marker_pfailure;
    hipLaunchKernelGGL(marker_grid_viscoplasticity_kernel, dim3(blocks256(marknum)), dim3(256), 0, ctx->stream, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

extern "C" int ebg_update_marker_strain(ebg_ctx* ctx, int64_t marknum, const int8_t* inside_flags, double timestep,
                                        double plastic_healing_rate, const double* marker_exx,
                                        const double* marker_exy, const double* marker_sxx,
                                        const double* marker_sxy, const double* marker_eta,
                                        const double* marker_eta_flow, const float* marker_pfailure,
                                        double* marker_GII, double* marker_strain_plastic,
                                        double* marker_strain_rate_plastic)
{
    EBG_CHECK(ctx != nullptr, EBG_ERR_ARG, "ebg_update_marker_strain: ctx is NULL");
    EBG_CHECK(marknum >= 0, EBG_ERR_ARG, "ebg_update_marker_strain: marknum < 0");
    if (marknum == 0) return EBG_OK;
    EBG_CHECK(marker_exx && marker_exy && marker_sxx && marker_sxy && marker_eta && marker_eta_flow &&
                  marker_pfailure && marker_GII && marker_strain_plastic && marker_strain_rate_plastic,
              EBG_ERR_ARG, "ebg_update_marker_strain: NULL array");
    EBG_HIP(hipSetDevice(ctx->device));
    StrainArgs a{marknum, inside_flags, timestep, plastic_healing_rate, marker_exx, marker_exy, marker_sxx,
                 marker_sxy, marker_eta, marker_eta_flow, marker_pfailure, marker_GII, marker_strain_plastic,
                 marker_strain_rate_plastic};
    hipLaunchKernelGGL(marker_strain_kernel, dim3(blocks256(marknum)), dim3(256), 0, ctx->stream, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

extern "C" int ebg_update_marker_plastic_yielding(ebg_ctx* ctx, const ebg_marker_plastic_in* in,
                                                  ebg_marker_plastic_out* out)
{
    EBG_CHECK(ctx != nullptr && in != nullptr && out != nullptr, EBG_ERR_ARG,
              "ebg_update_marker_plastic_yielding: NULL arguments");
    EBG_CHECK(in->marknum >= 0, EBG_ERR_ARG, "ebg_update_marker_plastic_yielding: marknum < 0");
    EBG_CHECK(in->forecast == 0 || in->forecast == 1, EBG_ERR_ARG,
              "ebg_update_marker_plastic_yielding: forecast must be 0 (viscoelastic) or 1 (elastic)");
    if (in->marknum == 0) return EBG_OK;
    EBG_CHECK(in->inside_flags && in->y && in->matid && in->mat_mu && in->nmat >= 1 && in->cohesion && in->fric &&
                  in->exx && in->exy && in->sr_ratio && in->pr,
              EBG_ERR_ARG, "ebg_update_marker_plastic_yielding: NULL input array");
    EBG_CHECK(out->sxx && out->sxy && out->eta && out->pfailure, EBG_ERR_ARG,
              "ebg_update_marker_plastic_yielding: NULL output array");
    EBG_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(marker_plastic_yield_kernel, dim3(blocks256(in->marknum)), dim3(256), 0, ctx->stream, *in,
                       *out);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}
