// Marker-side steps of the heat solve and the sub-grid diffusion of
// temperature and stress.
//
// Replaces GridTemperatureToMarkers.interpolate!
// (interpolation/utils/GridTemperatureToMarkers.jl:15-35),
// MarkerTemperature.add_remaining_grid_change_to_marker_temperature! and
// apply_sticky_temperature_correction! (heat_solver/marker_temperature/MarkerTemperature.jl:18-72),
// SubgridDiffusion.subgrid_heat_diffusion_update! with
// remove_subgrid_effects_from_grid_changes! (heat_solver/subgrid_heat/SubgridDiffusion.jl:52-185)
// and SubGridStress.Update.update_subgrid_stress!
// (stokes_continuity_solver/subgrid_stress/update/Update.jl:55-71 with
// SubgridShearStress.jl:25-96, SubgridNormalStress.jl:24-89, DiffusionTerm.jl:13-29,
// RemainingStress.jl:16-23).
//
// One thread per marker forms the basic-cell distances
// (MarkerGridMapping.jl:211-214) and, for the normal stress, the pressure
// cell (MarkerGridMapping.jl:311-350) in registers; the grid values reach the
// marker through the four-term sum of GridToMarker.get_marker_value
// (GridToMarker.jl:190-213).  The relaxed marker differences then go back to
// the grid through the library's own scatter (serial marker order, so the
// node averages equal the reference's loops bit for bit).
#include "ebg_common.hpp"
#include "ebg_krylov.hpp"

#include <cmath>

namespace ebg {

namespace {

__host__ __device__ inline unsigned nbk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

struct MarkerGrid {
    int xnum, ynum;
    double xmin, xmax, ymin, ymax;
    const double *gridx_b, *gridy_b, *xstp_b, *ystp_b;
    const double *gridx_vy, *xstp_vy, *gridy_vx, *ystp_vx;
};

MarkerGrid marker_grid(const Grid2D& g)
{
    MarkerGrid G{};
    G.xnum = (int)g.xnum;
    G.ynum = (int)g.ynum;
    G.xmin = g.xmin;
    G.xmax = g.xmax;
    G.ymin = g.ymin;
    G.ymax = g.ymax;
    G.gridx_b = g.gridx_b;
    G.gridy_b = g.gridy_b;
    G.xstp_b = g.xstp_b;
    G.ystp_b = g.ystp_b;
    G.gridx_vy = g.gridx_vy;
    G.xstp_vy = g.xstp_vy;
    G.gridy_vx = g.gridy_vx;
    G.ystp_vx = g.ystp_vx;
    return G;
}

__device__ __forceinline__ bool marker_inside(const MarkerGrid& G, const ebg_marker_cells& m, int64_t k)
{
    if (m.inside_flags) return m.inside_flags[k] == 1;
    const double x = m.x[k], y = m.y[k];
    return G.xmin <= x && x <= G.xmax && G.ymin <= y && y <= G.ymax;
}

// Stored upper-left cell of an inside marker (clamped into the arrays) and
// its normalized distances.
__device__ __forceinline__ void basic_cell(const MarkerGrid& G, const ebg_marker_cells& m, int64_t k, int* xn,
                                           int* yn, double* dx, double* dy)
{
    int a = m.marker_xn[k], b = m.marker_yn[k];
    a = a < 1 ? 1 : (a > G.xnum - 1 ? G.xnum - 1 : a);
    b = b < 1 ? 1 : (b > G.ynum - 1 ? G.ynum - 1 : b);
    *xn = a;
    *yn = b;
    *dx = (m.x[k] - G.gridx_b[a - 1]) / G.xstp_b[a - 1];
    *dy = (m.y[k] - G.gridy_b[b - 1]) / G.ystp_b[b - 1];
}

// GridToMarker.get_marker_value on a column-major (rows, *) array
__device__ __forceinline__ double grid_value(const double* __restrict__ a, int rows, int yn, int xn, double dy,
                                             double dx)
{
    const double wul = (1.0 - dx) * (1.0 - dy);
    const double wll = (1.0 - dx) * dy;
    const double wur = dx * (1.0 - dy);
    const double wlr = dx * dy;
    const double* c0 = a + (size_t)(xn - 1) * rows + (yn - 1);
    const double* c1 = c0 + rows;
    double v = 0.0;
    v += wul * c0[0];
    v += wll * c0[1];
    v += wur * c1[0];
    v += wlr * c1[1];
    return v;
}

__global__ void __launch_bounds__(256) grid_to_marker_temperature_kernel(MarkerGrid G, ebg_marker_cells m,
                                                                         const double* __restrict__ grid, int mode,
                                                                         double* __restrict__ TK)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.marknum || !marker_inside(G, m, k)) return;
    int xn, yn;
    double dx, dy;
    basic_cell(G, m, k, &xn, &yn, &dx, &dy);
    const double v = grid_value(grid, G.ynum, yn, xn, dy, dx);
    TK[k] = mode ? TK[k] + v : v;
}

__global__ void __launch_bounds__(256) sticky_temperature_kernel(MarkerGrid G, ebg_marker_cells m,
                                                                 const int16_t* __restrict__ matid, int air, int water,
                                                                 double ttop, double* __restrict__ TK)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.marknum || !marker_inside(G, m, k)) return;
    const int id = matid[k];
    if ((id == air || id == water) && TK[k] != ttop) TK[k] = ttop;
}

// SubgridDiffusion.jl:66-115 with :187-197 (timescale) and :218-246 (term)
__global__ void __launch_bounds__(256) subgrid_temperature_kernel(MarkerGrid G, ebg_marker_cells m,
                                                                  const double* __restrict__ tk1,
                                                                  const double* __restrict__ kt1,
                                                                  const double* __restrict__ rhocp1, double dt,
                                                                  double coef, double xstpavg, double ystpavg,
                                                                  double* __restrict__ TK, double* __restrict__ dT)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.marknum) return;
    if (!marker_inside(G, m, k)) {
        dT[k] = 0.0;
        return;
    }
    int xn, yn;
    double dx, dy;
    basic_cell(G, m, k, &xn, &yn, &dx, &dy);
    const double tn = grid_value(tk1, G.ynum, yn, xn, dy, dx);
    const double total = tn - TK[k];
    const double kk = grid_value(kt1, G.ynum, yn, xn, dy, dx);
    const double rc = grid_value(rhocp1, G.ynum, yn, xn, dy, dx);
    const double tsg = rc / kk / (2.0 / (xstpavg * xstpavg) + 2.0 / (ystpavg * ystpavg));
    double term = -coef * dt / tsg;
    term = term > -30.0 ? term : -30.0;
    const double d = total * (1.0 - exp(term));
    dT[k] = d;
    TK[k] += d;
}

__device__ __forceinline__ int pressure_index(int nnodes, int basic, double loc, const double* __restrict__ stag)
{
    int p = (loc < stag[basic]) ? basic - 1 : basic;   // stag[basic] is node basic+1 (1-based)
    if (p < 1) p = 1;
    if (p > nnodes - 2) p = nnodes - 2;
    return p;
}

__device__ __forceinline__ double stress_diffusion(double eta, double mu, double dt, double coef)
{
    const double maxwell = eta / mu;
    double s = -coef * dt / maxwell;
    s = s > -30.0 ? s : -30.0;
    return 1.0 - exp(s);
}

// SubgridShearStress.jl:60-95 and SubgridNormalStress.jl:58-88, marker parts
__global__ void __launch_bounds__(256) subgrid_stress_kernel(MarkerGrid G, ebg_subgrid_stress_in in,
                                                             double* __restrict__ msxy, double* __restrict__ msxx,
                                                             double* __restrict__ dxy, double* __restrict__ dxx)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const ebg_marker_cells& m = in.markers;
    if (k >= m.marknum) return;
    if (!marker_inside(G, m, k)) {
        dxy[k] = 0.0;
        dxx[k] = 0.0;
        return;
    }
    const int id = in.matid[k];
    const double mu = (id >= 1 && id <= in.nmat) ? in.mat_mu[id - 1] : __builtin_nan("");
    const double term = stress_diffusion(in.eta[k], mu, in.timestep, in.subgrid_diff_coef_stress);
    int xn, yn;
    double dx, dy;
    basic_cell(G, m, k, &xn, &yn, &dx, &dy);
    const double sxy_g = grid_value(in.sxy1, G.ynum, yn, xn, dy, dx);
    const double d1 = (sxy_g - msxy[k]) * term;
    msxy[k] += d1;
    dxy[k] = d1;
    const double x = m.x[k], y = m.y[k];
    const int yp = pressure_index(G.ynum, yn, y, G.gridy_vx);
    const int xp = pressure_index(G.xnum, xn, x, G.gridx_vy);
    const double dyp = (y - G.gridy_vx[yp]) / G.ystp_vx[yp];
    const double dxp = (x - G.gridx_vy[xp]) / G.xstp_vy[xp];
    const double sxx_g = grid_value(in.sxx1, G.ynum - 1, yp, xp, dyp, dxp);
    const double d2 = (sxx_g - msxx[k]) * term;
    msxx[k] += d2;
    dxx[k] = d2;
}

__global__ void fill_zero_kernel(double* __restrict__ a, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = 0.0;
}

__global__ void subtract_kernel(double* __restrict__ a, const double* __restrict__ b, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = a[t] - b[t];
}

struct SubgridState {
    DevBuf d1, d2;   // per-marker relaxed differences
};

SubgridState* state(ebg_ctx* ctx)
{
    if (!ctx->subgrid) ctx->subgrid = new SubgridState();
    return (SubgridState*)ctx->subgrid;
}

int check_cells(ebg_ctx* ctx, const ebg_marker_cells* m, const char* who)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, std::string(who) + ": no grid");
    EBG_CHECK(m && m->marknum >= 0, EBG_ERR_ARG, std::string(who) + ": bad marker arguments");
    EBG_CHECK(m->marknum == 0 || (m->x && m->y && m->marker_xn && m->marker_yn), EBG_ERR_ARG,
              std::string(who) + ": NULL marker arrays");
    EBG_CHECK(m->marknum < (int64_t)0xffffffffLL, EBG_ERR_ARG, std::string(who) + ": marknum out of range");
    return EBG_OK;
}

}  // namespace

void subgrid_release(ebg_ctx* ctx)
{
    if (ctx && ctx->subgrid) {
        delete (SubgridState*)ctx->subgrid;
        ctx->subgrid = nullptr;
    }
}

}  // namespace ebg

using namespace ebg;

extern "C" {

int ebg_grid_to_marker_temperature(ebg_ctx* ctx, const ebg_marker_cells* m, const double* grid, int mode,
                                   double* marker_TK)
{
    int rc = check_cells(ctx, m, "ebg_grid_to_marker_temperature");
    if (rc) return rc;
    EBG_CHECK(grid && marker_TK && (mode == 0 || mode == 1), EBG_ERR_ARG,
              "ebg_grid_to_marker_temperature: NULL arrays or mode not 0/1");
    if (m->marknum == 0) return EBG_OK;
    hipLaunchKernelGGL(grid_to_marker_temperature_kernel, dim3(nbk(m->marknum)), dim3(256), 0, ctx->stream,
                       marker_grid(ctx->grid), *m, grid, mode, marker_TK);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_apply_sticky_temperature_correction(ebg_ctx* ctx, const ebg_marker_cells* m, const int16_t* matid,
                                            int matid_sticky_air, int matid_sticky_water, double temperature_top,
                                            double* marker_TK)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_apply_sticky_temperature_correction: no grid");
    EBG_CHECK(m && m->marknum >= 0 && matid && marker_TK, EBG_ERR_ARG,
              "ebg_apply_sticky_temperature_correction: NULL arguments");
    EBG_CHECK(m->marknum == 0 || m->inside_flags || (m->x && m->y), EBG_ERR_ARG,
              "ebg_apply_sticky_temperature_correction: needs inside_flags or x and y");
    if (m->marknum == 0) return EBG_OK;
    hipLaunchKernelGGL(sticky_temperature_kernel, dim3(nbk(m->marknum)), dim3(256), 0, ctx->stream,
                       marker_grid(ctx->grid), *m, matid, matid_sticky_air, matid_sticky_water, temperature_top,
                       marker_TK);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_subgrid_heat_diffusion_update(ebg_ctx* ctx, const ebg_subgrid_heat_in* in, double* marker_TK,
                                      double* dtkn, double* dtk1, double* wtnodes)
{
    int rc = check_cells(ctx, in ? &in->markers : nullptr, "ebg_subgrid_heat_diffusion_update");
    if (rc) return rc;
    EBG_CHECK(in->tk1 && in->kt1 && in->rhocp1 && marker_TK && dtkn && dtk1, EBG_ERR_ARG,
              "ebg_subgrid_heat_diffusion_update: NULL arrays");
    EBG_CHECK(in->timestep > 0 && in->subgrid_diff_coef_temp > 0, EBG_ERR_ARG,
              "ebg_subgrid_heat_diffusion_update: timestep and subgrid_diff_coef_temp must be > 0");
    const Grid2D& g = ctx->grid;
    hipStream_t s = ctx->stream;
    const ebg_marker_cells& m = in->markers;
    SubgridState& S = *state(ctx);
    if ((rc = S.d1.ensure((size_t)std::max<int64_t>(m.marknum, 1) * 8))) return rc;
    double* dT = S.d1.as<double>();
    if (m.marknum > 0)
        hipLaunchKernelGGL(subgrid_temperature_kernel, dim3(nbk(m.marknum)), dim3(256), 0, s, marker_grid(g), m,
                           in->tk1, in->kt1, in->rhocp1, in->timestep, in->subgrid_diff_coef_temp, g.xstpavg,
                           g.ystpavg, marker_TK, dT);
    // dtkn: the inclusive scatter of dT, zero where no weight (SubgridDiffusion.jl:129-176)
    const int64_t nb = g.xnum * g.ynum;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(nbk(nb)), dim3(256), 0, s, dtkn, nb);
    ebg_marker_heat_props hp{};
    hp.marknum = m.marknum;
    hp.x = m.x;
    hp.y = m.y;
    hp.inside_flags = m.inside_flags;
    hp.TK = dT;
    ebg_heat_grid_fields hf{};
    hf.tk1 = dtkn;
    hf.wtnodes = wtnodes;
    if ((rc = ebg_interpolate_marker_heat_parameters_to_grid(ctx, &hp, &hf))) return rc;
    hipLaunchKernelGGL(subtract_kernel, dim3(nbk(nb)), dim3(256), 0, s, dtk1, (const double*)dtkn, nb);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_update_subgrid_stress(ebg_ctx* ctx, const ebg_subgrid_stress_in* in, ebg_subgrid_stress_out* out)
{
    int rc = check_cells(ctx, in ? &in->markers : nullptr, "ebg_update_subgrid_stress");
    if (rc) return rc;
    EBG_CHECK(out && in->matid && in->mat_mu && in->nmat >= 1 && in->eta && in->sxy1 && in->sxx1, EBG_ERR_ARG,
              "ebg_update_subgrid_stress: NULL input arrays");
    EBG_CHECK(out->marker_sxy && out->marker_sxx && out->dsxy && out->dsxx && out->dsxyn && out->dsxxn, EBG_ERR_ARG,
              "ebg_update_subgrid_stress: NULL output arrays");
    EBG_CHECK(in->timestep > 0 && in->subgrid_diff_coef_stress > 0, EBG_ERR_ARG,
              "ebg_update_subgrid_stress: timestep and subgrid_diff_coef_stress must be > 0");
    const Grid2D& g = ctx->grid;
    hipStream_t s = ctx->stream;
    const ebg_marker_cells& m = in->markers;
    SubgridState& S = *state(ctx);
    const size_t mb = (size_t)std::max<int64_t>(m.marknum, 1) * 8;
    if ((rc = S.d1.ensure(mb)) || (rc = S.d2.ensure(mb))) return rc;
    double* dxy = S.d1.as<double>();
    double* dxx = S.d2.as<double>();
    if (m.marknum > 0)
        hipLaunchKernelGGL(subgrid_stress_kernel, dim3(nbk(m.marknum)), dim3(256), 0, s, marker_grid(g), *in,
                           out->marker_sxy, out->marker_sxx, dxy, dxx);
    // dsxyn (exclusive rule, basic grid) and dsxxn (central rule, pressure
    // grid), zero where no weight (Update.jl:57-66, BilinearAverage.jl:234-251)
    const int64_t nb = g.xnum * g.ynum, np = (g.xnum - 1) * (g.ynum - 1);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(nbk(nb)), dim3(256), 0, s, out->dsxyn, nb);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(nbk(np)), dim3(256), 0, s, out->dsxxn, np);
    ebg_marker_stokes_props sp{};
    sp.marknum = m.marknum;
    sp.x = m.x;
    sp.y = m.y;
    sp.inside_flags = m.inside_flags;
    sp.sxy = dxy;
    sp.sxx = dxx;
    ebg_stokes_grid_fields sf{};
    sf.sxy1 = out->dsxyn;
    sf.sxx1 = out->dsxxn;
    sf.wtetas = out->wtetas;
    sf.wtetan = out->wtetan;
    if ((rc = ebg_interpolate_marker_stokes_parameters_to_grid(ctx, &sp, &sf))) return rc;
    // RemainingStress.calculate_remaining_stress_change!
    hipLaunchKernelGGL(subtract_kernel, dim3(nbk(nb)), dim3(256), 0, s, out->dsxy, (const double*)out->dsxyn, nb);
    hipLaunchKernelGGL(subtract_kernel, dim3(nbk(np)), dim3(256), 0, s, out->dsxx, (const double*)out->dsxxn, np);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

}  // extern "C"
