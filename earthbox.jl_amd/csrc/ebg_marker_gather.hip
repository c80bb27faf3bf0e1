// Grid -> marker gathers after a Stokes solve: marker strain rates, the
// stress-based / velocity-based strain-rate ratio and marker pressure, in one
// pass over the markers.
//
// Replaces MarkerStrainRate.calculate_marker_strain_rate!
// (stokes_continuity_solver/utils/strain_rate/MarkerStrainRate.jl:23-69),
// MarkerStrainRateRatio.calculate_marker_strain_rate_ratio!
// (stokes_continuity_solver/utils/strain_rate/MarkerStrainRateRatio.jl:71-188)
// MarkerPressure.calculate_marker_pressure!
// (stokes_continuity_solver/utils/MarkerPressure.jl:17-49) and
// StressCorrection.apply_marker_stress_correction_for_remaining_grid_stress!
// (stokes_continuity_solver/subgrid_stress/stress_correction/StressCorrection.jl:27-83).
// One thread owns
// one marker: it forms the basic-cell distances (MarkerGridMapping.jl:211-214),
// the pressure-cell indices and distances (MarkerGridMapping.jl:311-350) and
// the bilinear sums (interpolation/utils/GridToMarker.jl:190-226) in registers,
// so the reference's three marker loops and its stored marker_dx/dy become one
// read of the marker arrays and one write of each output.  Every expression
// keeps the reference's operand order and IEEE division, so the results are
// the serial loops' bits.
#include "ebg_common.hpp"

#include <cmath>

namespace ebg {

namespace {

struct GatherArgs {
    int xnum, ynum;
    double xmin, xmax, ymin, ymax;
    const double *gridx_b, *gridy_b, *xstp_b, *ystp_b;
    const double *gridx_vy, *xstp_vy, *gridy_vx, *ystp_vx;
    ebg_marker_gather_in in;
    ebg_marker_gather_out out;
    int want_ratio, want_stress;
};

// MarkerGridMapping.upr_left_index_pressure: the staggered node that starts
// the pressure cell holding the point, clamped to [1, nnodes-2] (1-based).
__device__ __forceinline__ int upr_left_index_pressure(int nnodes, int basic, double loc,
                                                       const double* __restrict__ stag)
{
    int p = (loc < stag[basic]) ? basic - 1 : basic;  // stag[basic] is node basic+1 (1-based)
    if (p < 1) p = 1;
    if (p > nnodes - 2) p = nnodes - 2;
    return p;
}

// GridToMarker.get_marker_value on a column-major (rows, *) array, 1-based
// upper-left node (yn, xn): the four weighted terms added to 0.0 in the
// reference's order (upper-left, lower-left, upper-right, lower-right).
__device__ __forceinline__ double bilinear(const double* __restrict__ a, int rows, int yn, int xn,
                                           double dy, double dx)
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

__global__ void __launch_bounds__(256) marker_gather_kernel(GatherArgs a)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.in.marknum) return;
    const double x = a.in.x[m], y = a.in.y[m];
    bool inside;
    if (a.in.inside_flags) inside = a.in.inside_flags[m] == 1;
    else inside = (a.xmin <= x && x <= a.xmax && a.ymin <= y && y <= a.ymax);
    if (!inside) return;
    // stored cells are 1..n-1 for inside markers; clamping keeps any other
    // value inside the arrays
    int xn = a.in.marker_xn[m], yn = a.in.marker_yn[m];
    xn = xn < 1 ? 1 : (xn > a.xnum - 1 ? a.xnum - 1 : xn);
    yn = yn < 1 ? 1 : (yn > a.ynum - 1 ? a.ynum - 1 : yn);
    const double dx = (x - a.gridx_b[xn - 1]) / a.xstp_b[xn - 1];
    const double dy = (y - a.gridy_b[yn - 1]) / a.ystp_b[yn - 1];
    const int ypr = upr_left_index_pressure(a.ynum, yn, y, a.gridy_vx);
    const int xpr = upr_left_index_pressure(a.xnum, xn, x, a.gridx_vy);
    const double dypr = (y - a.gridy_vx[ypr]) / a.ystp_vx[ypr];
    const double dxpr = (x - a.gridx_vy[xpr]) / a.xstp_vy[xpr];
    const int prow = a.ynum - 1, brow = a.ynum;
    if (a.in.exx) {
        const double exx = bilinear(a.in.exx, prow, ypr, xpr, dypr, dxpr);
        const double exy = bilinear(a.in.exy, brow, yn, xn, dy, dx);
        if (a.out.marker_exx) a.out.marker_exx[m] = exx;
        if (a.out.marker_exy) a.out.marker_exy[m] = exy;
        if (a.want_ratio) {
            const double sxx = bilinear(a.in.sxx2, prow, ypr, xpr, dypr, dxpr);
            const double sxy = bilinear(a.in.sxy2, brow, yn, xn, dy, dx);
            const double dsxx = bilinear(a.in.dsxx, prow, ypr, xpr, dypr, dxpr);
            const double dsxy = bilinear(a.in.dsxy, brow, yn, xn, dy, dx);
            if (a.want_stress) {
                a.out.marker_sxy[m] = a.out.marker_sxy[m] + dsxy;
                a.out.marker_sxx[m] = a.out.marker_sxx[m] + dsxx;
            }
            const int id = a.in.matid[m];
            const double mu = (id >= 1 && id <= a.in.nmat) ? a.in.mat_mu[id - 1] : __builtin_nan("");
            const double eta = a.in.eta[m];
            const double eii_v = sqrt(exx * exx + exy * exy);
            double r = 1.0;
            if (eii_v > 0) {
                const double tf = 1.0 / (a.in.timestep * mu);
                const double p = 0.5 * (sxx / eta + dsxx * tf);
                const double q = 0.5 * (sxy / eta + dsxy * tf);
                r = sqrt(p * p + q * q) / eii_v;
            }
            a.out.marker_sr_ratio[m] = r;
        }
    }
    if (a.out.marker_pr) a.out.marker_pr[m] = bilinear(a.in.pr1, prow, ypr, xpr, dypr, dxpr);
    if (a.want_stress && !a.want_ratio) {  // StressCorrection.jl:61-78
        a.out.marker_sxy[m] = a.out.marker_sxy[m] + bilinear(a.in.dsxy, brow, yn, xn, dy, dx);
        a.out.marker_sxx[m] = a.out.marker_sxx[m] + bilinear(a.in.dsxx, prow, ypr, xpr, dypr, dxpr);
    }
}

}  // namespace

}  // namespace ebg

using namespace ebg;

extern "C" int ebg_marker_strain_rate_and_pressure(ebg_ctx* ctx, const ebg_marker_gather_in* in,
                                                   ebg_marker_gather_out* out)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_marker_strain_rate_and_pressure: no grid");
    EBG_CHECK(in && out, EBG_ERR_ARG, "ebg_marker_strain_rate_and_pressure: NULL arguments");
    EBG_CHECK(in->marknum >= 0, EBG_ERR_ARG, "ebg_marker_strain_rate_and_pressure: marknum < 0");
    const bool want_strain = out->marker_exx || out->marker_exy || out->marker_sr_ratio;
    const bool want_stress = out->marker_sxx || out->marker_sxy;
    EBG_CHECK(!want_stress || (out->marker_sxx && out->marker_sxy && in->dsxx && in->dsxy), EBG_ERR_ARG,
              "ebg_marker_strain_rate_and_pressure: the stress correction needs marker_sxx, marker_sxy, "
              "dsxx and dsxy");
    EBG_CHECK(!want_strain || (in->exx && in->exy), EBG_ERR_ARG,
              "ebg_marker_strain_rate_and_pressure: strain-rate outputs need exx and exy");
    EBG_CHECK(!out->marker_pr || in->pr1, EBG_ERR_ARG,
              "ebg_marker_strain_rate_and_pressure: marker_pr needs pr1");
    EBG_CHECK(!out->marker_sr_ratio ||
                  (in->sxx2 && in->sxy2 && in->dsxx && in->dsxy && in->eta && in->matid && in->mat_mu &&
                   in->nmat > 0),
              EBG_ERR_ARG,
              "ebg_marker_strain_rate_and_pressure: marker_sr_ratio needs sxx2, sxy2, dsxx, dsxy, eta, "
              "matid and mat_mu");
    if (in->marknum == 0 || !(want_strain || want_stress || out->marker_pr)) return EBG_OK;
    EBG_CHECK(in->x && in->y && in->marker_xn && in->marker_yn, EBG_ERR_ARG,
              "ebg_marker_strain_rate_and_pressure: NULL marker arrays");
    const Grid2D& g = ctx->grid;
    GatherArgs a{};
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
    a.gridx_vy = g.gridx_vy;
    a.xstp_vy = g.xstp_vy;
    a.gridy_vx = g.gridy_vx;
    a.ystp_vx = g.ystp_vx;
    a.in = *in;
    a.out = *out;
    if (!want_strain) a.in.exx = nullptr;  // pressure only
    a.want_ratio = out->marker_sr_ratio != nullptr;
    a.want_stress = want_stress;
    const unsigned blocks = (unsigned)((in->marknum + 255) / 256);
    hipLaunchKernelGGL(marker_gather_kernel, dim3(blocks), dim3(256), 0, ctx->stream, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}
