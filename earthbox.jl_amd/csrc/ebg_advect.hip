// Marker-side kernels: inside flags, bit-exact cell indexing, the fused RK4
// gather-advect (velocity + spin + location + stress rotation) and grid spin.
//
// Compiled with -ffp-contract=off: every FP64 expression is evaluated in the
// reference's order without fused multiply-adds, and IEEE division is correctly
// rounded, so results are bit-identical to the reference loops (only the
// libm sin/cos of the stress rotation may differ in the last bit).
#include "ebg_common.hpp"
#include "ebg_device.hpp"

namespace ebg {

struct AxisDesc {
    const double* g;   // node coordinates, 0-based storage of a 1-based axis
    int n;             // number of nodes
    double x0, inv_h;  // uniform-axis guess parameters
    int uniform;
};

struct RK4Args {
    int64_t marknum;
    int order;
    double dt;
    double *x, *y;
    const int32_t *xn, *yn;
    const int8_t* flags;
    const double *vx1, *vy1, *esp;
    double *mvx, *mvy, *mspin;
    double *sxx, *sxy;
    int update_location;
    int xnum, ynum;
    AxisDesc ax, ay;
    const double *gridx_vy, *gridy_vx, *xstp_b, *ystp_b, *xstp_vy, *ystp_vx;
    double xmin, xmax, ymin, ymax;
};

// GridToMarker.get_marker_value (interpolation/utils/GridToMarker.jl:190-226):
// bilinear weights, summed UL, LL, UR, LR starting from 0.0.
__device__ __forceinline__ double marker_value(int yn, int xn, double dy, double dx,
                                               const double* __restrict__ a, int ld)
{
    const double wul = (1.0 - dx) * (1.0 - dy);
    const double wll = (1.0 - dx) * dy;
    const double wur = dx * (1.0 - dy);
    const double wlr = dx * dy;
    const double* p = a + (size_t)(xn - 1) * ld + (yn - 1);
    double v = 0.0;
    v += wul * p[0];
    v += wll * p[1];
    v += wur * p[ld];
    v += wlr * p[ld + 1];
    return v;
}

__device__ __forceinline__ int cell_index(double x, const AxisDesc& a)
{
    return left_node_index(x, a.g, a.n, a.x0, a.inv_h, a.uniform);
}

__global__ void __launch_bounds__(256) rk4_advect_kernel(RK4Args a)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.marknum) return;
    const double x0 = a.x[m];
    const double y0 = a.y[m];
    bool inside;
    if (a.flags) inside = (a.flags[m] == 1);
    else inside = (a.xmin <= x0 && x0 <= a.xmax && a.ymin <= y0 && y0 <= a.ymax);
    if (!inside) {  // VelocityAndSpin.jl:120-126; location and stress untouched
        if (a.mvx) a.mvx[m] = 0.0;
        if (a.mvy) a.mvy[m] = 0.0;
        if (a.mspin) a.mspin[m] = 0.0;
        return;
    }
    const int nst = (a.order == 4) ? 4 : 1;
    double xc = x0, yc = y0;
    double svx = 0.0, svy = 0.0, ssp = 0.0;
    double v1x = 0.0, v1y = 0.0, v1s = 0.0;
    double v2x = 0.0, v2y = 0.0, v2s = 0.0;
    double v3x = 0.0, v3y = 0.0, v3s = 0.0;
    for (int k = 1; k <= nst; ++k) {
        int xi, yi;
        if (k == 1 && a.xn) {
            xi = a.xn[m];
            yi = a.yn[m];
        } else {
            xi = cell_index(xc, a.ax);
            yi = cell_index(yc, a.ay);
        }
        // vx grid: x index = basic index, y remapped with strict '>'
        // (MarkerGridMapping.jl:378-409)
        const double gxb = a.ax.g[xi - 1];
        const double hxb = a.xstp_b[xi - 1];
        const double dxv = (xc - gxb) / hxb;
        int iyv = yi;
        if (yc > a.gridy_vx[iyv]) iyv += 1;  // gridy_vx[iyv+1] in 1-based terms
        if (iyv > a.ynum) iyv -= 1;
        const double dyv = (yc - a.gridy_vx[iyv - 1]) / a.ystp_vx[iyv - 1];
        const double vxm = marker_value(iyv, xi, dyv, dxv, a.vx1, a.ynum + 1);
        // vy grid: x remapped with strict '>', then min(xnum) (:413-428)
        int jxv = xi;
        if (xc > a.gridx_vy[jxv]) jxv += 1;
        if (jxv > a.xnum) jxv = a.xnum;
        const double dxy = (xc - a.gridx_vy[jxv - 1]) / a.xstp_vy[jxv - 1];
        const double gyb = a.ay.g[yi - 1];
        const double hyb = a.ystp_b[yi - 1];
        const double dyy = (yc - gyb) / hyb;
        const double vym = marker_value(yi, jxv, dyy, dxy, a.vy1, a.ynum);
        // spin on the basic grid (:402-428)
        const double dxs = (xc - gxb) / hxb;
        const double dys = (yc - gyb) / hyb;
        const double spm = marker_value(yi, xi, dys, dxs, a.esp, a.ynum);
        // RungeKutta.jl:9-47
        if (k == 1) { v1x = vxm; v1y = vym; v1s = spm; }
        else if (k == 2) { v2x = vxm; v2y = vym; v2s = spm; }
        else if (k == 3) { v3x = vxm; v3y = vym; v3s = spm; }
        else {
            svx = (v1x + 2.0 * v2x + 2.0 * v3x + vxm) / 6.0;
            svy = (v1y + 2.0 * v2y + 2.0 * v3y + vym) / 6.0;
            ssp = (v1s + 2.0 * v2s + 2.0 * v3s + spm) / 6.0;
        }
        if (k < 3) {
            xc = x0 + a.dt / 2.0 * vxm;
            yc = y0 + a.dt / 2.0 * vym;
        } else if (k == 3) {
            xc = x0 + a.dt * vxm;
            yc = y0 + a.dt * vym;
        }
    }
    if (nst == 1) { svx = v1x; svy = v1y; ssp = v1s; }
    if (a.mvx) a.mvx[m] = svx;
    if (a.mvy) a.mvy[m] = svy;
    if (a.mspin) a.mspin[m] = ssp;
    if (a.update_location) {  // MarkersLocation.jl:36-41
        a.x[m] = x0 + a.dt * svx;
        a.y[m] = y0 + a.dt * svy;
    }
    if (a.sxx) {  // MarkersStressRotation.jl:60-74
        const double t1 = 2.0 * ssp * a.dt;
        const double t2 = ssp * a.dt;
        double s1, c1, s2, c2;
        sincos(t1, &s1, &c1);
        sincos(t2, &s2, &c2);
        const double sxy_old = a.sxy[m], sxx_old = a.sxx[m];
        a.sxy[m] = sxx_old * s1 + sxy_old * c1;
        a.sxx[m] = sxx_old * (c2 * c2 - s2 * s2) - sxy_old * s1;
    }
}

__global__ void inside_flags_kernel(int64_t n, const double* __restrict__ x,
                                    const double* __restrict__ y, double xmin, double xmax,
                                    double ymin, double ymax, int8_t* __restrict__ f)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double xm = x[m], ym = y[m];
    f[m] = (xmin <= xm && xm <= xmax && ymin <= ym && ym <= ymax) ? 1 : -1;
}

__global__ void cell_indices_kernel(int64_t n, const double* __restrict__ x,
                                    const double* __restrict__ y, const int8_t* __restrict__ f,
                                    AxisDesc ax, AxisDesc ay, int32_t* __restrict__ xn,
                                    int32_t* __restrict__ yn)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    if (f && f[m] != 1) return;
    xn[m] = cell_index(x[m], ax);
    yn[m] = cell_index(y[m], ay);
}

// GridStrainRate.jl:161-182
__global__ void spin_kernel(int xnum, int ynum, const double* __restrict__ xstp_vy,
                            const double* __restrict__ ystp_vx, const double* __restrict__ vx1,
                            const double* __restrict__ vy1, double* __restrict__ esp)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)xnum * ynum) return;
    const int i = (int)(t % ynum) + 1, j = (int)(t / ynum) + 1;
    const double dvy = (vy1[(size_t)j * ynum + (i - 1)] - vy1[(size_t)(j - 1) * ynum + (i - 1)]) / xstp_vy[j - 1];
    const double dvx = (vx1[(size_t)(j - 1) * (ynum + 1) + i] - vx1[(size_t)(j - 1) * (ynum + 1) + (i - 1)]) / ystp_vx[i - 1];
    esp[t] = 0.5 * (dvy - dvx);
}

static AxisDesc axis_x(const Grid2D& g)
{
    return AxisDesc{g.gridx_b, (int)g.xnum, g.x0, g.inv_hx, g.uniform_x};
}
static AxisDesc axis_y(const Grid2D& g)
{
    return AxisDesc{g.gridy_b, (int)g.ynum, g.y0, g.inv_hy, g.uniform_y};
}

static unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace ebg

using namespace ebg;

extern "C" {

int ebg_marker_inside_flags(ebg_ctx* ctx, int64_t marknum, const double* x, const double* y,
                            int8_t* flags)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_marker_inside_flags: no grid set");
    EBG_CHECK(marknum >= 0, EBG_ERR_ARG, "ebg_marker_inside_flags: marknum < 0");
    if (marknum == 0) return EBG_OK;
    EBG_CHECK(x && y && flags, EBG_ERR_ARG, "ebg_marker_inside_flags: NULL array");
    const Grid2D& g = ctx->grid;
    hipLaunchKernelGGL(inside_flags_kernel, dim3(blocks_for(marknum, 256)), dim3(256), 0,
                       ctx->stream, marknum, x, y, g.xmin, g.xmax, g.ymin, g.ymax, flags);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_marker_cell_indices(ebg_ctx* ctx, int64_t marknum, const double* x, const double* y,
                            const int8_t* flags, int32_t* xn, int32_t* yn)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_marker_cell_indices: no grid set");
    EBG_CHECK(marknum >= 0, EBG_ERR_ARG, "ebg_marker_cell_indices: marknum < 0");
    if (marknum == 0) return EBG_OK;
    EBG_CHECK(x && y && xn && yn, EBG_ERR_ARG, "ebg_marker_cell_indices: NULL array");
    hipLaunchKernelGGL(cell_indices_kernel, dim3(blocks_for(marknum, 256)), dim3(256), 0,
                       ctx->stream, marknum, x, y, flags, axis_x(ctx->grid), axis_y(ctx->grid),
                       xn, yn);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_rk4_advect(ebg_ctx* ctx, int64_t marknum, int order, double timestep,
                   double* marker_x, double* marker_y,
                   const int32_t* marker_xn, const int32_t* marker_yn,
                   const int8_t* inside_flags,
                   const double* vx1, const double* vy1, const double* esp,
                   double* marker_vx, double* marker_vy, double* marker_spin,
                   double* marker_sxx, double* marker_sxy, int update_location)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_rk4_advect: no grid set");
    EBG_CHECK(marknum >= 0, EBG_ERR_ARG, "ebg_rk4_advect: marknum < 0");
    if (marknum == 0) return EBG_OK;
    EBG_CHECK(marker_x && marker_y && vx1 && vy1 && esp, EBG_ERR_ARG,
              "ebg_rk4_advect: NULL marker position or grid field");
    EBG_CHECK((marker_xn == nullptr) == (marker_yn == nullptr), EBG_ERR_ARG,
              "ebg_rk4_advect: marker_xn and marker_yn must both be given or both NULL");
    EBG_CHECK((marker_sxx == nullptr) == (marker_sxy == nullptr), EBG_ERR_ARG,
              "ebg_rk4_advect: marker_sxx and marker_sxy must both be given or both NULL");
    EBG_CHECK(std::isfinite(timestep), EBG_ERR_ARG, "ebg_rk4_advect: timestep not finite");
    const Grid2D& g = ctx->grid;
    RK4Args a;
    a.marknum = marknum;
    a.order = order;
    a.dt = timestep;
    a.x = marker_x;
    a.y = marker_y;
    a.xn = marker_xn;
    a.yn = marker_yn;
    a.flags = inside_flags;
    a.vx1 = vx1;
    a.vy1 = vy1;
    a.esp = esp;
    a.mvx = marker_vx;
    a.mvy = marker_vy;
    a.mspin = marker_spin;
    a.sxx = marker_sxx;
    a.sxy = marker_sxy;
    a.update_location = update_location;
    a.xnum = (int)g.xnum;
    a.ynum = (int)g.ynum;
    a.ax = axis_x(g);
    a.ay = axis_y(g);
    a.gridx_vy = g.gridx_vy;
    a.gridy_vx = g.gridy_vx;
    a.xstp_b = g.xstp_b;
    a.ystp_b = g.ystp_b;
    a.xstp_vy = g.xstp_vy;
    a.ystp_vx = g.ystp_vx;
    a.xmin = g.xmin;
    a.xmax = g.xmax;
    a.ymin = g.ymin;
    a.ymax = g.ymax;
    hipLaunchKernelGGL(rk4_advect_kernel, dim3(blocks_for(marknum, 256)), dim3(256), 0,
                       ctx->stream, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_calculate_spin(ebg_ctx* ctx, const double* vx1, const double* vy1, double* esp)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_calculate_spin: no grid set");
    EBG_CHECK(vx1 && vy1 && esp, EBG_ERR_ARG, "ebg_calculate_spin: NULL array");
    const Grid2D& g = ctx->grid;
    const int64_t n = g.xnum * g.ynum;
    hipLaunchKernelGGL(spin_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream,
                       (int)g.xnum, (int)g.ynum, g.xstp_vy, g.ystp_vx, vx1, vy1, esp);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

}  // extern "C"
