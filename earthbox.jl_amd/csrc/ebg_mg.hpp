// Geometric multigrid for the 2-D staggered Stokes-continuity equations
// (the reference's experimental solver, src/stokes_continuity_solver/multigrid/),
// as HIP kernels templated on the working precision T (double: plain FP64;
// float: the mixed-precision preconditioner).
//
//  - smoother: Gauss-Seidel velocity updates v += dR/diag*theta_s and the
//    pressure update p += etan*dRc*theta_c (SolveStokes2dOpt.jl:9-76), coloured
//    by (i+j) parity and phased per colour as vx, then vy, then p so every
//    launch is race-free and deterministic;
//  - residuals (Residuals.jl:101-591, 2-D variants), ghost-node boundary
//    conditions (BoundaryConditions.jl:22-53), pressure normalisation
//    (PressureUpdate.jl:24-37);
//  - restriction (Restriction.jl:377-505), viscosity restriction
//    (ViscosityRestriction.jl:68-298) and prolongation (Prolongation.jl:150-185)
//    in gather form: every coarse node sums its fine contributors in the order
//    the reference's serial scatter would add them, so no atomics are needed and
//    FP64 results equal the serial loops.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace ebg {

template <typename T>
struct MGLevelDev {
    int xnum, ynum;
    const T *xstp_b, *ystp_b, *xstp_vy, *ystp_vx;
    T *etas, *etan;
    T *vx, *vy, *pr;
    // second solution buffers (levels below the finest): the additive tiled
    // smoother writes every tile's update here, then the two are swapped
    T *vx_alt = nullptr, *vy_alt = nullptr, *pr_alt = nullptr;
    T *RX, *RY, *RC;
    T *rx, *ry, *rc;
    // optional fixed-node masks (finest level of the Stokes preconditioner):
    // nodes whose correction is known are neither smoothed nor corrected
    const int8_t *fix_vx = nullptr, *fix_vy = nullptr, *fix_p = nullptr;
};

// Fine->coarse transfer description for one axis of one grid type: for each
// fine node its left coarse node index (1-based) and normalised distance, and
// for each coarse node the contiguous fine range [lo, hi] that contributes.
template <typename T>
struct AxisMapDev {
    const int* idx;   // [n_fine]
    const T* d;       // [n_fine]
    const int* lo;    // [n_coarse]
    const int* hi;    // [n_coarse]
};

template <typename T>
struct TransferDev {
    AxisMapDev<T> vx_y, vx_x, vy_y, vy_x, pr_y, pr_x;  // residual restriction ranges
    AxisMapDev<T> b_y, b_x, e_y, e_x;                  // viscosity restriction ranges
};

#define MG_VX(l, i, j) (l).vx[(size_t)((j) - 1) * ((l).ynum + 1) + ((i) - 1)]
#define MG_VY(l, i, j) (l).vy[(size_t)((j) - 1) * (l).ynum + ((i) - 1)]
#define MG_PR(l, i, j) (l).pr[(size_t)((j) - 1) * ((l).ynum - 1) + ((i) - 1)]
#define MG_ES(l, i, j) (l).etas[(size_t)((j) - 1) * (l).ynum + ((i) - 1)]
#define MG_EN(l, i, j) (l).etan[(size_t)((j) - 1) * ((l).ynum - 1) + ((i) - 1)]

// Field access through a view, so the same residual arithmetic runs on the
// level arrays in HBM (GlobalView) and on LDS tiles (TileView, ebg_mg_tile.hpp).
// Step accessors take the 0-based index into the reference's 1-based arrays.
template <typename T>
struct GlobalView {
    static constexpr bool fast_div = false;
    const MGLevelDev<T>& l;
    __device__ T vx(int i, int j) const { return MG_VX(l, i, j); }
    __device__ T vy(int i, int j) const { return MG_VY(l, i, j); }
    __device__ T pr(int i, int j) const { return MG_PR(l, i, j); }
    __device__ T es(int i, int j) const { return MG_ES(l, i, j); }
    __device__ T en(int i, int j) const { return MG_EN(l, i, j); }
    __device__ T RX(int i, int j) const { return l.RX[(size_t)(j - 1) * (l.ynum + 1) + (i - 1)]; }
    __device__ T RY(int i, int j) const { return l.RY[(size_t)(j - 1) * l.ynum + (i - 1)]; }
    __device__ T RC(int i, int j) const { return l.RC[(size_t)(j - 1) * (l.ynum - 1) + (i - 1)]; }
    __device__ T xb(int k) const { return l.xstp_b[k]; }
    __device__ T yb(int k) const { return l.ystp_b[k]; }
    __device__ T xv(int k) const { return l.xstp_vy[k]; }
    __device__ T yv(int k) const { return l.ystp_vx[k]; }
};

// Division policy of the residual arithmetic: V::fast_div selects a times
// the hardware reciprocal of b for FP32 (the mixed-precision preconditioner:
// two instructions, about 1 ulp, where the IEEE quotient takes ten); every
// other view divides with IEEE rounding, as the reference.
template <class V, typename T>
__device__ __forceinline__ T vdiv(T a, T b)
{
    if constexpr (V::fast_div && sizeof(T) == 4) return a * __builtin_amdgcn_rcpf(b);
    else return a / b;
}

// x-Stokes residual and diagonal at vx(i, j) (Residuals.jl:227-287, 2-D)
template <typename T, class V>
__device__ __forceinline__ T vx_residual_v(const V& v, int i, int j, T* coef)
{
    auto D = [](T a, T b) { return vdiv<V, T>(a, b); };
    const T dxL = v.xb(j - 2), dxR = v.xb(j - 1), dxC = v.xv(j - 1);
    const T dyU = v.yv(i - 2), dyD = v.yv(i - 1), dyC = v.yb(i - 2);
    const T nR = v.en(i - 1, j), nL = v.en(i - 1, j - 1);
    const T sU = v.es(i - 1, j), sD = v.es(i, j);
    const T pL = v.pr(i - 1, j - 1), pR = v.pr(i - 1, j);
    const T vxL = v.vx(i, j - 1), vxC = v.vx(i, j), vxR = v.vx(i, j + 1);
    const T vxU = v.vx(i - 1, j), vxD = v.vx(i + 1, j);
    const T vyUL = v.vy(i - 1, j), vyUR = v.vy(i - 1, j + 1);
    const T vyLL = v.vy(i, j), vyLR = v.vy(i, j + 1);
    const T dp_dx = D(pR - pL, dxC);
    const T dsxx = D(T(2) * (D(nR * (vxR - vxC), dxR) - D(nL * (vxC - vxL), dxL)), dxC);
    const T dsxy = D(sD * (D(vxD - vxC, dyD) + D(vyLR - vyLL, dxC)) -
                     sU * (D(vxC - vxU, dyU) + D(vyUR - vyUL, dxC)), dyC);
    const T dR = v.RX(i, j) + dp_dx - dsxx - dsxy;
    if (coef) *coef = T(-2) * (D(D(nR, dxR), dxC) + D(D(nL, dxL), dxC)) - (D(D(sD, dyD), dyC) + D(D(sU, dyU), dyC));
    return dR;
}

// y-Stokes residual and diagonal at vy(i, j) (Residuals.jl:373-433, 2-D)
template <typename T, class V>
__device__ __forceinline__ T vy_residual_v(const V& v, int i, int j, T* coef)
{
    auto D = [](T a, T b) { return vdiv<V, T>(a, b); };
    const T dxL = v.xb(j - 2), dxR = v.xv(j - 1), dxC = v.xv(j - 2);
    const T dyU = v.yb(i - 2), dyD = v.yb(i - 1), dyC = v.yv(i - 1);
    const T pD = v.pr(i, j - 1), pU = v.pr(i - 1, j - 1);
    const T nD = v.en(i, j - 1), nU = v.en(i - 1, j - 1);
    const T sR = v.es(i, j), sL = v.es(i, j - 1);
    const T vyU = v.vy(i - 1, j), vyC = v.vy(i, j), vyD = v.vy(i + 1, j);
    const T vyR = v.vy(i, j + 1), vyL = v.vy(i, j - 1);
    const T vxLR = v.vx(i + 1, j), vxUR = v.vx(i, j);
    const T vxLL = v.vx(i + 1, j - 1), vxUL = v.vx(i, j - 1);
    const T dp_dy = D(pD - pU, dyC);
    const T dsyy = D(T(2) * (D(nD * (vyD - vyC), dyD) - D(nU * (vyC - vyU), dyU)), dyC);
    const T dsyx = D(sR * (D(vyR - vyC, dxR) + D(vxLR - vxUR, dyC)) -
                     sL * (D(vyC - vyL, dxL) + D(vxLL - vxUL, dyC)), dxC);
    const T dR = v.RY(i, j) + dp_dy - dsyy - dsyx;
    if (coef) *coef = T(-2) * (D(D(nD, dyD), dyC) + D(D(nU, dyU), dyC)) - (D(D(sR, dxR), dxC) + D(D(sL, dxL), dxC));
    return dR;
}

// continuity residual at pr(i, j) (Residuals.jl:562-591, 2-D)
template <typename T, class V>
__device__ __forceinline__ T pr_residual_v(const V& v, int i, int j)
{
    auto D = [](T a, T b) { return vdiv<V, T>(a, b); };
    const T dxC = v.xb(j - 1), dyC = v.yb(i - 1);
    return v.RC(i, j) - (D(v.vx(i + 1, j + 1) - v.vx(i + 1, j), dxC) +
                         D(v.vy(i + 1, j + 1) - v.vy(i, j + 1), dyC));
}

template <typename T>
__device__ __forceinline__ T mg_vx_residual(const MGLevelDev<T>& l, int i, int j, T* coef)
{
    return vx_residual_v<T>(GlobalView<T>{l}, i, j, coef);
}
template <typename T>
__device__ __forceinline__ T mg_vy_residual(const MGLevelDev<T>& l, int i, int j, T* coef)
{
    return vy_residual_v<T>(GlobalView<T>{l}, i, j, coef);
}
template <typename T>
__device__ __forceinline__ T mg_pr_residual(const MGLevelDev<T>& l, int i, int j)
{
    return pr_residual_v<T>(GlobalView<T>{l}, i, j);
}

// ---------------------------------------------------------------- kernels
template <typename T>
__global__ void mg_bc_kernel(MGLevelDev<T> l, T d_top, T d_bottom, T d_left, T d_right)
{
    // One thread per boundary node.  The serial reference pass (columns outer,
    // rows inner) lets the top corner ghosts read the pre-pass value of the
    // node below them and the later corners the already-zeroed neighbour; the
    // threads that own those corners also own the neighbour they read, so the
    // result equals the serial pass and no two threads touch the same node.
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int xnum = l.xnum, ynum = l.ynum;
    if (t < xnum) {  // vx top/bottom ghosts of column j
        const int j = t + 1;
        if (j == 1 || j == xnum) {
            const T old = MG_VX(l, 2, j);
            MG_VX(l, 1, j) = d_top * old;
            MG_VX(l, 2, j) = T(0);
            MG_VX(l, ynum + 1, j) = d_bottom * T(0);
        } else {
            MG_VX(l, 1, j) = d_top * MG_VX(l, 2, j);
            MG_VX(l, ynum + 1, j) = d_bottom * MG_VX(l, ynum, j);
        }
        return;
    }
    int u = t - xnum;
    if (u < ynum - 2) {  // vx left/right boundary columns, rows 3..ynum
        const int i = u + 3;
        MG_VX(l, i, 1) = T(0);
        MG_VX(l, i, xnum) = T(0);
        return;
    }
    u -= (ynum - 2);
    if (u < ynum) {  // vy left/right ghosts of row i
        const int i = u + 1;
        if (i == 1 || i == ynum) {
            const T old = MG_VY(l, i, 2);
            MG_VY(l, i, 1) = d_left * old;
            MG_VY(l, i, 2) = T(0);
            MG_VY(l, i, xnum + 1) = d_right * T(0);
        } else {
            MG_VY(l, i, 1) = d_left * MG_VY(l, i, 2);
            MG_VY(l, i, xnum + 1) = d_right * MG_VY(l, i, xnum);
        }
        return;
    }
    u -= ynum;
    if (u < xnum - 2) {  // vy top/bottom boundary rows, columns 3..xnum
        const int j = u + 3;
        MG_VY(l, 1, j) = T(0);
        MG_VY(l, ynum, j) = T(0);
    }
}
inline int mg_bc_threads(int xnum, int ynum) { return xnum + (ynum - 2) + ynum + (xnum - 2); }

// PHASE 0: vx, 1: vy, 2: p; interior nodes of colour (i+j)&1 == color
template <typename T, int PHASE>
__global__ void __launch_bounds__(256) mg_sweep_kernel(MGLevelDev<T> l, int color, T theta_s, T theta_c)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (PHASE == 0) {  // i in 2..ynum, j in 2..xnum-1
        const int ni = l.ynum - 1;
        if (t >= ni * (l.xnum - 2)) return;
        const int i = 2 + t % ni, j = 2 + t / ni;
        if (((i + j) & 1) != color) return;
        if (l.fix_vx && l.fix_vx[(size_t)(j - 1) * (l.ynum + 1) + (i - 1)]) return;
        T coef;
        const T dR = mg_vx_residual(l, i, j, &coef);
        MG_VX(l, i, j) += dR / coef * theta_s;
    } else if (PHASE == 1) {  // i in 2..ynum-1, j in 2..xnum
        const int ni = l.ynum - 2;
        if (t >= ni * (l.xnum - 1)) return;
        const int i = 2 + t % ni, j = 2 + t / ni;
        if (((i + j) & 1) != color) return;
        if (l.fix_vy && l.fix_vy[(size_t)(j - 1) * l.ynum + (i - 1)]) return;
        T coef;
        const T dR = mg_vy_residual(l, i, j, &coef);
        MG_VY(l, i, j) += dR / coef * theta_s;
    } else {
        const int ni = l.ynum - 1;
        if (t >= ni * (l.xnum - 1)) return;
        const int i = 1 + t % ni, j = 1 + t / ni;
        if (((i + j) & 1) != color) return;
        if (l.fix_p && l.fix_p[(size_t)(j - 1) * (l.ynum - 1) + (i - 1)]) return;
        const T dR = mg_pr_residual(l, i, j);
        MG_PR(l, i, j) += MG_EN(l, i, j) * dR * theta_c;
    }
}

// dp = prnorm - pr[1,1] computed once into *scratch, then added everywhere
template <typename T>
__global__ void mg_pressure_shift_kernel(const T* pr, const T* prnorm_dev, T prnorm, T* scratch)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) scratch[0] = (prnorm_dev ? prnorm_dev[0] : prnorm) - pr[0];
}
template <typename T>
__global__ void mg_pressure_add_kernel(T* pr, int n, const T* scratch)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) pr[t] += scratch[0];
}

// Level residuals at entry t of the (ynum+1) x (xnum+1) index space (the vx
// grid; vy and p handled by the same thread where their ranges overlap)
template <typename T, class V>
__device__ __forceinline__ void mg_residual_at(const MGLevelDev<T>& l, const V& v, int t)
{
    const int xnum = l.xnum, ynum = l.ynum;
    const int ni = ynum + 1;
    if (t >= ni * (xnum + 1)) return;
    const int i = 1 + t % ni, j = 1 + t / ni;
    if (j <= xnum) {
        T r = T(0);
        if (i >= 2 && i <= ynum && j >= 2 && j <= xnum - 1) r = vx_residual_v<T>(v, i, j, (T*)nullptr);
        if (l.fix_vx && l.fix_vx[(size_t)(j - 1) * (ynum + 1) + (i - 1)]) r = T(0);
        l.rx[(size_t)(j - 1) * (ynum + 1) + (i - 1)] = r;
    }
    if (i <= ynum) {
        T r = T(0);
        if (i >= 2 && i <= ynum - 1 && j >= 2 && j <= xnum) r = vy_residual_v<T>(v, i, j, (T*)nullptr);
        if (l.fix_vy && l.fix_vy[(size_t)(j - 1) * ynum + (i - 1)]) r = T(0);
        l.ry[(size_t)(j - 1) * ynum + (i - 1)] = r;
    }
    if (i <= ynum - 1 && j <= xnum - 1) {
        const size_t k = (size_t)(j - 1) * (ynum - 1) + (i - 1);
        l.rc[k] = (l.fix_p && l.fix_p[k]) ? T(0) : pr_residual_v<T>(v, i, j);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) mg_residual_kernel(MGLevelDev<T> l)
{
    mg_residual_at<T>(l, GlobalView<T>{l}, blockIdx.x * blockDim.x + threadIdx.x);
}

// The level arrays with a pending gauge shift added to every pressure read
template <typename T>
struct ShiftedView : GlobalView<T> {
    T shift;
    __device__ T pr(int i, int j) const { return MG_PR(this->l, i, j) + shift; }
};

// The gauge shift p += shift[0] followed by the residuals, in one pass: the
// residual arithmetic reads p + shift (the values the shift would store) and
// each thread writes its shifted pressure to pnew, which the caller swaps in
// for p.  Same values, bit for bit, as the separate shift and residual passes.
template <typename T>
__global__ void __launch_bounds__(256) mg_shift_residual_kernel(MGLevelDev<T> l, const T* __restrict__ shift,
                                                                T* __restrict__ pnew)
{
    const T s = shift[0];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    mg_residual_at<T>(l, ShiftedView<T>{{l}, s}, t);
    const int ni = l.ynum + 1;
    if (t < ni * (l.xnum + 1)) {
        const int i = 1 + t % ni, j = 1 + t / ni;
        if (i <= l.ynum - 1 && j <= l.xnum - 1) {
            const size_t k = (size_t)(j - 1) * (l.ynum - 1) + (i - 1);
            pnew[k] = l.pr[k] + s;
        }
    }
}

// Gather restriction of one grid type.  val(i,j) is read from fine array f
// (scaled by etan for the continuity residual).  Loop order j-outer, i-inner
// matches Restriction.jl:444-463.
template <typename T, int KIND>  // 0 vx, 1 vy, 2 pr
__device__ __forceinline__ void mg_restrict_at(const MGLevelDev<T>& f, const MGLevelDev<T>& c,
                                               const AxisMapDev<T>& my, const AxisMapDev<T>& mx, int t)
{
    const int yc = c.ynum, xc = c.xnum;
    int nyc, nxc;
    if (KIND == 0) { nyc = yc + 1; nxc = xc; }
    else if (KIND == 1) { nyc = yc; nxc = xc + 1; }
    else { nyc = yc - 1; nxc = xc - 1; }
    if (t >= nyc * nxc) return;
    const int ic = 1 + t % nyc, jc = 1 + t / nyc;
    bool interior;
    if (KIND == 0) interior = (ic > 1 && ic < yc + 1 && jc > 1 && jc < xc);
    else if (KIND == 1) interior = (ic > 1 && ic < yc && jc > 1 && jc < xc + 1);
    else interior = true;
    T num = T(0), den = T(0);
    if (interior) {
        const int ilo = my.lo[ic - 1], ihi = my.hi[ic - 1];
        const int jlo = mx.lo[jc - 1], jhi = mx.hi[jc - 1];
        for (int j = jlo; j <= jhi; ++j) {
            const int jx = mx.idx[j - 1];
            const T dx = mx.d[j - 1];
            const T wx = (jx == jc) ? (T(1) - dx) : dx;
            for (int i = ilo; i <= ihi; ++i) {
                const int iy = my.idx[i - 1];
                const T dy = my.d[i - 1];
                const T wy = (iy == ic) ? (T(1) - dy) : dy;
                const T w = wx * wy;
                T val;
                if (KIND == 0) val = f.rx[(size_t)(j - 1) * (f.ynum + 1) + (i - 1)];
                else if (KIND == 1) val = f.ry[(size_t)(j - 1) * f.ynum + (i - 1)];
                else val = MG_EN(f, i, j) * f.rc[(size_t)(j - 1) * (f.ynum - 1) + (i - 1)];
                num += w * val;
                den += w;
            }
        }
    }
    if (KIND == 0) c.RX[(size_t)(jc - 1) * (yc + 1) + (ic - 1)] = (interior && den != T(0)) ? num / den : T(0);
    else if (KIND == 1) c.RY[(size_t)(jc - 1) * yc + (ic - 1)] = (interior && den != T(0)) ? num / den : T(0);
    else c.RC[(size_t)(jc - 1) * (yc - 1) + (ic - 1)] = (den != T(0)) ? num / den / MG_EN(c, ic, jc) : T(0);
}
template <typename T, int KIND>
__global__ void __launch_bounds__(256) mg_restrict_kernel(MGLevelDev<T> f, MGLevelDev<T> c,
                                                           AxisMapDev<T> my, AxisMapDev<T> mx)
{
    mg_restrict_at<T, KIND>(f, c, my, mx, blockIdx.x * blockDim.x + threadIdx.x);
}

// The three restrictions of mg_restrict_kernel in one launch (threads over
// the coarse vx, then vy, then pressure grids, the same arithmetic per
// entry), each thread also clearing its entry of the coarse solution, which
// the V-cycle's smoothing of that level starts from.
template <typename T>
__global__ void __launch_bounds__(256) mg_restrict_zero_kernel(MGLevelDev<T> f, MGLevelDev<T> c, TransferDev<T> x)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nvx = (c.ynum + 1) * c.xnum, nvy = c.ynum * (c.xnum + 1), np = (c.ynum - 1) * (c.xnum - 1);
    if (t < nvx) {
        mg_restrict_at<T, 0>(f, c, x.vx_y, x.vx_x, t);
        c.vx[t] = T(0);
        return;
    }
    t -= nvx;
    if (t < nvy) {
        mg_restrict_at<T, 1>(f, c, x.vy_y, x.vy_x, t);
        c.vy[t] = T(0);
        return;
    }
    t -= nvy;
    if (t < np) {
        mg_restrict_at<T, 2>(f, c, x.pr_y, x.pr_x, t);
        c.pr[t] = T(0);
    }
}

// Gather viscosity restriction; loop order i-outer, j-inner as
// ViscosityRestriction.jl:202-215.  KIND 0: etas (basic), 1: etan (pressure).
template <typename T, int KIND>
__global__ void mg_visc_restrict_kernel(MGLevelDev<T> f, MGLevelDev<T> c, AxisMapDev<T> my,
                                        AxisMapDev<T> mx)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nyc = KIND == 0 ? c.ynum : c.ynum - 1;
    const int nxc = KIND == 0 ? c.xnum : c.xnum - 1;
    if (t >= nyc * nxc) return;
    const int ic = 1 + t % nyc, jc = 1 + t / nyc;
    const int ilo = my.lo[ic - 1], ihi = my.hi[ic - 1];
    const int jlo = mx.lo[jc - 1], jhi = mx.hi[jc - 1];
    T num = T(0), den = T(0);
    for (int i = ilo; i <= ihi; ++i) {
        const int iy = my.idx[i - 1];
        const T dy = my.d[i - 1];
        const T wy = (iy == ic) ? (T(1) - dy) : dy;
        for (int j = jlo; j <= jhi; ++j) {
            const int jx = mx.idx[j - 1];
            const T dx = mx.d[j - 1];
            const T wx = (jx == jc) ? (T(1) - dx) : dx;
            const T w = wx * wy;
            const T val = KIND == 0 ? MG_ES(f, i, j) : MG_EN(f, i, j);
            num += w * val;
            den += w;
        }
    }
    const T out = (den != T(0)) ? num / den : T(0);
    if (KIND == 0) MG_ES(c, ic, jc) = out;
    else MG_EN(c, ic, jc) = out;
}

template <typename T>
__device__ __forceinline__ T mg_c2f(const T* a, int ld, int iy, int jx, T dy, T dx)
{
    const T omx = T(1) - dx, omy = T(1) - dy;
    const T w00 = omx * omy, w01 = omx * dy, w10 = dx * omy, w11 = dx * dy;
    const T* p = a + (size_t)(jx - 1) * ld + (iy - 1);
    T v = T(0);
    v += w00 * p[0];
    v += w01 * p[1];
    v += w10 * p[ld];
    v += w11 * p[ld + 1];
    return v;
}

template <typename T>
__global__ void __launch_bounds__(256) mg_prolong_kernel(MGLevelDev<T> c, MGLevelDev<T> f,
                                                          TransferDev<T> m, T relax_v, T relax_p)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int xf = f.xnum, yf = f.ynum;
    const int ni = yf + 1;
    if (t >= ni * (xf + 1)) return;
    const int i = 1 + t % ni, j = 1 + t / ni;
    if (j < xf + 1) {
        const T dv = mg_c2f(c.vx, c.ynum + 1, m.vx_y.idx[i - 1], m.vx_x.idx[j - 1],
                            m.vx_y.d[i - 1], m.vx_x.d[j - 1]);
        if (!(f.fix_vx && f.fix_vx[(size_t)(j - 1) * (f.ynum + 1) + (i - 1)])) MG_VX(f, i, j) += dv * relax_v;
    }
    if (i < yf + 1) {
        const T dv = mg_c2f(c.vy, c.ynum, m.vy_y.idx[i - 1], m.vy_x.idx[j - 1],
                            m.vy_y.d[i - 1], m.vy_x.d[j - 1]);
        if (!(f.fix_vy && f.fix_vy[(size_t)(j - 1) * f.ynum + (i - 1)])) MG_VY(f, i, j) += dv * relax_v;
    }
    if (i < yf && j < xf) {
        const T dp = mg_c2f(c.pr, c.ynum - 1, m.pr_y.idx[i - 1], m.pr_x.idx[j - 1],
                            m.pr_y.d[i - 1], m.pr_x.d[j - 1]);
        if (!(f.fix_p && f.fix_p[(size_t)(j - 1) * (f.ynum - 1) + (i - 1)])) MG_PR(f, i, j) += dp * relax_p;
    }
}

template <typename T>
__global__ void fill_kernel(T* a, int64_t n, T v)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// ------------------------------------------------------- host-side hierarchy
struct HostAxis {
    std::vector<double> g;  // node coordinates of one axis of one grid type
};

// Host description of one uniform level (cumulative-sum coordinates as the
// reference's Initialization.jl:27-43 builds them).
struct HostLevelGrid {
    int xnum, ynum;
    std::vector<double> gridx_b, gridy_b, xstp_b, ystp_b, gridx_vy, xstp_vy, gridy_vx,
        ystp_vx, gridx_pr, gridy_pr;
};

HostLevelGrid make_uniform_level_grid(int xnum, int ynum, double xsize, double ysize);

// One axis of a uniform level grid: basic nodes by cumulative sum
// (Initialization.jl:27-43), staggered nodes with ghosts and their steps
// (VxGridTools.jl:40-99), pressure nodes (PressureGridTools.jl:34-86).
void basic_axis(int n, double step, std::vector<double>& g, std::vector<double>& stp);
void staggered_axis(int n, const std::vector<double>& g, const std::vector<double>& stp,
                    std::vector<double>& gs, std::vector<double>& stps);
void pressure_axis(int n, const std::vector<double>& g, std::vector<double>& gp);

// clamp(searchsortedlast(g, x), 1, n-1) (MarkerGridMapping.jl:78-86)
inline int left_index_host(double x, const std::vector<double>& g)
{
    const int n = (int)g.size();
    int lo = 0, hi = n + 1;
    while (lo < hi - 1) {
        const int m = lo + (hi - lo) / 2;
        if (x < g[m - 1]) hi = m;
        else lo = m;
    }
    return lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
}

template <typename T>
struct HostAxisMap {
    std::vector<int> idx, lo, hi;
    std::vector<T> d;
};

// GridMappingManager.jl:289-338 for one axis, plus the contiguous fine range
// [lo, hi] within [rlo, rhi] that feeds each coarse node.
template <typename T>
HostAxisMap<T> axis_map(const std::vector<double>& ga, const std::vector<double>& gb, int rlo, int rhi,
                        bool clamp)
{
    HostAxisMap<T> m;
    const int na = (int)ga.size(), nb = (int)gb.size();
    m.idx.resize(na);
    m.d.resize(na);
    for (int k = 0; k < na; ++k) {
        const int ib = left_index_host(ga[k], gb);
        const double x0 = gb[ib - 1];
        const double w = gb[ib] - x0;
        m.idx[k] = ib;
        double d = (ga[k] - x0) / w;
        if (clamp) d = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
        m.d[k] = (T)d;
    }
    m.lo.assign(nb, 1);
    m.hi.assign(nb, 0);
    for (int kc = 1; kc <= nb; ++kc) {
        int lo = -1, hi = -2;
        for (int i = rlo; i <= rhi; ++i) {
            const int ix = m.idx[i - 1];
            if (ix == kc - 1 || ix == kc) {
                if (lo < 0) lo = i;
                hi = i;
            }
        }
        if (lo > 0) {
            m.lo[kc - 1] = lo;
            m.hi[kc - 1] = hi;
        }
    }
    return m;
}
std::vector<std::pair<int, int>> mg_level_sizes(int xnum, int ynum, int max_levels);

// One device-resident hierarchy of precision T.
template <typename T>
struct MGHierarchy {
    std::vector<HostLevelGrid> hgrid;
    std::vector<MGLevelDev<T>> lev;
    std::vector<TransferDev<T>> xfer;   // xfer[k]: level k -> k+1 maps
    std::vector<void*> allocations;
    T* scratch = nullptr;               // pressure-shift scalar
    bool normalize_finest = true;       // pressure gauge p[1,1] = prnorm on level 0
    // LDS-tiled smoother plan per level (mg_plan_tiles)
    struct TilePlan {
        int ti, tj, nti, ntj;  // the regular tiling of the level
        size_t lds;            // LDS bytes of one regular tile
        bool whole_ok;         // the whole level fits one workgroup's LDS ...
        size_t whole_lds;      // ... with this many bytes
    };
    std::vector<TilePlan> tiles;
    bool fast_div = false;              // tiled smoother: FP32 hardware-reciprocal division
    bool additive_tiles = false;        // tiled smoother below the finest level: all tiles per launch
    // optional direct solve of the coarsest level (mg_coarse_setup): the
    // row-major inverse of its operator and two vectors of coarse_n entries;
    // the memory belongs to the caller
    const T* coarse_inv = nullptr;
    T *coarse_b = nullptr, *coarse_x = nullptr;
    int coarse_n = 0;
    int nlev() const { return (int)lev.size(); }
    ~MGHierarchy();
};

// clamp_weights: clamp normalised transfer distances to [0, 1] so boundary
// nodes of non-nested pressure grids interpolate instead of extrapolating
// (the reference's maps extrapolate; the preconditioner clamps).
template <typename T>
int mg_build(MGHierarchy<T>& H, int xnum, int ynum, double xsize, double ysize, int nlevels,
             const std::vector<HostLevelGrid>* finest_override, bool clamp_weights = false);
template <typename T>
void mg_restrict_viscosity(MGHierarchy<T>& H, hipStream_t s);
template <typename T>
void mg_smoother(MGHierarchy<T>& H, int k, int nsweeps, T theta_s, T theta_c, T prnorm,
                 const T* prnorm_dev, const T* bc_d, hipStream_t s);
template <typename T>
void mg_vcycle(MGHierarchy<T>& H, const std::vector<int>& smoothing, T theta_s, T theta_c,
               T relax_v, T relax_p, T prnorm, const T* prnorm_dev, const T* bc_d, hipStream_t s);
template <typename T>
void mg_zero_solution(MGHierarchy<T>& H, int k, hipStream_t s);

std::vector<int> mg_smoothing_counts(int nlev, int finest, int multiplier, int cap);

// Tiled smoother (ebg_mg_tile.hpp): plan the tiles of every level (a level
// whose whole window fits the workgroup LDS limit becomes one tile), then a
// V-cycle with the same structure as mg_vcycle whose smoothing calls run
// local_sweeps sweeps per tile launch.
template <typename T>
int mg_plan_tiles(MGHierarchy<T>& H, int tile, bool allow_single, bool fast_div);
template <typename T>
void mg_smoother_tiled(MGHierarchy<T>& H, int k, int nsweeps, int local_sweeps, T theta_s,
                       T theta_c, T prnorm, const T* prnorm_dev, const T* bc_d, hipStream_t s);
template <typename T>
void mg_vcycle_tiled(MGHierarchy<T>& H, const std::vector<int>& smoothing, int local_sweeps,
                     T theta_s, T theta_c, T relax_v, T relax_p, T prnorm, const T* prnorm_dev,
                     const T* bc_d, hipStream_t s);

// ---- direct solve of the coarsest level (ebg_coarse.hip) ----
// Unknowns of a level: interior vx (i 2..ynum, j 2..xnum-1), interior vy
// (i 2..ynum-1, j 2..xnum), every pressure cell.
inline int64_t mg_level_unknowns(int xnum, int ynum)
{
    return (int64_t)(ynum - 1) * (xnum - 2) + (int64_t)(ynum - 2) * (xnum - 1) + (int64_t)(ynum - 1) * (xnum - 1);
}
// Assembles the operator of level k in FP64 (the smoother's ghost rules with
// the mirror factors bc_d, the pressure gauge p(1,1) = 0 in place of the
// continuity row of cell (1,1)) into A (n x n, column-major, device), factors
// it with rocSOLVER, solves for the inverse in work (n x n FP64) and stores it
// row-major in precision T into inv.  Returns EBG_ERR_NUMERIC when the
// operator is singular.
template <typename T>
int mg_coarse_setup(const MGHierarchy<T>& H, int k, const double* bc_d, double* A, double* work, int* ipiv,
                    int* info, T* inv, void* solver_handle, hipStream_t s);
// x = inverse * b on the coarsest level: gathers RX, RY, RC, writes vx, vy, pr
// (zero elsewhere) and refreshes the ghosts.
template <typename T>
// zeroed: the coarsest solution is already cleared (fused restriction)
void mg_coarse_solve(MGHierarchy<T>& H, const T* bc_d, hipStream_t s, bool zeroed = false);

// ---- pieces of the reference's stand-alone multigrid driver (2-D and 3-D) ----
// Viscosity continuation: the initial range and step factor of
// MultigridDataManager.jl:349-373, the jump of ViscosityScaling.jl:267-309 and
// the exponent of scale_viscosity_for_all_levels! (ViscosityScaling.jl:41-79).
struct ViscRamp {
    double min_o = 0.0, max_o = 0.0, min_cur = 0.0, max_cur = 0.0, min_factor = 1.0, max_factor = 1.0;
    void init(double vmin, double vmax, double njumps);
    bool below_full() const { return max_cur < max_o || min_cur > min_o; }
    void jump();
    double coef() const;
};

constexpr int MG_RUN_BLOCKS = 256;
// Folds the minimum and maximum of device array a[0..n) into lo and hi
// (synchronises the stream); partial holds 2 * MG_RUN_BLOCKS doubles.
int mg_device_range(const double* a, int64_t n, double* partial, hipStream_t s, double& lo, double& hi);
// eta = min_cur * (orig / min_o)^coef on n entries
void mg_scale_viscosity(const double* orig, double* eta, int64_t n, double min_cur, double inv_min_o, double coef,
                        hipStream_t s);
// calculate_max_principle_residual_{2d,3d} (MultigridVCycle.jl:436-468): the
// largest 10^res of the cycle's n log10 residuals from the third cycle on,
// 1e38 before; NaN when any residual is NaN (Julia's max propagates NaN).
double mg_max_principal_residual(const double* res, int n, int ivcycle);

}  // namespace ebg
