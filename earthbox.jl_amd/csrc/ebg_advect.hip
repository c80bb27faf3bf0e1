// Marker-side kernels: inside flags, bit-exact cell indexing, the fused RK4
// gather-advect (velocity + spin + location + stress rotation) and grid spin.
//
// Compiled with -ffp-contract=off: every FP64 expression is evaluated in the
// reference's order without fused multiply-adds, and every division returns
// the correctly rounded quotient (rn_div: reciprocal plus Markstein
// correction, which equals IEEE division), so results are bit-identical to
// the reference loops (only the libm sin/cos of the stress rotation may differ
// in the last bit).
#include "ebg_common.hpp"
#include "ebg_device.hpp"

namespace ebg {

struct AxisDesc {
    const double* g;         // node coordinates, 0-based storage of a 1-based axis
    int n;                   // number of nodes
    double x0, inv_h;        // uniform-axis guess parameters
    int uniform;
    const AxisCell* cells;   // n-1 per-cell records (lo, hi, steps, reciprocals)
};

struct RK4Args {
    int64_t marknum;
    double dt, half_dt;
    double *x, *y;
    const int32_t *xn, *yn;
    const int8_t* flags;
    const double *vx1, *vy1, *esp;
    double *mvx, *mvy, *mspin;
    double *sxx, *sxy;
    int update_location;
    int xnum, ynum;
    AxisDesc ax, ay;
    double xmin, xmax, ymin, ymax;
};

// Correctly rounded a/b given rb = RN(1/b): q0 = RN(a*rb) is within 1.5 ulp
// of a/b, the first remainder step makes it faithful, and Markstein's theorem
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, "Newton-Raphson-based division with an FMA") makes
// the second step return RN(a/b): with rb within half an ulp of 1/b and q
// within one ulp of a/b, r = a - b*q is exact and RN(q + r*rb) = RN(a/b).
// The theorem needs no underflow or overflow in r and q: callers use it only
// for numerators that pass div_in_range() and divisors in [2^-100, 2^100]
// (ebg_grid2d_set rejects grid steps outside that range), and take the plain
// IEEE division otherwise, so the result always equals a / b bit for bit.
// The divisors are positive, so a / b has the sign of a; copying that sign
// onto q also turns the +0 the steps return for a = -0 into -0.
__device__ __forceinline__ double rn_div(double a, double b, double rb)
{
    double q = a * rb;
    double r = __builtin_fma(-b, q, a);
    q = __builtin_fma(r, rb, q);
    r = __builtin_fma(-b, q, a);
    q = __builtin_fma(r, rb, q);
    return __builtin_copysign(q, a);
}

// Numerators for which rn_div is exact with divisors and reciprocals between
// 2^-100 and 2^100 (grid steps): zero, or magnitude in [2^-900, 2^900].
__device__ __forceinline__ bool div_in_range(double a)
{
    const double m = fabs(a);
    return m == 0.0 || (m >= 0x1p-900 && m <= 0x1p900);
}

// GridToMarker.get_marker_value (interpolation/utils/GridToMarker.jl:190-226):
// bilinear weights, summed UL, LL, UR, LR starting from 0.0; p points at the
// upper-left node of a column-major array with leading dimension ld.
__device__ __forceinline__ double bilinear(const double* p, int ld, double dy, double dx)
{
    const double wul = (1.0 - dx) * (1.0 - dy);
    const double wll = (1.0 - dx) * dy;
    const double wur = dx * (1.0 - dy);
    const double wlr = dx * dy;
    double v = 0.0;
    v += wul * p[0];
    v += wll * p[1];
    v += wur * p[ld];
    v += wlr * p[ld + 1];
    return v;
}

__device__ __forceinline__ double marker_value(int yn, int xn, double dy, double dx,
                                               const double* __restrict__ a, int ld)
{
    return bilinear(a + (unsigned)((xn - 1) * ld + (yn - 1)), ld, dy, dx);
}

__device__ __forceinline__ int cell_index(double x, const AxisDesc& a)
{
    return left_node_index(x, a.g, a.n, a.x0, a.inv_h, a.uniform);
}

// clamp(searchsortedlast(g, x), 1, n-1) (MarkerGridMapping.jl:78-86) for the
// RK stages.  Uniform axes: every node lies within a quarter cell of
// x0 + (i-1)*h (checked when the grid is set), so for x in the true cell k,
// t = (x - x0)/h is in [k-1.25, k+0.25) and the guess floor(t)+1 is k-1, k or
// k+1; one exact comparison against the guessed cell's stored bounds
// (lo = g[c-1], hi = g[c]) then gives k.  Non-uniform axes: binary search.
// Written without branches: the guess clamps t to [0, n-2] (a NaN t becomes
// 0), which gives 1 for t < 1, n-1 for t >= n-2 and floor(t)+1 otherwise;
// a NaN x selects n-1 at the end.
__device__ __forceinline__ int uniform_guess(double x, const AxisDesc& a)
{
    const double t = (x - a.x0) * a.inv_h;
    return (int)fmin(fmax(t, 0.0), (double)(a.n - 2)) + 1;
}

__device__ __forceinline__ int uniform_fixup(double x, int c, double lo, double hi, int n)
{
    const int dn = (c > 1 && x < lo) ? 1 : 0;
    const int up = (c < n - 1 && x >= hi) ? 1 : 0;
    c = dn ? c - 1 : c + up;
    return (x != x) ? n - 1 : c;
}

template <bool UNIFORM>
__device__ __forceinline__ int stage_cell(double x, const AxisDesc& a)
{
    if (!UNIFORM) return left_node_index(x, a.g, a.n, 0.0, 0.0, 0);
    const int c = uniform_guess(x, a);
    const double2 lh = *reinterpret_cast<const double2*>(&a.cells[c - 1].lo);
    return uniform_fixup(x, c, lh.x, lh.y, a.n);
}

// Velocity and spin at one point (VelocityAndSpin.jl:343-428 with the
// MarkerGridMapping.jl:354-441 index remaps): vx on the vx grid (basic x
// cell, y remapped with strict '>'), vy on the vy grid (x remapped with
// strict '>'), spin on the basic grid.  The reference's clamps of the remapped
// indices (iyv <= ynum, jxv <= xnum) cannot bind: the basic cell is at most
// n-1, so the remap gives at most n.
struct StageWeights {
    double dxb, dyb, dyv, dxy;
    bool rx, ry;
};

__device__ __forceinline__ StageWeights stage_weights(const AxisCell& cx, const AxisCell& cy, double xc,
                                                      double yc)
{
    StageWeights w;
    w.ry = yc > cy.sb;
    w.rx = xc > cx.sb;
    const double nxb = xc - cx.lo, nyb = yc - cy.lo;
    const double nyv = yc - (w.ry ? cy.sb : cy.sa), nxy = xc - (w.rx ? cx.sb : cx.sa);
    const double hyv = w.ry ? cy.shb : cy.sha, hxy = w.rx ? cx.shb : cx.sha;
    if (div_in_range(nxb) && div_in_range(nyb) && div_in_range(nyv) && div_in_range(nxy)) {
        w.dxb = rn_div(nxb, cx.h, cx.rh);
        w.dyb = rn_div(nyb, cy.h, cy.rh);
        w.dyv = rn_div(nyv, hyv, w.ry ? cy.rshb : cy.rsha);
        w.dxy = rn_div(nxy, hxy, w.rx ? cx.rshb : cx.rsha);
    } else {
        w.dxb = nxb / cx.h;
        w.dyb = nyb / cy.h;
        w.dyv = nyv / hyv;
        w.dxy = nxy / hxy;
    }
    return w;
}

template <bool SPIN>
__device__ __forceinline__ void point_values(const RK4Args& a, double xc, double yc, int xi,
                                             int yi, double& vx, double& vy, double& sp)
{
    const AxisCell cx = a.ax.cells[xi - 1];
    const AxisCell cy = a.ay.cells[yi - 1];
    const StageWeights w = stage_weights(cx, cy, xc, yc);
    vx = marker_value(yi + (w.ry ? 1 : 0), xi, w.dyv, w.dxb, a.vx1, a.ynum + 1);
    vy = marker_value(yi, xi + (w.rx ? 1 : 0), w.dyb, w.dxy, a.vy1, a.ynum);
    if (SPIN) sp = marker_value(yi, xi, w.dyb, w.dxb, a.esp, a.ynum);
}

// (v1 + 2 v2 + 2 v3 + v4)/6 in the reference's order (RungeKutta.jl:36-47),
// i.e. ((v1 + 2 v2) + 2 v3) + v4: the kernel accumulates the partial sum
// stage by stage, which is the same sequence of roundings, and divides here.
__device__ __forceinline__ double rk4_sixth(double s)
{
    constexpr double r6 = 1.0 / 6.0;
    return div_in_range(s) ? rn_div(s, 6.0, r6) : s / 6.0;
}

// RK stages (RungeKutta.jl:9-47, VelocityAndSpin.jl:45-154), location update
// (MarkersLocation.jl:36-41) and stress rotation (MarkersStressRotation.jl:60-74)
// fused per marker.  F: RK_RK4 (runge_kutta_order_max 4, else one stage),
// RK_XN (stage-1 cells given), RK_STRESS (rotate sxx, sxy), RK_UNIFORM (both
// axes uniform), RK_SPIN (marker spin wanted; the stress rotation implies
// it).  Without spin outputs the spin field is not gathered.
enum { RK_RK4 = 1, RK_XN = 2, RK_STRESS = 4, RK_UNIFORM = 8, RK_SPIN = 16 };

template <int F>
__global__ void __launch_bounds__(256) rk4_advect_kernel(RK4Args a)
{
    constexpr int NST = (F & RK_RK4) ? 4 : 1;
    constexpr bool XN = (F & RK_XN) != 0, STRESS = (F & RK_STRESS) != 0;
    constexpr bool UNI = (F & RK_UNIFORM) != 0;
    constexpr bool SPIN = (F & (RK_SPIN | RK_STRESS)) != 0;
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
    double svx, svy, ssp;
    int xi, yi;
    if (XN) {
        xi = a.xn[m];
        yi = a.yn[m];
    } else {
        xi = stage_cell<UNI>(x0, a.ax);
        yi = stage_cell<UNI>(y0, a.ay);
    }
    double vx, vy, vs = 0.0;
    point_values<SPIN>(a, x0, y0, xi, yi, vx, vy, vs);
    if (NST == 4) {
        double ax = vx, ay = vy, as = vs;   // running ((v1 + 2 v2) + 2 v3) + v4
        double xc = x0 + a.half_dt * vx;
        double yc = y0 + a.half_dt * vy;
        point_values<SPIN>(a, xc, yc, stage_cell<UNI>(xc, a.ax), stage_cell<UNI>(yc, a.ay), vx, vy, vs);
        ax = ax + 2.0 * vx;
        ay = ay + 2.0 * vy;
        if (SPIN) as = as + 2.0 * vs;
        xc = x0 + a.half_dt * vx;
        yc = y0 + a.half_dt * vy;
        point_values<SPIN>(a, xc, yc, stage_cell<UNI>(xc, a.ax), stage_cell<UNI>(yc, a.ay), vx, vy, vs);
        ax = ax + 2.0 * vx;
        ay = ay + 2.0 * vy;
        if (SPIN) as = as + 2.0 * vs;
        xc = x0 + a.dt * vx;
        yc = y0 + a.dt * vy;
        point_values<SPIN>(a, xc, yc, stage_cell<UNI>(xc, a.ax), stage_cell<UNI>(yc, a.ay), vx, vy, vs);
        svx = rk4_sixth(ax + vx);
        svy = rk4_sixth(ay + vy);
        ssp = SPIN ? rk4_sixth(as + vs) : 0.0;
    } else {
        svx = vx;
        svy = vy;
        ssp = vs;
    }
    if (a.mvx) a.mvx[m] = svx;
    if (a.mvy) a.mvy[m] = svy;
    if (SPIN && a.mspin) a.mspin[m] = ssp;
    if (a.update_location) {
        a.x[m] = x0 + a.dt * svx;
        a.y[m] = y0 + a.dt * svy;
    }
    if (STRESS) {
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

// ---------------------------------------------------------------------------
// Tolerance-mode RK4 on nominally uniform axes (both axes uniform to 1e-9 of
// a step, checked when the grid is set).  Same stages, remaps, clamps and
// outputs as rk4_advect_kernel, with three changes of arithmetic:
//  - the cell search is one FMA: t2 = 2(x - x0)/h, and the half-cell index
//    k = clamp(floor(t2), 0, 2n-3) gives the basic cell (k >> 1) and the
//    velocity-grid remap (k & 1, the side of the cell centre) at once;
//  - the normalized distances come from the nominal node positions, so the
//    basic one is t - c and the staggered one differs from it by +-1/2;
//  - the bilinear value is the nested linear interpolation in FMAs, and the
//    RK4 sum is multiplied by 1/6.
// Spin is gathered only when an output needs it (SPIN).  The bilinear field
// is continuous, so a different cell choice at a node or centre (where the
// nominal and the stored coordinates round differently) changes the value by
// rounding only.  The grids are read through buffer descriptors (32-bit
// offsets, no 64-bit address arithmetic).  On the MI355X the kernel is bound
// by the vector memory address unit: each 16-byte corner-pair gather takes it
// a cycle per four lanes (16 gathers per marker).  Two markers per thread
// measured no faster; a per-wavefront LDS window of the grids gained 7 % on
// markers in their initial order and lost 27 % once they had drifted.
enum { RKF_RK4 = 1, RKF_XN = 2, RKF_STRESS = 4, RKF_SPIN = 8 };

struct RK4FastArgs {
    int64_t marknum;
    double dt, half_dt;
    double *x, *y;
    const int32_t *xn, *yn;
    const int8_t* flags;
    const double *vx1, *vy1, *esp;
    double *mvx, *mvy, *mspin;
    double *sxx, *sxy;
    int update_location;
    int xnum, ynum;
    double inv_hx2, nx02, inv_hy2, ny02;   // t2 = x * inv_h2 + n02 = 2 (x - x0) / h
    double kmax_x, kmax_y;                 // 2n - 3: last half cell
    double xmin, xmax, ymin, ymax;
};

struct FastCell {
    int c;        // 0-based basic cell
    int r;        // 1: right of (x) / below (y) the cell centre
    double db;    // distance from the basic node, in cells
    double ds;    // distance from the staggered node the remap picks
};

// +0.5 (r = 0) or -0.5 (r = 1) from the sign bit
__device__ __forceinline__ double signed_half(int r)
{
    return __builtin_bit_cast(double, (unsigned long long)(0x3fe00000u | ((unsigned)r << 31)) << 32);
}

// t2 = 2t is exact (a doubled FMA), so db = t2/2 - c rounds as t - c
__device__ __forceinline__ FastCell fast_cell(double x, double inv_h2, double n02, double kmax)
{
    FastCell f;
    const double t2 = __builtin_fma(x, inv_h2, n02);
    const int k = (int)fmin(fmax(t2, 0.0), kmax);   // NaN -> 0
    f.c = k >> 1;
    f.r = k & 1;
    f.db = __builtin_fma(t2, 0.5, -(double)f.c);
    f.ds = f.db + signed_half(f.r);
    return f;
}

// Cell from a given 1-based basic index (the stage-1 marker_xn/yn, clamped
// to the grid): the remap side follows from the nominal distance.
__device__ __forceinline__ FastCell fast_cell_given(double x, int c1, double inv_h2, double n02, int n)
{
    FastCell f;
    const double t2 = __builtin_fma(x, inv_h2, n02);
    f.c = min(max(c1 - 1, 0), n - 2);
    f.db = __builtin_fma(t2, 0.5, -(double)f.c);
    f.r = f.db > 0.5;
    f.ds = f.db + signed_half(f.r);
    return f;
}

// Grid reads through buffer descriptors: 32-bit byte offsets (the launcher
// keeps every grid below 2 GiB) and the column step as the scalar offset of
// the second load, so no 64-bit address arithmetic per read.
typedef unsigned int ebg_u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t grid_rsrc(const double* p, int64_t count)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, 0, (int)(count * 8), 0x00020000);
}

__device__ __forceinline__ double lo_f64(ebg_u32x4 v)
{
    return __builtin_bit_cast(double, ((unsigned long long)v.y << 32) | v.x);
}
__device__ __forceinline__ double hi_f64(ebg_u32x4 v)
{
    return __builtin_bit_cast(double, ((unsigned long long)v.w << 32) | v.z);
}

// nested linear interpolation from the four corner values
__device__ __forceinline__ double lerp4(double p00, double p10, double p01, double p11, double dy, double dx)
{
    const double a = __builtin_fma(dy, p10 - p00, p00);
    const double b = __builtin_fma(dy, p11 - p01, p01);
    return __builtin_fma(dx, b - a, a);
}

// the same at node idx (upper left) of a column-major grid with column step
// ld8 bytes, read through its buffer descriptor
__device__ __forceinline__ double lerp2(__amdgpu_buffer_rsrc_t r, int idx, int ld8, double dy, double dx)
{
    const ebg_u32x4 c0 = __builtin_amdgcn_raw_buffer_load_b128(r, idx << 3, 0, 0);
    const ebg_u32x4 c1 = __builtin_amdgcn_raw_buffer_load_b128(r, idx << 3, ld8, 0);
    return lerp4(lo_f64(c0), hi_f64(c0), lo_f64(c1), hi_f64(c1), dy, dx);
}

struct FastGrids {
    __amdgpu_buffer_rsrc_t vx, vy, sp;
    int ldx, ldy, ldx8, ldy8;
};

template <bool SPIN>
__device__ __forceinline__ void fast_values(const FastGrids& g, const FastCell& cx, const FastCell& cy,
                                            double& vx, double& vy, double& sp)
{
    // vx grid: basic x cell, y remapped; vy grid: x remapped, basic y cell
    vx = lerp2(g.vx, __umul24(cx.c, g.ldx) + cy.c + cy.r, g.ldx8, cy.ds, cx.db);
    vy = lerp2(g.vy, __umul24(cx.c + cx.r, g.ldy) + cy.c, g.ldy8, cy.db, cx.ds);
    if (SPIN) sp = lerp2(g.sp, __umul24(cx.c, g.ldy) + cy.c, g.ldy8, cy.db, cx.db);
}

template <int F>
__global__ void __launch_bounds__(256) rk4_fast_kernel(RK4FastArgs a)
{
    constexpr bool RK4 = (F & RKF_RK4) != 0, XN = (F & RKF_XN) != 0;
    constexpr bool STRESS = (F & RKF_STRESS) != 0;
    constexpr bool SPIN = (F & (RKF_SPIN | RKF_STRESS)) != 0;   // the rotation needs it
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
    FastGrids g;
    g.ldx = a.ynum + 1;
    g.ldy = a.ynum;
    g.ldx8 = g.ldx * 8;
    g.ldy8 = g.ldy * 8;
    g.vx = grid_rsrc(a.vx1, (int64_t)g.ldx * a.xnum);
    g.vy = grid_rsrc(a.vy1, (int64_t)g.ldy * (a.xnum + 1));
    if (SPIN) g.sp = grid_rsrc(a.esp, (int64_t)g.ldy * a.xnum);
    FastCell cx, cy;
    if (XN) {
        cx = fast_cell_given(x0, a.xn[m], a.inv_hx2, a.nx02, a.xnum);
        cy = fast_cell_given(y0, a.yn[m], a.inv_hy2, a.ny02, a.ynum);
    } else {
        cx = fast_cell(x0, a.inv_hx2, a.nx02, a.kmax_x);
        cy = fast_cell(y0, a.inv_hy2, a.ny02, a.kmax_y);
    }
    double vx, vy, vs = 0.0;
    fast_values<SPIN>(g, cx, cy, vx, vy, vs);
    double svx = vx, svy = vy, ssp = vs;
    if (RK4) {
        double ax = vx, ay = vy, as = vs;
        double xc = __builtin_fma(a.half_dt, vx, x0);
        double yc = __builtin_fma(a.half_dt, vy, y0);
        fast_values<SPIN>(g, fast_cell(xc, a.inv_hx2, a.nx02, a.kmax_x),
                          fast_cell(yc, a.inv_hy2, a.ny02, a.kmax_y), vx, vy, vs);
        ax = __builtin_fma(2.0, vx, ax);
        ay = __builtin_fma(2.0, vy, ay);
        if (SPIN) as = __builtin_fma(2.0, vs, as);
        xc = __builtin_fma(a.half_dt, vx, x0);
        yc = __builtin_fma(a.half_dt, vy, y0);
        fast_values<SPIN>(g, fast_cell(xc, a.inv_hx2, a.nx02, a.kmax_x),
                          fast_cell(yc, a.inv_hy2, a.ny02, a.kmax_y), vx, vy, vs);
        ax = __builtin_fma(2.0, vx, ax);
        ay = __builtin_fma(2.0, vy, ay);
        if (SPIN) as = __builtin_fma(2.0, vs, as);
        xc = __builtin_fma(a.dt, vx, x0);
        yc = __builtin_fma(a.dt, vy, y0);
        fast_values<SPIN>(g, fast_cell(xc, a.inv_hx2, a.nx02, a.kmax_x),
                          fast_cell(yc, a.inv_hy2, a.ny02, a.kmax_y), vx, vy, vs);
        constexpr double r6 = 1.0 / 6.0;
        svx = (ax + vx) * r6;
        svy = (ay + vy) * r6;
        if (SPIN) ssp = (as + vs) * r6;
    }
    if (a.mvx) a.mvx[m] = svx;
    if (a.mvy) a.mvy[m] = svy;
    if (SPIN && a.mspin) a.mspin[m] = ssp;
    if (a.update_location) {
        a.x[m] = __builtin_fma(a.dt, svx, x0);
        a.y[m] = __builtin_fma(a.dt, svy, y0);
    }
    if (STRESS) {
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

static const void* const rk4_fast_kernels[16] = {
    (const void*)rk4_fast_kernel<0>,  (const void*)rk4_fast_kernel<1>,
    (const void*)rk4_fast_kernel<2>,  (const void*)rk4_fast_kernel<3>,
    (const void*)rk4_fast_kernel<4>,  (const void*)rk4_fast_kernel<5>,
    (const void*)rk4_fast_kernel<6>,  (const void*)rk4_fast_kernel<7>,
    (const void*)rk4_fast_kernel<8>,  (const void*)rk4_fast_kernel<9>,
    (const void*)rk4_fast_kernel<10>, (const void*)rk4_fast_kernel<11>,
    (const void*)rk4_fast_kernel<12>, (const void*)rk4_fast_kernel<13>,
    (const void*)rk4_fast_kernel<14>, (const void*)rk4_fast_kernel<15>};

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
    return AxisDesc{g.gridx_b, (int)g.xnum, g.x0, g.inv_hx, g.uniform_x, g.cells_x};
}
static AxisDesc axis_y(const Grid2D& g)
{
    return AxisDesc{g.gridy_b, (int)g.ynum, g.y0, g.inv_hy, g.uniform_y, g.cells_y};
}

// nominally uniform axes, and grids the fast kernel can address with 32-bit
// byte offsets
static bool fast_kernel_fits(const Grid2D& g)
{
    return g.nominal_x && g.nominal_y && (int64_t)(g.ynum + 1) * (g.xnum + 1) * 8 < ((int64_t)1 << 31);
}

static unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

#define EBG_RK4_K(f) (const void*)rk4_advect_kernel<f>
#define EBG_RK4_K8(b) EBG_RK4_K(b), EBG_RK4_K(b + 1), EBG_RK4_K(b + 2), EBG_RK4_K(b + 3), \
    EBG_RK4_K(b + 4), EBG_RK4_K(b + 5), EBG_RK4_K(b + 6), EBG_RK4_K(b + 7)
// rk4_advect_kernel<F> for every flag combination F in [0, 32)
static const void* const rk4_kernels[32] = {EBG_RK4_K8(0), EBG_RK4_K8(8), EBG_RK4_K8(16),
                                            EBG_RK4_K8(24)};
#undef EBG_RK4_K8
#undef EBG_RK4_K

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
    if (ctx->rk4_arithmetic == EBG_RK4_TOLERANCE && fast_kernel_fits(g)) {
        RK4FastArgs a;
        a.marknum = marknum;
        a.dt = timestep;
        a.half_dt = timestep / 2.0;
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
        // doubled exactly: t2 = 2t
        a.inv_hx2 = 2.0 * g.inv_hx;
        a.nx02 = 2.0 * (-g.x0 * g.inv_hx);
        a.inv_hy2 = 2.0 * g.inv_hy;
        a.ny02 = 2.0 * (-g.y0 * g.inv_hy);
        a.kmax_x = (double)(2 * g.xnum - 3);
        a.kmax_y = (double)(2 * g.ynum - 3);
        a.xmin = g.xmin;
        a.xmax = g.xmax;
        a.ymin = g.ymin;
        a.ymax = g.ymax;
        const int f = (order == 4 ? RKF_RK4 : 0) | (marker_xn ? RKF_XN : 0) |
                      (marker_sxx ? RKF_STRESS : 0) | (marker_spin ? RKF_SPIN : 0);
        void* args[] = {&a};
        EBG_HIP(hipLaunchKernel(rk4_fast_kernels[f], dim3(blocks_for(marknum, 256)), dim3(256), args,
                                0, ctx->stream));
        EBG_HIP(hipGetLastError());
        return EBG_OK;
    }
    RK4Args a;
    a.marknum = marknum;
    a.dt = timestep;
    a.half_dt = timestep / 2.0;
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
    a.xmin = g.xmin;
    a.xmax = g.xmax;
    a.ymin = g.ymin;
    a.ymax = g.ymax;
    const int f = (order == 4 ? RK_RK4 : 0) | (marker_xn ? RK_XN : 0) |
                  (marker_sxx ? RK_STRESS : 0) | (g.uniform_x && g.uniform_y ? RK_UNIFORM : 0) |
                  (marker_spin ? RK_SPIN : 0);
    void* args[] = {&a};
    EBG_HIP(hipLaunchKernel(rk4_kernels[f], dim3(blocks_for(marknum, 256)), dim3(256), args, 0,
                            ctx->stream));
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_rk4_set_arithmetic(ebg_ctx* ctx, int mode)
{
    EBG_CHECK(ctx != nullptr, EBG_ERR_ARG, "ebg_rk4_set_arithmetic: ctx is NULL");
    EBG_CHECK(mode == EBG_RK4_EXACT || mode == EBG_RK4_TOLERANCE, EBG_ERR_ARG,
              "ebg_rk4_set_arithmetic: mode must be EBG_RK4_EXACT or EBG_RK4_TOLERANCE");
    ctx->rk4_arithmetic = mode;
    return EBG_OK;
}

int ebg_rk4_tolerance_kernel_available(ebg_ctx* ctx)
{
    if (!ctx || !ctx->has_grid) return -1;
    return fast_kernel_fits(ctx->grid) ? 1 : 0;
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
