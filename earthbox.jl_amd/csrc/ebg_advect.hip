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
// outputs as rk4_advect_kernel, with these changes of arithmetic:
//  - positions are carried in cell coordinates t = (x - x0)/h (one FMA from
//    the nominal origin and reciprocal step), and each stage point is one FMA
//    from the marker's t: t + (h_stage/h) v;
//  - the basic cell is floor(t) (clamped to the grid), its normalized
//    distance t - floor(t); the velocity-grid remap takes the staggered node
//    of the cell's far half when that distance is at least 1/2 (computed
//    cells) or above 1/2 (the stage-1 cells a caller passes), and its
//    distance differs from the basic one by -+1/2;
//  - the bilinear value is the nested linear interpolation in FMAs from each
//    node's value and its difference to the next node down the column, and
//    the RK4 sum is multiplied by 1/6.
// The bilinear field is continuous, so a different cell choice at a node or
// centre (where nominal and stored coordinates round differently) changes the
// value by rounding only; positions stay within 1e-10 of the domain size of
// the exact kernel's.
//
// Grid reads.  Markers kept in their creation order put a wavefront's 64
// markers in a few neighbouring cells of one column, so the 16 corner pairs a
// marker reads (24 with spin) come from a small patch of nodes.  Each
// wavefront takes the cells of its first and last marker (or, when some of
// its markers lie outside that box, the least and greatest cells of all of
// them), widens that box by one cell on every side, and copies the patch of
// every grid it reads into its
// own LDS tile as (value, difference to the next node down the column) pairs,
// with one 16-byte load per node.  A stage point whose cell lies in the box
// reads its two pairs per grid from LDS; any other point (a marker far from
// its neighbours in memory, a stage leaving the box) computes the clamped cell
// and reads global memory, with the same arithmetic, so every marker's result
// is the same whether its reads hit the tile or not, and the same whatever
// the marker order.  A box larger than the tile sends the whole wavefront to
// global reads.  Global reads go through buffer descriptors (32-bit offsets;
// reads past a grid's end return zero and are never used).
enum { RKF_RK4 = 1, RKF_XN = 2, RKF_STRESS = 4, RKF_SPIN = 8 };

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


struct RK4TolArgs {
    int64_t marknum;
    double dt;
    double *x, *y;
    const int32_t *xn, *yn;
    const int8_t* flags;
    const double *vx1, *vy1, *esp;
    double *mvx, *mvy, *mspin;
    double *sxx, *sxy;
    int update_location;
    int xnum, ynum;
    double ixh, tx0, iyh, ty0;        // cell coordinates: t = x * ixh + tx0
    double hdx, hdy, ddx, ddy;        // stage steps in cells per unit velocity: dt/2/h, dt/h
    double xmin, xmax, ymin, ymax;
    int tiles;                        // 0: every read from global memory
};

constexpr int RKT_WAVES = 4;   // wavefronts per 256-thread block
constexpr int RKT_TILE = 160;  // node pairs per grid per wavefront (2.5 KB)

// Cell of a stage point: c, remap side r, basic and staggered distances.
struct TolCell {
    int c, r;
    double db, ds;
};

__device__ __forceinline__ double half_toward(bool r)   // -1/2 when r, else +1/2
{
    return r ? -0.5 : 0.5;
}

// clamp(floor(t), 0, n-2), clamped in FP64 before the conversion (a NaN t
// gives cell 0)
__device__ __forceinline__ TolCell tol_cell_clamped(double t, int nm2)
{
    TolCell p;
    p.c = (int)fmin(fmax(__builtin_floor(t), 0.0), (double)nm2);
    p.db = t - (double)p.c;
    p.r = p.db >= 0.5;
    p.ds = p.db + half_toward(p.r);
    return p;
}

// the caller's 1-based stage-1 cell, clamped to the grid
__device__ __forceinline__ TolCell tol_cell_given(double t, int c1, int nm2)
{
    TolCell p;
    p.c = min(max(c1 - 1, 0), nm2);
    p.db = t - (double)p.c;
    p.r = p.db > 0.5;
    p.ds = p.db + half_toward(p.r);
    return p;
}

// nested linear interpolation from two (value, difference down the column)
// pairs: q0 at the upper-left node, q1 one column to the right
__device__ __forceinline__ double lerp_pairs(double2 q0, double2 q1, double dy, double dx)
{
    const double a = __builtin_fma(dy, q0.y, q0.x);
    const double b = __builtin_fma(dy, q1.y, q1.x);
    return __builtin_fma(dx, b - a, a);
}

// the same from global memory: node idx (upper left) of a column-major grid
// with column step ld8 bytes, read through its buffer descriptor
__device__ __forceinline__ double lerp_global(__amdgpu_buffer_rsrc_t r, int idx, int ld8, double dy, double dx)
{
    const ebg_u32x4 c0 = __builtin_amdgcn_raw_buffer_load_b128(r, idx << 3, 0, 0);
    const ebg_u32x4 c1 = __builtin_amdgcn_raw_buffer_load_b128(r, idx << 3, ld8, 0);
    const double p00 = lo_f64(c0), p01 = lo_f64(c1);
    return lerp_pairs(make_double2(p00, hi_f64(c0) - p00), make_double2(p01, hi_f64(c1) - p01), dy, dx);
}

struct TolGrids {
    __amdgpu_buffer_rsrc_t vx, vy, sp;
    int ldx, ldy, ldx8, ldy8;
};

// A wavefront's LDS tile: cells X0 .. X0+ncv-1 by Y0 .. Y0+nrv-1 (ncv = 0:
// no tile), nodes X0 .. X0+ncv+1 by Y0 .. Y0+P-1 in columns of P pairs.
struct TolTile {
    const double2 *vx, *vy, *sp;
    int X0, Y0, P;
    unsigned ncv, nrv;
    double X0d, Y0d, ncvd, nrvd;   // the same in FP64
};

// global reads at a clamped (or given) cell: the fallback of the tile
template <bool SPIN>
__device__ __forceinline__ void tol_values_global(const TolGrids& g, const TolCell& px, const TolCell& py,
                                                  double& vx, double& vy, double& sp)
{
    vx = lerp_global(g.vx, __umul24(px.c, g.ldx) + py.c + py.r, g.ldx8, py.ds, px.db);
    vy = lerp_global(g.vy, __umul24(px.c + px.r, g.ldy) + py.c, g.ldy8, py.db, px.ds);
    if (SPIN) sp = lerp_global(g.sp, __umul24(px.c, g.ldy) + py.c, g.ldy8, py.db, px.db);
}

// tile reads at tile cell (lc, lr)
template <bool SPIN>
__device__ __forceinline__ void tol_values_tile(const TolTile& w, int lc, int lr, const TolCell& px,
                                                const TolCell& py, double& vx, double& vy, double& sp)
{
    int i = lc * w.P + lr + py.r;   // vx: basic x cell, y remapped
    vx = lerp_pairs(w.vx[i], w.vx[i + w.P], py.ds, px.db);
    i = (lc + px.r) * w.P + lr;      // vy: x remapped, basic y cell
    vy = lerp_pairs(w.vy[i], w.vy[i + w.P], py.db, px.ds);
    if (SPIN) {
        i = lc * w.P + lr;
        sp = lerp_pairs(w.sp[i], w.sp[i + w.P], py.db, px.db);
    }
}

// velocity (and spin) at the stage point with cell coordinates (tx, ty).
// The tile test is done on floor(t) in FP64 (false for NaN), so the integer
// conversion only sees cells of the box; any other point takes the clamped
// cell, which equals floor(t) for every cell of the grid.
template <bool SPIN>
__device__ __forceinline__ void tol_values(const TolTile& w, const TolGrids& g, double tx, double ty, int nxm2,
                                           int nym2, double& vx, double& vy, double& sp)
{
    const double fx = __builtin_floor(tx), fy = __builtin_floor(ty);
    const double ox = fx - w.X0d, oy = fy - w.Y0d;
    const bool in = (ox >= 0.0) & (ox < w.ncvd) & (oy >= 0.0) & (oy < w.nrvd);
    if (in) {
        TolCell px, py;
        px.db = tx - fx;
        py.db = ty - fy;
        px.r = px.db >= 0.5;
        py.r = py.db >= 0.5;
        px.ds = px.db + half_toward(px.r);
        py.ds = py.db + half_toward(py.r);
        tol_values_tile<SPIN>(w, (int)ox, (int)oy, px, py, vx, vy, sp);
    } else {
        tol_values_global<SPIN>(g, tol_cell_clamped(tx, nxm2), tol_cell_clamped(ty, nym2), vx, vy, sp);
    }
}

__device__ __forceinline__ bool tol_in_tile(const TolTile& w, int cx, int cy)
{
    return ((unsigned)(cx - w.X0) < w.ncv) & ((unsigned)(cy - w.Y0) < w.nrv);
}

// copy nodes (X0 + col, Y0 + row), col < nc, row < P, of one grid into a
// tile as (value, next value down the column - value)
__device__ __forceinline__ void tol_fill(double2* t, __amdgpu_buffer_rsrc_t r, int ld, int X0, int Y0, int P,
                                         int ne, int lane)
{
    const float rP = __builtin_amdgcn_rcpf((float)P);
    for (int e = lane; e < ne; e += 64) {
        const int col = (int)(((float)e + 0.5f) * rP);   // floor(e / P): e < 2^7, |error| < 1/(2P)
        const int row = e - col * P;
        const ebg_u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(r, ((X0 + col) * ld + Y0 + row) << 3, 0, 0);
        t[e] = make_double2(lo_f64(q), hi_f64(q) - lo_f64(q));
    }
}

// wavefront-level ordering of the tile writes before the reads of other lanes
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int F>
__global__ void __launch_bounds__(256) rk4_tol_kernel(RK4TolArgs a)
{
    constexpr bool RK4 = (F & RKF_RK4) != 0, XN = (F & RKF_XN) != 0;
    constexpr bool STRESS = (F & RKF_STRESS) != 0;
    constexpr bool SPIN = (F & (RKF_SPIN | RKF_STRESS)) != 0;   // the rotation needs it
    constexpr int NG = SPIN ? 3 : 2;
    __shared__ double2 tiles[RKT_WAVES][NG][RKT_TILE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = m < a.marknum;
    double x0 = 0.0, y0 = 0.0;
    if (live) {
        x0 = a.x[m];
        y0 = a.y[m];
    }
    bool inside = false;
    if (live) {
        if (a.flags) inside = (a.flags[m] == 1);
        else inside = (a.xmin <= x0) & (x0 <= a.xmax) & (a.ymin <= y0) & (y0 <= a.ymax);
        if (!inside) {  // VelocityAndSpin.jl:120-126; location and stress untouched
            if (a.mvx) a.mvx[m] = 0.0;
            if (a.mvy) a.mvy[m] = 0.0;
            if (a.mspin) a.mspin[m] = 0.0;
        }
    }
    const int nxm2 = a.xnum - 2, nym2 = a.ynum - 2;
    TolGrids g;
    g.ldx = a.ynum + 1;
    g.ldy = a.ynum;
    g.ldx8 = g.ldx * 8;
    g.ldy8 = g.ldy * 8;
    g.vx = grid_rsrc(a.vx1, (int64_t)g.ldx * a.xnum);
    g.vy = grid_rsrc(a.vy1, (int64_t)g.ldy * (a.xnum + 1));
    if (SPIN) g.sp = grid_rsrc(a.esp, (int64_t)g.ldy * a.xnum);
    const double tx = __builtin_fma(x0, a.ixh, a.tx0);
    const double ty = __builtin_fma(y0, a.iyh, a.ty0);
    // the tile box: the cells of the wavefront's first and last marker; when
    // a marker inside the domain lies outside that box, the least and
    // greatest cells of all its markers inside the domain.  Widened by one
    // cell and clipped to the grid.
    const int ix = (int)fmin(fmax(tx, -2.0), 1.0e9), iy = (int)fmin(fmax(ty, -2.0), 1.0e9);
    const int fx = __builtin_amdgcn_readlane(ix, 0), lx = __builtin_amdgcn_readlane(ix, 63);
    const int fy = __builtin_amdgcn_readlane(iy, 0), ly = __builtin_amdgcn_readlane(iy, 63);
    int bx0 = min(fx, lx), bx1 = max(fx, lx), by0 = min(fy, ly), by1 = max(fy, ly);
    const bool outside_box = inside & ((ix < bx0) | (ix > bx1) | (iy < by0) | (iy > by1));
    if (__ballot(outside_box) != 0) {
        bx0 = __reduce_min_sync(~0ull, inside ? ix : 0x7fffffff);
        bx1 = __reduce_max_sync(~0ull, inside ? ix : -1);
        by0 = __reduce_min_sync(~0ull, inside ? iy : 0x7fffffff);
        by1 = __reduce_max_sync(~0ull, inside ? iy : -1);
    }
    TolTile w;
    w.X0 = max(bx0 - 1, 0);
    w.Y0 = max(by0 - 1, 0);
    const int ncv = min(bx1 + 1, nxm2) - w.X0 + 1;
    const int nrv = min(by1 + 1, nym2) - w.Y0 + 1;
    w.P = nrv + 2;
    const bool on = a.tiles && ncv >= 1 && nrv >= 1 && ncv <= RKT_TILE && nrv <= RKT_TILE &&
                    (ncv + 2) * w.P <= RKT_TILE;
    w.ncv = on ? (unsigned)ncv : 0u;
    w.nrv = on ? (unsigned)nrv : 0u;
    w.X0d = (double)w.X0;
    w.Y0d = (double)w.Y0;
    w.ncvd = (double)w.ncv;
    w.nrvd = (double)w.nrv;
    w.vx = tiles[wave][0];
    w.vy = tiles[wave][1];
    w.sp = tiles[wave][NG - 1];
    if (on) {
        const int ne = (ncv + 2) * w.P;
        tol_fill(tiles[wave][0], g.vx, g.ldx, w.X0, w.Y0, w.P, ne, lane);
        tol_fill(tiles[wave][1], g.vy, g.ldy, w.X0, w.Y0, w.P, ne, lane);
        if (SPIN) tol_fill(tiles[wave][NG - 1], g.sp, g.ldy, w.X0, w.Y0, w.P, ne, lane);
    }
    wave_lds_sync();
    if (!inside) return;
    double vx, vy, vs = 0.0;
    if (XN) {
        const TolCell px = tol_cell_given(tx, a.xn[m], nxm2);
        const TolCell py = tol_cell_given(ty, a.yn[m], nym2);
        if (tol_in_tile(w, px.c, py.c)) tol_values_tile<SPIN>(w, px.c - w.X0, py.c - w.Y0, px, py, vx, vy, vs);
        else tol_values_global<SPIN>(g, px, py, vx, vy, vs);
    } else {
        tol_values<SPIN>(w, g, tx, ty, nxm2, nym2, vx, vy, vs);
    }
    double svx = vx, svy = vy, ssp = vs;
    if (RK4) {
        double ax = vx, ay = vy, as = vs;
        tol_values<SPIN>(w, g, __builtin_fma(a.hdx, vx, tx), __builtin_fma(a.hdy, vy, ty), nxm2, nym2, vx, vy,
                         vs);
        ax = __builtin_fma(2.0, vx, ax);
        ay = __builtin_fma(2.0, vy, ay);
        if (SPIN) as = __builtin_fma(2.0, vs, as);
        tol_values<SPIN>(w, g, __builtin_fma(a.hdx, vx, tx), __builtin_fma(a.hdy, vy, ty), nxm2, nym2, vx, vy,
                         vs);
        ax = __builtin_fma(2.0, vx, ax);
        ay = __builtin_fma(2.0, vy, ay);
        if (SPIN) as = __builtin_fma(2.0, vs, as);
        tol_values<SPIN>(w, g, __builtin_fma(a.ddx, vx, tx), __builtin_fma(a.ddy, vy, ty), nxm2, nym2, vx, vy,
                         vs);
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

static const void* const rk4_tol_kernels[16] = {
    (const void*)rk4_tol_kernel<0>,  (const void*)rk4_tol_kernel<1>,
    (const void*)rk4_tol_kernel<2>,  (const void*)rk4_tol_kernel<3>,
    (const void*)rk4_tol_kernel<4>,  (const void*)rk4_tol_kernel<5>,
    (const void*)rk4_tol_kernel<6>,  (const void*)rk4_tol_kernel<7>,
    (const void*)rk4_tol_kernel<8>,  (const void*)rk4_tol_kernel<9>,
    (const void*)rk4_tol_kernel<10>, (const void*)rk4_tol_kernel<11>,
    (const void*)rk4_tol_kernel<12>, (const void*)rk4_tol_kernel<13>,
    (const void*)rk4_tol_kernel<14>, (const void*)rk4_tol_kernel<15>};

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
        RK4TolArgs a;
        a.marknum = marknum;
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
        a.ixh = g.inv_hx;
        a.tx0 = -g.x0 * g.inv_hx;
        a.iyh = g.inv_hy;
        a.ty0 = -g.y0 * g.inv_hy;
        a.hdx = (timestep / 2.0) * g.inv_hx;
        a.hdy = (timestep / 2.0) * g.inv_hy;
        a.ddx = timestep * g.inv_hx;
        a.ddy = timestep * g.inv_hy;
        a.xmin = g.xmin;
        a.xmax = g.xmax;
        a.ymin = g.ymin;
        a.ymax = g.ymax;
        a.tiles = ctx->rk4_tiles;
        const int f = (order == 4 ? RKF_RK4 : 0) | (marker_xn ? RKF_XN : 0) |
                      (marker_sxx ? RKF_STRESS : 0) | (marker_spin ? RKF_SPIN : 0);
        void* args[] = {&a};
        EBG_HIP(hipLaunchKernel(rk4_tol_kernels[f], dim3(blocks_for(marknum, 256)), dim3(256), args, 0,
                                ctx->stream));
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

int ebg_rk4_set_tiling(ebg_ctx* ctx, int on)
{
    EBG_CHECK(ctx != nullptr, EBG_ERR_ARG, "ebg_rk4_set_tiling: ctx is NULL");
    EBG_CHECK(on == 0 || on == 1, EBG_ERR_ARG, "ebg_rk4_set_tiling: on must be 0 or 1");
    ctx->rk4_tiles = on;
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
