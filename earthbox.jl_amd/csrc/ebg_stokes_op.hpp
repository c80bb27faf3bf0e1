// Matrix-free rows of the assembled Stokes-continuity system.
//
// For cell (i, j) (1-based, i = row/y) the three rows x-Stokes, y-Stokes and
// continuity are evaluated with exactly the coefficients the reference's COO
// build produces (solver/build_system/StokesBuildManager.jl:661-706 and the
// stencil files under discretize_{x,y}_stokes_equation/ and
// discretize_continuity_equation/), and the products are summed in the COO
// triplet order of that row, so y = L*x matches the reference's SpMV bit for
// bit.  Ghost-node boundary conditions are folded through the b* arrays
// (v_ghost = C + D*v); Kbond rows carry 2*pscale/(xstpavg+ystpavg).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ebg {

struct StokesOpDev {
    int xnum, ynum;
    const double *xstp, *ystp, *xstpc, *ystpc;  // xstp_b, ystp_b, xstp_vy, ystp_vx
    double xstpavg, ystpavg;
    const double *etan, *etas, *rho1_vy;          // (ny-1,nx-1), (ny,nx), (ny,nx+1)
    const double *RX, *RY, *RC;                   // (ny+1,nx), (ny,nx+1), (ny-1,nx-1)
    const double *btopx, *btopy, *bbottomx, *bbottomy;  // (nx+1) x 2
    const double *bleftx, *blefty, *brightx, *brighty;  // (ny+1) x 2
    int zx_j, zx_i0, zx_i1, zy_j, zy_i0, zy_i1;   // bintern_zone[1..3], [5..7]
    double zvx, zvy;                              // bintern_velocity[4], [8]
    int pressure_bc_mode;
    double pressure_bc, pscale, timestep, gravity_y;
    int stab;
};

// Unknown vector layouts: the reference's interleaved order (3c-2, 3c-1, 3c,
// cell c = (j-1)(ynum-1)+i) and the planar order used inside the solver.
struct LayoutInterleaved {
    int nc;
    __device__ __forceinline__ int64_t vx(int c) const { return 3 * (int64_t)c; }
    __device__ __forceinline__ int64_t vy(int c) const { return 3 * (int64_t)c + 1; }
    __device__ __forceinline__ int64_t p(int c) const { return 3 * (int64_t)c + 2; }
};
struct LayoutPlanar {
    int nc;
    __device__ __forceinline__ int64_t vx(int c) const { return c; }
    __device__ __forceinline__ int64_t vy(int c) const { return (int64_t)nc + c; }
    __device__ __forceinline__ int64_t p(int c) const { return 2 * (int64_t)nc + c; }
};

#define EBG_N(a, i, j) (s.etan[(size_t)((j) - 1) * (s.ynum - 1) + ((i) - 1)])
#define EBG_S(a, i, j) (s.etas[(size_t)((j) - 1) * s.ynum + ((i) - 1)])
#define EBG_BX(arr, a, col) (s.arr[(size_t)((col) - 1) * (s.xnum + 1) + ((a) - 1)])
#define EBG_BY(arr, a, col) (s.arr[(size_t)((col) - 1) * (s.ynum + 1) + ((a) - 1)])
#define EBG_V1(arr, k) (s.arr[(k) - 1])

__device__ __forceinline__ int cell_of(const StokesOpDev& s, int i, int j)
{
    return (j - 1) * (s.ynum - 1) + (i - 1);  // 0-based cell index
}

__device__ __forceinline__ bool out_zone_x(const StokesOpDev& s, int i, int j)
{
    return (j != s.zx_j || i < s.zx_i0 || i > s.zx_i1);
}
__device__ __forceinline__ bool out_zone_y(const StokesOpDev& s, int i, int j)
{
    return (j != s.zy_j || i < s.zy_i0 || i > s.zy_i1);
}
__device__ __forceinline__ bool internal_continuity(const StokesOpDev& s, int i, int j)
{
    if (s.pressure_bc_mode == 0) return (j > 1 || i > 1);
    if (s.pressure_bc_mode == 1) return (1 < i && i < s.ynum - 1);
    if (s.pressure_bc_mode == 2) return (1 < j && j < s.xnum - 1);
    return false;
}

__device__ __forceinline__ double kbond(const StokesOpDev& s)
{
    return 2.0 * s.pscale / (s.xstpavg + s.ystpavg);
}

// ---- x-Stokes row of cell (i, j): returns L*x (WANT_R=false) or R (true) ----
template <bool WANT_R, class Lay>
__device__ double x_row(const StokesOpDev& s, int i, int j, const double* __restrict__ x, Lay L)
{
    const int xnum = s.xnum, ynum = s.ynum;
    const int c = cell_of(s, i, j);
    if (!(j < xnum - 1 && out_zone_x(s, i, j))) {
        if (WANT_R) return out_zone_x(s, i, j) ? 0.0 : 2.0 * s.pscale / (s.xstpavg + s.ystpavg) * s.zvx;
        return kbond(s) * x[L.vx(c)];
    }
    const bool top = (i == 1), low = (i >= ynum - 1), left = (j == 1), right2 = (j >= xnum - 2);
    const double xc1 = EBG_V1(xstpc, j + 1), yi = EBG_V1(ystp, i);
    if (WANT_R) {
        double R = s.RX[(size_t)j * (ynum + 1) + i];  // RX1[i+1, j+1]
        if (left) R = R - EBG_BY(bleftx, i + 1, 1) * 2.0 * EBG_N(, i, j) / EBG_V1(xstp, j) / xc1;
        if (right2) R = R - EBG_BY(brightx, i + 1, 1) * 2.0 * EBG_N(, i, j + 1) / EBG_V1(xstp, j + 1) / xc1;
        if (top) R = R - EBG_BX(btopx, j + 1, 1) * EBG_S(, i, j + 1) / EBG_V1(ystpc, i) / yi;
        if (low) R = R - EBG_BX(bbottomx, j + 1, 1) * EBG_S(, i + 1, j + 1) / EBG_V1(ystpc, i + 1) / yi;
        if (top) R = R - EBG_BX(btopy, j + 1, 1) * EBG_S(, i, j + 1) / xc1 / yi;
        if (top) R = R + EBG_BX(btopy, j + 2, 1) * EBG_S(, i, j + 1) / xc1 / yi;
        if (low) R = R + EBG_BX(bbottomy, j + 1, 1) * EBG_S(, i + 1, j + 1) / xc1 / yi;
        if (low) R = R - EBG_BX(bbottomy, j + 2, 1) * EBG_S(, i + 1, j + 1) / xc1 / yi;
        return R;
    }
    // central coefficient with folds in node order Left, Right, Upper, Bottom
    double cC = -2.0 * (EBG_N(, i, j + 1) / EBG_V1(xstp, j + 1) + EBG_N(, i, j) / EBG_V1(xstp, j)) / xc1 -
                (EBG_S(, i + 1, j + 1) / EBG_V1(ystpc, i + 1) + EBG_S(, i, j + 1) / EBG_V1(ystpc, i)) / yi;
    if (left) cC = cC + EBG_BY(bleftx, i + 1, 2) * 2.0 * EBG_N(, i, j) / EBG_V1(xstp, j) / xc1;
    if (right2) cC = cC + EBG_BY(brightx, i + 1, 2) * 2.0 * EBG_N(, i, j + 1) / EBG_V1(xstp, j + 1) / xc1;
    if (top) cC = cC + EBG_BX(btopx, j + 1, 2) * EBG_S(, i, j + 1) / EBG_V1(ystpc, i) / yi;
    if (low) cC = cC + EBG_BX(bbottomx, j + 1, 2) * EBG_S(, i + 1, j + 1) / EBG_V1(ystpc, i + 1) / yi;
    double y = 0.0;
    y += cC * x[L.vx(c)];
    if (!left) y += (2.0 * EBG_N(, i, j) / EBG_V1(xstp, j) / xc1) * x[L.vx(c - (ynum - 1))];
    if (!right2) y += (2.0 * EBG_N(, i, j + 1) / EBG_V1(xstp, j + 1) / xc1) * x[L.vx(c + (ynum - 1))];
    if (!top) y += (EBG_S(, i, j + 1) / EBG_V1(ystpc, i) / yi) * x[L.vx(c - 1)];
    if (!low) y += (EBG_S(, i + 1, j + 1) / EBG_V1(ystpc, i + 1) / yi) * x[L.vx(c + 1)];
    if (!top) {
        double vUL = EBG_S(, i, j + 1) / xc1 / yi;
        if (low) vUL = vUL - EBG_BX(bbottomy, j + 1, 2) * EBG_S(, i + 1, j + 1) / xc1 / yi;
        y += vUL * x[L.vy(c - 1)];
        double vUR = -EBG_S(, i, j + 1) / xc1 / yi;
        if (low) vUR = vUR + EBG_BX(bbottomy, j + 2, 2) * EBG_S(, i + 1, j + 1) / xc1 / yi;
        y += vUR * x[L.vy(c - 1 + (ynum - 1))];
        if (!low) {
            y += (-EBG_S(, i + 1, j + 1) / xc1 / yi) * x[L.vy(c)];
            y += (EBG_S(, i + 1, j + 1) / xc1 / yi) * x[L.vy(c + (ynum - 1))];
        }
    } else {
        double vLL = EBG_BX(btopy, j + 1, 2) * EBG_S(, i, j + 1) / xc1 / yi;
        if (!low) vLL = vLL - EBG_S(, i + 1, j + 1) / xc1 / yi;
        y += vLL * x[L.vy(c)];
        double vLR = -EBG_BX(btopy, j + 2, 2) * EBG_S(, i, j + 1) / xc1 / yi;
        if (!low) vLR = vLR + EBG_S(, i + 1, j + 1) / xc1 / yi;
        y += vLR * x[L.vy(c + (ynum - 1))];
    }
    y += (s.pscale / xc1) * x[L.p(c)];
    y += (-s.pscale / xc1) * x[L.p(c + (ynum - 1))];
    return y;
}

// ---- y-Stokes row (with drunken-sailor terms) ----
template <bool WANT_R, class Lay>
__device__ double y_row(const StokesOpDev& s, int i, int j, const double* __restrict__ x, Lay L)
{
    const int xnum = s.xnum, ynum = s.ynum;
    const int c = cell_of(s, i, j);
    if (!(i < ynum - 1 && out_zone_y(s, i, j))) {
        if (WANT_R) return out_zone_y(s, i, j) ? 0.0 : 2.0 * s.pscale / (s.xstpavg + s.ystpavg) * s.zvy;
        return kbond(s) * x[L.vy(c)];
    }
    double dy_g = 0.0, dx_g = 0.0;
    if (s.stab == 1) {
        const double* rho = s.rho1_vy;
        const int ld = ynum;
#define RHO(a, b) rho[(size_t)((b) - 1) * ld + ((a) - 1)]
        double drhody = ((RHO(i + 2, j + 1) - RHO(i + 1, j + 1)) / EBG_V1(ystp, i + 1) +
                         (RHO(i + 1, j + 1) - RHO(i, j + 1)) / EBG_V1(ystp, i)) / 2.0;
        dy_g = drhody * s.gravity_y * s.timestep;
        double drhodx = ((RHO(i + 1, j + 2) - RHO(i + 1, j + 1)) / EBG_V1(xstpc, j + 1) +
                         (RHO(i + 1, j + 1) - RHO(i + 1, j)) / EBG_V1(xstpc, j)) / 2.0;
        dx_g = drhodx * s.gravity_y * s.timestep;
#undef RHO
    }
    const bool top = (i == 1), low2 = (i >= ynum - 2), left = (j == 1), right = (j >= xnum - 1);
    const double yc1 = EBG_V1(ystpc, i + 1), xj = EBG_V1(xstp, j);
    // the four vx coupling magnitudes (before BC scaling)
    const double aUL = EBG_S(, i + 1, j) / yc1 / xj - dx_g / 4.0;
    const double aLL = -EBG_S(, i + 1, j) / yc1 / xj - dx_g / 4.0;
    const double aUR = -EBG_S(, i + 1, j + 1) / yc1 / xj - dx_g / 4.0;
    const double aLR = EBG_S(, i + 1, j + 1) / yc1 / xj - dx_g / 4.0;
    if (WANT_R) {
        double R = s.RY[(size_t)j * ynum + i];  // RY1[i+1, j+1]
        if (top) R = R - EBG_BX(btopy, j + 1, 1) * 2.0 * EBG_N(, i, j) / EBG_V1(ystp, i) / yc1;
        if (low2) R = R - EBG_BX(bbottomy, j + 1, 1) * 2.0 * EBG_N(, i + 1, j) / EBG_V1(ystp, i + 1) / yc1;
        if (left) R = R - EBG_BY(blefty, i + 1, 1) * EBG_S(, i + 1, j) / EBG_V1(xstpc, j) / xj;
        if (right) R = R - EBG_BY(brighty, i + 1, 1) * EBG_S(, i + 1, j + 1) / EBG_V1(xstpc, j + 1) / xj;
        if (left) R = R - EBG_BY(bleftx, i + 1, 1) * aUL;
        if (left) R = R - EBG_BY(bleftx, i + 2, 1) * aLL;
        if (right) R = R - EBG_BY(brightx, i + 1, 1) * aUR;
        if (right) R = R - EBG_BY(brightx, i + 2, 1) * aLR;
        return R;
    }
    double cC = -2.0 * (EBG_N(, i + 1, j) / EBG_V1(ystp, i + 1) + EBG_N(, i, j) / EBG_V1(ystp, i)) / yc1 -
                (EBG_S(, i + 1, j + 1) / EBG_V1(xstpc, j + 1) + EBG_S(, i + 1, j) / EBG_V1(xstpc, j)) / xj - dy_g;
    if (top) cC = cC + EBG_BX(btopy, j + 1, 2) * 2.0 * EBG_N(, i, j) / EBG_V1(ystp, i) / yc1;
    if (low2) cC = cC + EBG_BX(bbottomy, j + 1, 2) * 2.0 * EBG_N(, i + 1, j) / EBG_V1(ystp, i + 1) / yc1;
    if (left) cC = cC + EBG_BY(blefty, i + 1, 2) * EBG_S(, i + 1, j) / EBG_V1(xstpc, j) / xj;
    if (right) cC = cC + EBG_BY(brighty, i + 1, 2) * EBG_S(, i + 1, j + 1) / EBG_V1(xstpc, j + 1) / xj;
    double y = 0.0;
    y += cC * x[L.vy(c)];
    if (!top) y += (2.0 * EBG_N(, i, j) / EBG_V1(ystp, i) / yc1) * x[L.vy(c - 1)];
    if (!low2) y += (2.0 * EBG_N(, i + 1, j) / EBG_V1(ystp, i + 1) / yc1) * x[L.vy(c + 1)];
    if (!left) y += (EBG_S(, i + 1, j) / EBG_V1(xstpc, j) / xj) * x[L.vy(c - (ynum - 1))];
    if (!right) y += (EBG_S(, i + 1, j + 1) / EBG_V1(xstpc, j + 1) / xj) * x[L.vy(c + (ynum - 1))];
    if (!left) {
        double vUL = aUL;
        if (right) vUL = vUL + EBG_BY(brightx, i + 1, 2) * aUR;
        y += vUL * x[L.vx(c - (ynum - 1))];
        double vLL = aLL;
        if (right) vLL = vLL + EBG_BY(brightx, i + 2, 2) * aLR;
        y += vLL * x[L.vx(c + 1 - (ynum - 1))];
        if (!right) {
            y += aUR * x[L.vx(c)];
            y += aLR * x[L.vx(c + 1)];
        }
    } else {
        double vUR = EBG_BY(bleftx, i + 1, 2) * aUL;
        if (!right) vUR = vUR - EBG_S(, i + 1, j + 1) / yc1 / xj - dx_g / 4.0;
        y += vUR * x[L.vx(c)];
        double vLR = EBG_BY(bleftx, i + 2, 2) * aLL;
        if (!right) vLR = vLR + EBG_S(, i + 1, j + 1) / yc1 / xj - dx_g / 4.0;
        y += vLR * x[L.vx(c + 1)];
    }
    y += (s.pscale / yc1) * x[L.p(c)];
    y += (-s.pscale / yc1) * x[L.p(c + 1)];
    return y;
}

// ---- continuity row ----
template <bool WANT_R, class Lay>
__device__ double p_row(const StokesOpDev& s, int i, int j, const double* __restrict__ x, Lay L)
{
    const int xnum = s.xnum, ynum = s.ynum;
    const int c = cell_of(s, i, j);
    if (!internal_continuity(s, i, j)) {
        if (WANT_R) {
            const double pbc = 2.0 * s.pressure_bc / (s.xstpavg + s.ystpavg);
            if (s.pressure_bc_mode == 0) return pbc;
            if (s.pressure_bc_mode == 1) return (i == 1) ? pbc : 0.0;
            return (j == 1) ? pbc : 0.0;
        }
        return kbond(s) * x[L.p(c)];
    }
    const double xj = EBG_V1(xstp, j), yi = EBG_V1(ystp, i), ps = s.pscale;
    if (WANT_R) {
        double R = s.RC[(size_t)(j - 1) * (ynum - 1) + (i - 1)] * ps;
        if (j != 1 && j >= xnum - 1) R = R - EBG_BY(brightx, i + 1, 1) * ps / xj;
        if (!(j >= xnum - 1) && j == 1) R = R + EBG_BY(bleftx, i + 1, 1) * ps / xj;
        if (i != 1 && i >= ynum - 1) R = R - EBG_BX(bbottomy, j + 1, 1) * ps / yi;
        if (!(i >= ynum - 1) && i == 1) R = R + EBG_BX(btopy, j + 1, 1) * ps / yi;
        return R;
    }
    double y = 0.0;
    if (j != 1) {
        double v = -ps / xj;
        if (j >= xnum - 1) v = v + EBG_BY(brightx, i + 1, 2) * ps / xj;
        y += v * x[L.vx(c - (ynum - 1))];
    }
    if (!(j >= xnum - 1)) {
        double v = ps / xj;
        if (j == 1) v = v - EBG_BY(bleftx, i + 1, 2) * ps / xj;
        y += v * x[L.vx(c)];
    }
    if (i != 1) {
        double v = -ps / yi;
        if (i >= ynum - 1) v = v + EBG_BX(bbottomy, j + 1, 2) * ps / yi;
        y += v * x[L.vy(c - 1)];
    }
    if (!(i >= ynum - 1)) {
        double v = ps / yi;
        if (i == 1) v = v - EBG_BX(btopy, j + 1, 2) * ps / yi;
        y += v * x[L.vy(c)];
    }
    return y;
}

}  // namespace ebg
