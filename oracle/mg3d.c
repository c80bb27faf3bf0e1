/*
 * Oracle (test infrastructure only): the reference's experimental 3-D
 * geometric multigrid (src/stokes_continuity_solver/multigrid/), restated.
 *
 * Sweep order.  The reference smoother (smoother/solver/solve/
 * SolveStokes3dOpt.jl:12-157) visits, per colour (i+j+k)&1, the z-planes k
 * (Threads.@threads when znum >= 16), the columns j and every second row i,
 * updating vx, vy, vz and p of one point in turn.  Same-colour points are
 * coupled through the staggered stencil (vx(i,j,k) reads vy(i-1,j+1,k) and
 * p(i-1,j-1,k-1)), so the threaded result depends on the schedule.  Two
 * orders are restated:
 *   - ebo_mg3_sweep: per colour all vx, then all vy, then all vz, then all p
 *     (data-parallel and deterministic: the order of the HIP kernels);
 *   - ebo_mg3_sweep_serial: the reference's own single-thread order.
 * Every update uses the reference's formulas (Residuals.jl:145-560).
 * Boundary conditions are the free-slip ghost passes of
 * BoundaryConditions.jl:92-174 in the reference's visiting order.
 */
#include "eb_mg3d_oracle.h"
#include <math.h>
#include <stdlib.h>

#define VX(l, i, j, k) EBO_A3((l)->vx, (l)->ynum + 1, (l)->xnum, i, j, k)
#define VY(l, i, j, k) EBO_A3((l)->vy, (l)->ynum, (l)->xnum + 1, i, j, k)
#define VZ(l, i, j, k) EBO_A3((l)->vz, (l)->ynum + 1, (l)->xnum + 1, i, j, k)
#define PR(l, i, j, k) EBO_A3((l)->pr, (l)->ynum - 1, (l)->xnum - 1, i, j, k)
#define EN(l, i, j, k) EBO_A3((l)->etan, (l)->ynum - 1, (l)->xnum - 1, i, j, k)
#define EXY(l, i, j, k) EBO_A3((l)->etaxy, (l)->ynum, (l)->xnum, i, j, k)
#define EXZ(l, i, j, k) EBO_A3((l)->etaxz, (l)->ynum - 1, (l)->xnum, i, j, k)
#define EYZ(l, i, j, k) EBO_A3((l)->etayz, (l)->ynum, (l)->xnum - 1, i, j, k)
#define XB(l, j) EBO_V((l)->xstp_b, j)
#define YB(l, i) EBO_V((l)->ystp_b, i)
#define ZB(l, k) EBO_V((l)->zstp_b, k)
#define YV(l, i) EBO_V((l)->ystp_vx, i)
#define ZV(l, k) EBO_V((l)->zstp_vx, k)
#define XV(l, j) EBO_V((l)->xstp_vy, j)

/* multigrid/utils/Domain.jl:203-234 (on_v*_boundary3d_limits) */
static int on_vx_bnd(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k)
{
    return i == 1 || i == l->ynum + 1 || j == 1 || j == l->xnum || k == 1 || k == l->znum + 1;
}
static int on_vy_bnd(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k)
{
    return i == 1 || i == l->ynum || j == 1 || j == l->xnum + 1 || k == 1 || k == l->znum + 1;
}
static int on_vz_bnd(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k)
{
    return i == 1 || i == l->ynum + 1 || j == 1 || j == l->xnum + 1 || k == 1 || k == l->znum;
}

/* Residuals.jl:145-225 calculate_vx_residual (3-D) */
double ebo_mg3_vx_residual(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k, double *coef)
{
    const double dxL = XB(l, j - 1), dxR = XB(l, j), dxC = XV(l, j);
    const double dyU = YV(l, i - 1), dyD = YV(l, i), dyC = YB(l, i - 1);
    const double dzF = ZV(l, k), dzB = ZV(l, k - 1), dzC = ZB(l, k - 1);
    const double nR = EN(l, i - 1, j, k - 1), nL = EN(l, i - 1, j - 1, k - 1);
    const double xyU = EXY(l, i - 1, j, k - 1), xyD = EXY(l, i, j, k - 1);
    const double xzF = EXZ(l, i - 1, j, k), xzB = EXZ(l, i - 1, j, k - 1);
    const double pL = PR(l, i - 1, j - 1, k - 1), pR = PR(l, i - 1, j, k - 1);
    const double vxL = VX(l, i, j - 1, k), vxC = VX(l, i, j, k), vxR = VX(l, i, j + 1, k);
    const double vxU = VX(l, i - 1, j, k), vxD = VX(l, i + 1, j, k);
    const double vxF = VX(l, i, j, k + 1), vxB = VX(l, i, j, k - 1);
    const double vyUL = VY(l, i - 1, j, k), vyUR = VY(l, i - 1, j + 1, k);
    const double vyLL = VY(l, i, j, k), vyLR = VY(l, i, j + 1, k);
    const double vzRF = VZ(l, i, j + 1, k), vzLF = VZ(l, i, j, k);
    const double vzRB = VZ(l, i, j + 1, k - 1), vzLB = VZ(l, i, j, k - 1);
    const double dp_dx = (pR - pL) / dxC;
    const double dsxx = 2.0 * (nR * (vxR - vxC) / dxR - nL * (vxC - vxL) / dxL) / dxC;
    const double dsxy = (xyD * ((vxD - vxC) / dyD + (vyLR - vyLL) / dxC) -
                         xyU * ((vxC - vxU) / dyU + (vyUR - vyUL) / dxC)) / dyC;
    const double dsxz = (xzF * ((vzRF - vzLF) / dxC + (vxF - vxC) / dzF) -
                         xzB * ((vzRB - vzLB) / dxC + (vxC - vxB) / dzB)) / dzC;
    const double dR = EBO_A3(l->RX, l->ynum + 1, l->xnum, i, j, k) + dp_dx - dsxx - dsxy - dsxz;
    if (coef)
        *coef = -2.0 * (nR / dxR / dxC + nL / dxL / dxC) - (xyD / dyD / dyC + xyU / dyU / dyC) -
                (xzF / dzF / dzC + xzB / dzB / dzC);
    return dR;
}

/* Residuals.jl:289-371 calculate_vy_residual (3-D) */
double ebo_mg3_vy_residual(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k, double *coef)
{
    const double dxL = XB(l, j - 1), dxR = XV(l, j), dxC = XV(l, j - 1);
    const double dyU = YB(l, i - 1), dyD = YB(l, i), dyC = YV(l, i);
    const double dzF = ZV(l, k), dzB = ZV(l, k - 1), dzC = ZB(l, k - 1);
    const double pD = PR(l, i, j - 1, k - 1), pU = PR(l, i - 1, j - 1, k - 1);
    const double nD = EN(l, i, j - 1, k - 1), nU = EN(l, i - 1, j - 1, k - 1);
    const double xyR = EXY(l, i, j, k - 1), xyL = EXY(l, i, j - 1, k - 1);
    const double yzF = EYZ(l, i, j - 1, k), yzB = EYZ(l, i, j - 1, k - 1);
    const double vyU = VY(l, i - 1, j, k), vyC = VY(l, i, j, k), vyD = VY(l, i + 1, j, k);
    const double vyR = VY(l, i, j + 1, k), vyL = VY(l, i, j - 1, k);
    const double vyF = VY(l, i, j, k + 1), vyB = VY(l, i, j, k - 1);
    const double vxLR = VX(l, i + 1, j, k), vxUR = VX(l, i, j, k);
    const double vxLL = VX(l, i + 1, j - 1, k), vxUL = VX(l, i, j - 1, k);
    const double vzDF = VZ(l, i + 1, j, k), vzUF = VZ(l, i, j, k);
    const double vzDB = VZ(l, i + 1, j, k - 1), vzUB = VZ(l, i, j, k - 1);
    const double dp_dy = (pD - pU) / dyC;
    const double dsyy = 2.0 * (nD * (vyD - vyC) / dyD - nU * (vyC - vyU) / dyU) / dyC;
    const double dsyx = (xyR * ((vyR - vyC) / dxR + (vxLR - vxUR) / dyC) -
                         xyL * ((vyC - vyL) / dxL + (vxLL - vxUL) / dyC)) / dxC;
    const double dsyz = (yzF * ((vyF - vyC) / dzF + (vzDF - vzUF) / dyC) -
                         yzB * ((vyC - vyB) / dzB + (vzDB - vzUB) / dyC)) / dzC;
    const double dR = EBO_A3(l->RY, l->ynum, l->xnum + 1, i, j, k) + dp_dy - dsyy - dsyx - dsyz;
    if (coef)
        *coef = -2.0 * (nD / dyD / dyC + nU / dyU / dyC) - (xyR / dxR / dxC + xyL / dxL / dxC) -
                (yzF / dzF / dzC + yzB / dzB / dzC);
    return dR;
}

/* Residuals.jl:435-521 calculate_vz_residual (3-D) */
double ebo_mg3_vz_residual(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k, double *coef)
{
    const double dzR = ZB(l, k), dzL = ZB(l, k - 1), dzC = ZV(l, k);
    const double dxF = XV(l, j), dxB = XV(l, j - 1), dxC = XB(l, j - 1);
    const double dyU = YV(l, i - 1), dyD = YV(l, i), dyC = YB(l, i - 1);
    const double pF = PR(l, i - 1, j - 1, k), pB = PR(l, i - 1, j - 1, k - 1);
    const double nF = EN(l, i - 1, j - 1, k), nB = EN(l, i - 1, j - 1, k - 1);
    const double xzR = EXZ(l, i - 1, j, k), xzL = EXZ(l, i - 1, j - 1, k);
    const double yzU = EYZ(l, i - 1, j - 1, k), yzD = EYZ(l, i, j - 1, k);
    const double vxRF = VX(l, i, j, k + 1), vxLF = VX(l, i, j, k);
    const double vxRB = VX(l, i, j - 1, k + 1), vxLB = VX(l, i, j - 1, k);
    const double vyLR = VY(l, i, j, k + 1), vyLL = VY(l, i, j, k);
    const double vyUR = VY(l, i - 1, j, k + 1), vyUL = VY(l, i - 1, j, k);
    const double vzL = VZ(l, i, j, k - 1), vzC = VZ(l, i, j, k), vzR = VZ(l, i, j, k + 1);
    const double vzF = VZ(l, i, j + 1, k), vzB = VZ(l, i, j - 1, k);
    const double vzD = VZ(l, i + 1, j, k), vzU = VZ(l, i - 1, j, k);
    const double dp_dz = (pF - pB) / dzC;
    const double dszz = 2.0 * (nF * (vzR - vzC) / dzR - nB * (vzC - vzL) / dzL) / dzC;
    const double dszx = (xzR * ((vxRF - vxLF) / dzC + (vzF - vzC) / dxF) -
                         xzL * ((vxRB - vxLB) / dzC + (vzC - vzB) / dxB)) / dxC;
    const double dszy = (yzD * ((vzD - vzC) / dyD + (vyLR - vyLL) / dzC) -
                         yzU * ((vzC - vzU) / dyU + (vyUR - vyUL) / dzC)) / dyC;
    const double dR = EBO_A3(l->RZ, l->ynum + 1, l->xnum + 1, i, j, k) + dp_dz - dszz - dszx - dszy;
    if (coef)
        *coef = -2.0 * (nF / dzR / dzC + nB / dzL / dzC) - (xzR / dxF / dxC + xzL / dxB / dxC) -
                (yzD / dyD / dyC + yzU / dyU / dyC);
    return dR;
}

/* Residuals.jl:525-560 calculate_pressure_residual (3-D) */
double ebo_mg3_pr_residual(const ebo_mg3level *l, int64_t i, int64_t j, int64_t k)
{
    const double dxC = XB(l, j), dyC = YB(l, i), dzC = ZB(l, k);
    const double vxR = VX(l, i + 1, j + 1, k + 1), vxL = VX(l, i + 1, j, k + 1);
    const double vyD = VY(l, i + 1, j + 1, k + 1), vyU = VY(l, i, j + 1, k + 1);
    const double vzF = VZ(l, i + 1, j + 1, k + 1), vzB = VZ(l, i + 1, j + 1, k);
    return EBO_A3(l->RC, l->ynum - 1, l->xnum - 1, i, j, k) -
           ((vxR - vxL) / dxC + (vyD - vyU) / dyC + (vzF - vzB) / dzC);
}

/* core/VxBoundaryConditions.jl:6-88 (free slip) */
static void set_vx(ebo_mg3level *l, int64_t i, int64_t j, int64_t k)
{
    const int64_t Y = l->ynum + 1, Z = l->znum + 1;
    VX(l, i, j, k) = 0.0;
    if (i == 1) VX(l, i, j, k) = VX(l, i + 1, j, k);
    if (i == Y) VX(l, i, j, k) = VX(l, i - 1, j, k);
    if (k == 1) VX(l, i, j, k) = VX(l, i, j, k + 1);
    if (k == Z) VX(l, i, j, k) = VX(l, i, j, k - 1);
}
/* core/VyBoundaryConditions.jl:6-86 (free slip) */
static void set_vy(ebo_mg3level *l, int64_t i, int64_t j, int64_t k)
{
    const int64_t X = l->xnum + 1, Z = l->znum + 1;
    VY(l, i, j, k) = 0.0;
    if (j == 1) VY(l, i, j, k) = VY(l, i, j + 1, k);
    if (j == X) VY(l, i, j, k) = VY(l, i, j - 1, k);
    if (k == 1) VY(l, i, j, k) = VY(l, i, j, k + 1);
    if (k == Z) VY(l, i, j, k) = VY(l, i, j, k - 1);
}
/* core/VzBoundaryConditions.jl:6-87 (free slip) */
static void set_vz(ebo_mg3level *l, int64_t i, int64_t j, int64_t k)
{
    const int64_t Y = l->ynum + 1, X = l->xnum + 1;
    VZ(l, i, j, k) = 0.0;
    if (i == 1) VZ(l, i, j, k) = VZ(l, i + 1, j, k);
    if (i == Y) VZ(l, i, j, k) = VZ(l, i - 1, j, k);
    if (j == 1) VZ(l, i, j, k) = VZ(l, i, j + 1, k);
    if (j == X) VZ(l, i, j, k) = VZ(l, i, j - 1, k);
}

/* boundary_conditions/BoundaryConditions.jl:92-174: six face loops per
 * component, each boundary node visited once, in place. */
void ebo_mg3_set_bc(ebo_mg3level *l)
{
    const int64_t xnum = l->xnum, ynum = l->ynum, znum = l->znum;
    /* vx: i in 1..ynum+1, j in 1..xnum, k in 1..znum+1 */
    {
        const int64_t Y = ynum + 1, X = xnum, Z = znum + 1;
        for (int64_t j = 1; j <= X; ++j)
            for (int64_t i = 1; i <= Y; ++i) { set_vx(l, i, j, 1); set_vx(l, i, j, Z); }
        for (int64_t k = 2; k <= Z - 1; ++k)
            for (int64_t j = 1; j <= X; ++j) { set_vx(l, 1, j, k); set_vx(l, Y, j, k); }
        for (int64_t k = 2; k <= Z - 1; ++k)
            for (int64_t i = 2; i <= Y - 1; ++i) { set_vx(l, i, 1, k); set_vx(l, i, X, k); }
    }
    /* vy: i in 1..ynum, j in 1..xnum+1, k in 1..znum+1 */
    {
        const int64_t Y = ynum, X = xnum + 1, Z = znum + 1;
        for (int64_t j = 1; j <= X; ++j)
            for (int64_t i = 1; i <= Y; ++i) { set_vy(l, i, j, 1); set_vy(l, i, j, Z); }
        for (int64_t k = 2; k <= Z - 1; ++k)
            for (int64_t j = 1; j <= X; ++j) { set_vy(l, 1, j, k); set_vy(l, Y, j, k); }
        for (int64_t k = 2; k <= Z - 1; ++k)
            for (int64_t i = 2; i <= Y - 1; ++i) { set_vy(l, i, 1, k); set_vy(l, i, X, k); }
    }
    /* vz: i in 1..ynum+1, j in 1..xnum+1, k in 1..znum */
    {
        const int64_t Y = ynum + 1, X = xnum + 1, Z = znum;
        for (int64_t j = 1; j <= X; ++j)
            for (int64_t i = 1; i <= Y; ++i) { set_vz(l, i, j, 1); set_vz(l, i, j, Z); }
        for (int64_t k = 2; k <= Z - 1; ++k)
            for (int64_t j = 1; j <= X; ++j) { set_vz(l, 1, j, k); set_vz(l, Y, j, k); }
        for (int64_t k = 2; k <= Z - 1; ++k)
            for (int64_t i = 2; i <= Y - 1; ++i) { set_vz(l, i, 1, k); set_vz(l, i, X, k); }
    }
}

/* SolveStokes3dOpt.jl:108-157 update_vx!/update_vy!/update_vz!/update_pressure! */
static void upd_vx(ebo_mg3level *l, int64_t i, int64_t j, int64_t k, double ts)
{
    double coef, dR = ebo_mg3_vx_residual(l, i, j, k, &coef);
    VX(l, i, j, k) += dR / coef * ts;
}
static void upd_vy(ebo_mg3level *l, int64_t i, int64_t j, int64_t k, double ts)
{
    double coef, dR = ebo_mg3_vy_residual(l, i, j, k, &coef);
    VY(l, i, j, k) += dR / coef * ts;
}
static void upd_vz(ebo_mg3level *l, int64_t i, int64_t j, int64_t k, double ts)
{
    double coef, dR = ebo_mg3_vz_residual(l, i, j, k, &coef);
    VZ(l, i, j, k) += dR / coef * ts;
}
static void upd_pr(ebo_mg3level *l, int64_t i, int64_t j, int64_t k, double tc)
{
    double dR = ebo_mg3_pr_residual(l, i, j, k);
    PR(l, i, j, k) += EN(l, i, j, k) * dR * tc;
}

/* One sweep, both colours, phased order (see the file comment). */
void ebo_mg3_sweep(ebo_mg3level *l, double theta_s, double theta_c)
{
    const int64_t xnum = l->xnum, ynum = l->ynum, znum = l->znum;
    for (int color = 0; color <= 1; ++color) {
        for (int64_t k = 2; k <= znum; ++k)
            for (int64_t j = 2; j <= xnum - 1; ++j)
                for (int64_t i = 2; i <= ynum; ++i)
                    if (((i + j + k) & 1) == color) upd_vx(l, i, j, k, theta_s);
        for (int64_t k = 2; k <= znum; ++k)
            for (int64_t j = 2; j <= xnum; ++j)
                for (int64_t i = 2; i <= ynum - 1; ++i)
                    if (((i + j + k) & 1) == color) upd_vy(l, i, j, k, theta_s);
        for (int64_t k = 2; k <= znum - 1; ++k)
            for (int64_t j = 2; j <= xnum; ++j)
                for (int64_t i = 2; i <= ynum; ++i)
                    if (((i + j + k) & 1) == color) upd_vz(l, i, j, k, theta_s);
        for (int64_t k = 1; k <= znum - 1; ++k)
            for (int64_t j = 1; j <= xnum - 1; ++j)
                for (int64_t i = 1; i <= ynum - 1; ++i)
                    if (((i + j + k) & 1) == color) upd_pr(l, i, j, k, theta_c);
    }
}

/* SolveStokes3dOpt.jl:12-54 and :56-105 with one thread: per colour, planes
 * k, columns j, rows i of that colour; vx, vy, vz, p of each point in turn. */
void ebo_mg3_sweep_serial(ebo_mg3level *l, double theta_s, double theta_c)
{
    const int64_t xnum = l->xnum, ynum = l->ynum, znum = l->znum;
    for (int color = 0; color <= 1; ++color)
        for (int64_t k = 1; k <= znum + 1; ++k)
            for (int64_t j = 1; j <= xnum + 1; ++j) {
                const int64_t i0 = (((1 + j + k) & 1) == color) ? 1 : 2;
                for (int64_t i = i0; i <= ynum + 1; i += 2) {
                    if (j < xnum + 1 && !on_vx_bnd(l, i, j, k)) upd_vx(l, i, j, k, theta_s);
                    if (i < ynum + 1 && !on_vy_bnd(l, i, j, k)) upd_vy(l, i, j, k, theta_s);
                    if (k < znum + 1 && !on_vz_bnd(l, i, j, k)) upd_vz(l, i, j, k, theta_s);
                    if (i < ynum && j < xnum && k < znum) upd_pr(l, i, j, k, theta_c);
                }
            }
}

/* pressure_update/PressureUpdate.jl:5-21 */
void ebo_mg3_normalize_pressure(ebo_mg3level *l, double prnorm)
{
    const int64_t n = (l->ynum - 1) * (l->xnum - 1) * (l->znum - 1);
    const double dp = prnorm - l->pr[0];
    for (int64_t q = 0; q < n; ++q) l->pr[q] += dp;
}

/* Residuals.jl:10-100 (compute_residuals! 3-D): interior residuals, zeros on
 * boundary nodes. */
void ebo_mg3_residuals(ebo_mg3level *l)
{
    const int64_t xnum = l->xnum, ynum = l->ynum, znum = l->znum;
    for (int64_t k = 1; k <= znum + 1; ++k)
        for (int64_t j = 1; j <= xnum + 1; ++j)
            for (int64_t i = 1; i <= ynum + 1; ++i) {
                if (j < xnum + 1)
                    EBO_A3(l->rx, ynum + 1, xnum, i, j, k) =
                        on_vx_bnd(l, i, j, k) ? 0.0 : ebo_mg3_vx_residual(l, i, j, k, NULL);
                if (i < ynum + 1)
                    EBO_A3(l->ry, ynum, xnum + 1, i, j, k) =
                        on_vy_bnd(l, i, j, k) ? 0.0 : ebo_mg3_vy_residual(l, i, j, k, NULL);
                if (k < znum + 1)
                    EBO_A3(l->rz, ynum + 1, xnum + 1, i, j, k) =
                        on_vz_bnd(l, i, j, k) ? 0.0 : ebo_mg3_vz_residual(l, i, j, k, NULL);
                if (i < ynum && j < xnum && k < znum)
                    EBO_A3(l->rc, ynum - 1, xnum - 1, i, j, k) = ebo_mg3_pr_residual(l, i, j, k);
            }
}

/* Velocity block only (no reference counterpart: the smoother of the
 * block-triangular preconditioner's velocity V-cycle): the sweep without the
 * pressure phase, so the momentum equations are relaxed with p held fixed. */
void ebo_mg3_smoother_velocity(ebo_mg3level *l, int64_t nsweeps, double theta_s)
{
    const int64_t xnum = l->xnum, ynum = l->ynum, znum = l->znum;
    for (int64_t s = 0; s < nsweeps; ++s) {
        ebo_mg3_set_bc(l);
        for (int color = 0; color <= 1; ++color) {
            for (int64_t k = 2; k <= znum; ++k)
                for (int64_t j = 2; j <= xnum - 1; ++j)
                    for (int64_t i = 2; i <= ynum; ++i)
                        if (((i + j + k) & 1) == color) upd_vx(l, i, j, k, theta_s);
            for (int64_t k = 2; k <= znum; ++k)
                for (int64_t j = 2; j <= xnum; ++j)
                    for (int64_t i = 2; i <= ynum - 1; ++i)
                        if (((i + j + k) & 1) == color) upd_vy(l, i, j, k, theta_s);
            for (int64_t k = 2; k <= znum - 1; ++k)
                for (int64_t j = 2; j <= xnum; ++j)
                    for (int64_t i = 2; i <= ynum; ++i)
                        if (((i + j + k) & 1) == color) upd_vz(l, i, j, k, theta_s);
        }
    }
    ebo_mg3_residuals(l);
}

/* smoother/Smoother.jl:50-64 with StokesSolverGaussSeidel.jl:16-28:
 * nsweeps x [BC pass; sweep], pressure normalisation, residuals. */
void ebo_mg3_smoother(ebo_mg3level *l, int64_t nsweeps, double theta_s, double theta_c,
                      double prnorm, int serial)
{
    for (int64_t s = 0; s < nsweeps; ++s) {
        ebo_mg3_set_bc(l);
        if (serial) ebo_mg3_sweep_serial(l, theta_s, theta_c);
        else ebo_mg3_sweep(l, theta_s, theta_c);
    }
    ebo_mg3_normalize_pressure(l, prnorm);
    ebo_mg3_residuals(l);
}

/* TriLinearInterpolation.jl:23-70 add_to_numerator_and_denominator! */
static void add_num_den3(int64_t n1, int64_t n2, double *num, double *den, int32_t iy, int32_t jx,
                         int32_t kz, double dy, double dx, double dz, double val)
{
    const double omx = 1.0 - dx, omy = 1.0 - dy, omz = 1.0 - dz;
    const double w000 = omx * omy * omz, w001 = omx * omy * dz, w010 = omx * dy * omz,
                 w011 = omx * dy * dz, w100 = dx * omy * omz, w101 = dx * omy * dz,
                 w110 = dx * dy * omz, w111 = dx * dy * dz;
#define ACC(di, dj, dk, w)                                           \
    do {                                                             \
        EBO_A3(num, n1, n2, iy + di, jx + dj, kz + dk) += (w) * val; \
        EBO_A3(den, n1, n2, iy + di, jx + dj, kz + dk) += (w);       \
    } while (0)
    ACC(0, 0, 0, w000);
    ACC(0, 0, 1, w001);
    ACC(1, 0, 0, w010);
    ACC(1, 0, 1, w011);
    ACC(0, 1, 0, w100);
    ACC(0, 1, 1, w101);
    ACC(1, 1, 0, w110);
    ACC(1, 1, 1, w111);
#undef ACC
}

/* TriLinearInterpolation.jl:72-107 coarse2fine_trilinear_interpolation! */
static double c2f3(int64_t n1, int64_t n2, const double *a, int32_t iy, int32_t jx, int32_t kz,
                   double dy, double dx, double dz)
{
    const double omx = 1.0 - dx, omy = 1.0 - dy, omz = 1.0 - dz;
    const double w000 = omx * omy * omz, w001 = omx * omy * dz, w010 = omx * dy * omz,
                 w011 = omx * dy * dz, w100 = dx * omy * omz, w101 = dx * omy * dz,
                 w110 = dx * dy * omz, w111 = dx * dy * dz;
    double v = 0.0;
    v += w000 * EBO_A3(a, n1, n2, iy, jx, kz);
    v += w001 * EBO_A3(a, n1, n2, iy, jx, kz + 1);
    v += w010 * EBO_A3(a, n1, n2, iy + 1, jx, kz);
    v += w011 * EBO_A3(a, n1, n2, iy + 1, jx, kz + 1);
    v += w100 * EBO_A3(a, n1, n2, iy, jx + 1, kz);
    v += w101 * EBO_A3(a, n1, n2, iy, jx + 1, kz + 1);
    v += w110 * EBO_A3(a, n1, n2, iy + 1, jx + 1, kz);
    v += w111 * EBO_A3(a, n1, n2, iy + 1, jx + 1, kz + 1);
    return v;
}

enum { KB = 0, KV = 1, KP = 2, KS = 3 };
#define MAPARGS(m, ky, kx, kzz, i, j, k)                                                   \
    EBO_V((m)->iy[ky], i), EBO_V((m)->jx[kx], j), EBO_V((m)->kz[kzz], k), EBO_V((m)->dy[ky], i), \
        EBO_V((m)->dx[kx], j), EBO_V((m)->dz[kzz], k)

/* restriction/Restriction.jl:64-376 (3-D, one thread): fine nodes visited
 * k, j, i from 2; continuity residuals scaled by the fine etan before and by
 * the coarse etan after; wt* are scratch arrays of the coarse shapes. */
void ebo_mg3_restrict(const ebo_mg3level *f, ebo_mg3level *c, const ebo_mg3map *m,
                      double *wtx, double *wty, double *wtz, double *wtc)
{
    const int64_t xf = f->xnum, yf = f->ynum, zf = f->znum;
    const int64_t xc = c->xnum, yc = c->ynum, zc = c->znum;
    const int64_t nx = (yc + 1) * xc * (zc + 1), ny = yc * (xc + 1) * (zc + 1),
                  nz = (yc + 1) * (xc + 1) * zc, np = (yc - 1) * (xc - 1) * (zc - 1);
    for (int64_t q = 0; q < nx; ++q) { c->RX[q] = 0.0; wtx[q] = 0.0; }
    for (int64_t q = 0; q < ny; ++q) { c->RY[q] = 0.0; wty[q] = 0.0; }
    for (int64_t q = 0; q < nz; ++q) { c->RZ[q] = 0.0; wtz[q] = 0.0; }
    for (int64_t q = 0; q < np; ++q) { c->RC[q] = 0.0; wtc[q] = 0.0; }
    for (int64_t k = 2; k <= zf; ++k)
        for (int64_t j = 2; j <= xf; ++j)
            for (int64_t i = 2; i <= yf; ++i) {
                if (j < xf)
                    add_num_den3(yc + 1, xc, c->RX, wtx, MAPARGS(m, KV, KB, KV, i, j, k),
                                 EBO_A3(f->rx, yf + 1, xf, i, j, k));
                if (i < yf)
                    add_num_den3(yc, xc + 1, c->RY, wty, MAPARGS(m, KB, KV, KV, i, j, k),
                                 EBO_A3(f->ry, yf, xf + 1, i, j, k));
                if (k < zf)
                    add_num_den3(yc + 1, xc + 1, c->RZ, wtz, MAPARGS(m, KV, KV, KB, i, j, k),
                                 EBO_A3(f->rz, yf + 1, xf + 1, i, j, k));
                if (i < yf && j < xf && k < zf) {
                    const double v = EN(f, i, j, k) * EBO_A3(f->rc, yf - 1, xf - 1, i, j, k);
                    add_num_den3(yc - 1, xc - 1, c->RC, wtc, MAPARGS(m, KP, KP, KP, i, j, k), v);
                }
            }
    for (int64_t kc = 1; kc <= zc + 1; ++kc)
        for (int64_t jc = 1; jc <= xc + 1; ++jc)
            for (int64_t ic = 1; ic <= yc + 1; ++ic) {
                if (jc < xc + 1) {
                    double *r = &EBO_A3(c->RX, yc + 1, xc, ic, jc, kc);
                    const double w = EBO_A3(wtx, yc + 1, xc, ic, jc, kc);
                    if (w != 0 && ic > 1 && ic < yc + 1 && jc > 1 && jc < xc && kc > 1 && kc < zc + 1)
                        *r = *r / w;
                    else
                        *r = 0.0;
                }
                if (ic < yc + 1) {
                    double *r = &EBO_A3(c->RY, yc, xc + 1, ic, jc, kc);
                    const double w = EBO_A3(wty, yc, xc + 1, ic, jc, kc);
                    if (w != 0 && ic > 1 && ic < yc && jc > 1 && jc < xc + 1 && kc > 1 && kc < zc + 1)
                        *r = *r / w;
                    else
                        *r = 0.0;
                }
                if (kc < zc + 1) {
                    double *r = &EBO_A3(c->RZ, yc + 1, xc + 1, ic, jc, kc);
                    const double w = EBO_A3(wtz, yc + 1, xc + 1, ic, jc, kc);
                    if (w != 0 && ic > 1 && ic < yc + 1 && jc > 1 && jc < xc + 1 && kc > 1 && kc < zc)
                        *r = *r / w;
                    else
                        *r = 0.0;
                }
                if (ic < yc && jc < xc && kc < zc) {
                    double *r = &EBO_A3(c->RC, yc - 1, xc - 1, ic, jc, kc);
                    const double w = EBO_A3(wtc, yc - 1, xc - 1, ic, jc, kc);
                    if (w != 0) *r = *r / w / EN(c, ic, jc, kc);
                    else *r = 0.0;
                }
            }
}

/* prolongation/Prolongation.jl:60-140 and vcycle/core/SmoothAndProlongate.jl:
 * fine += relax * trilinear(coarse) on every fine node (ghosts included). */
void ebo_mg3_prolong(const ebo_mg3level *c, ebo_mg3level *f, const ebo_mg3map *m,
                     double relax_v, double relax_p)
{
    const int64_t xf = f->xnum, yf = f->ynum, zf = f->znum;
    const int64_t xc = c->xnum, yc = c->ynum;
    for (int64_t k = 1; k <= zf + 1; ++k)
        for (int64_t j = 1; j <= xf + 1; ++j)
            for (int64_t i = 1; i <= yf + 1; ++i) {
                if (j < xf + 1)
                    VX(f, i, j, k) += c2f3(yc + 1, xc, c->vx, MAPARGS(m, KV, KB, KV, i, j, k)) * relax_v;
                if (i < yf + 1)
                    VY(f, i, j, k) += c2f3(yc, xc + 1, c->vy, MAPARGS(m, KB, KV, KV, i, j, k)) * relax_v;
                if (k < zf + 1)
                    VZ(f, i, j, k) += c2f3(yc + 1, xc + 1, c->vz, MAPARGS(m, KV, KV, KB, i, j, k)) * relax_v;
                if (i < yf && j < xf && k < zf)
                    PR(f, i, j, k) += c2f3(yc - 1, xc - 1, c->pr, MAPARGS(m, KP, KP, KP, i, j, k)) * relax_p;
            }
}

/* interpolation/ViscosityRestriction.jl:55-230 (3-D): fine nodes visited
 * i, j, k; coarse = sum(w*eta)/sum(w) where sum(w) != 0. */
void ebo_mg3_viscosity_restrict(const ebo_mg3level *f, ebo_mg3level *c, const ebo_mg3map *m,
                                double *wxy, double *wxz, double *wyz, double *wn)
{
    const int64_t xf = f->xnum, yf = f->ynum, zf = f->znum;
    const int64_t xc = c->xnum, yc = c->ynum, zc = c->znum;
    const int64_t nxy = yc * xc * (zc - 1), nxz = (yc - 1) * xc * zc, nyz = yc * (xc - 1) * zc,
                  nn = (yc - 1) * (xc - 1) * (zc - 1);
    for (int64_t q = 0; q < nxy; ++q) { c->etaxy[q] = 0.0; wxy[q] = 0.0; }
    for (int64_t q = 0; q < nxz; ++q) { c->etaxz[q] = 0.0; wxz[q] = 0.0; }
    for (int64_t q = 0; q < nyz; ++q) { c->etayz[q] = 0.0; wyz[q] = 0.0; }
    for (int64_t q = 0; q < nn; ++q) { c->etan[q] = 0.0; wn[q] = 0.0; }
    for (int64_t i = 1; i <= yf; ++i)
        for (int64_t j = 1; j <= xf; ++j)
            for (int64_t k = 1; k <= zf; ++k) {
                if (k < zf)
                    add_num_den3(yc, xc, c->etaxy, wxy, MAPARGS(m, KB, KB, KS, i, j, k), EXY(f, i, j, k));
                if (i < yf)
                    add_num_den3(yc - 1, xc, c->etaxz, wxz, MAPARGS(m, KS, KB, KB, i, j, k), EXZ(f, i, j, k));
                if (j < xf)
                    add_num_den3(yc, xc - 1, c->etayz, wyz, MAPARGS(m, KB, KS, KB, i, j, k), EYZ(f, i, j, k));
                if (i < yf && j < xf && k < zf)
                    add_num_den3(yc - 1, xc - 1, c->etan, wn, MAPARGS(m, KP, KP, KP, i, j, k), EN(f, i, j, k));
            }
    for (int64_t q = 0; q < nxy; ++q) if (wxy[q] != 0) c->etaxy[q] = c->etaxy[q] / wxy[q];
    for (int64_t q = 0; q < nxz; ++q) if (wxz[q] != 0) c->etaxz[q] = c->etaxz[q] / wxz[q];
    for (int64_t q = 0; q < nyz; ++q) if (wyz[q] != 0) c->etayz[q] = c->etayz[q] / wyz[q];
    for (int64_t q = 0; q < nn; ++q) if (wn[q] != 0) c->etan[q] = c->etan[q] / wn[q];
}

/* utils/CalculateMeanResiduals.jl:14-70 (accumulate_principle_residuals_3d!):
 * out[0..3] = log10(sqrt(mean((r / scale)^2))) of the interior x-, y-, z-Stokes
 * residuals (scale stokesscale) and of the continuity residuals (continscale),
 * summed serially in the order k, j, i. */
void ebo_mg3_mean_residuals(const ebo_mg3level *l, double stokesscale, double continscale, double *out)
{
    const int64_t xnum = l->xnum, ynum = l->ynum, znum = l->znum;
    double sx = 0.0, sy = 0.0, sz = 0.0, sc = 0.0;
    const double inv_ss = 1.0 / stokesscale, inv_cc = 1.0 / continscale;
    for (int64_t k = 1; k <= znum; ++k)
        for (int64_t j = 1; j <= xnum; ++j)
            for (int64_t i = 1; i <= ynum; ++i) {
                if (i > 1 && j > 1 && k > 1 && j < xnum) {
                    const double r = EBO_A3(l->rx, ynum + 1, xnum, i, j, k) * inv_ss;
                    sx += r * r;
                }
                if (i > 1 && j > 1 && k > 1 && i < ynum) {
                    const double r = EBO_A3(l->ry, ynum, xnum + 1, i, j, k) * inv_ss;
                    sy += r * r;
                }
                if (i > 1 && j > 1 && k > 1 && k < znum) {
                    const double r = EBO_A3(l->rz, ynum + 1, xnum + 1, i, j, k) * inv_ss;
                    sz += r * r;
                }
                if (j < xnum && i < ynum && k < znum) {
                    const double r = EBO_A3(l->rc, ynum - 1, xnum - 1, i, j, k) * inv_cc;
                    sc += r * r;
                }
            }
    out[0] = log10(sqrt(sx / (double)((ynum - 1) * (xnum - 2) * (znum - 1))));
    out[1] = log10(sqrt(sy / (double)((ynum - 2) * (xnum - 1) * (znum - 1))));
    out[2] = log10(sqrt(sz / (double)((ynum - 1) * (xnum - 1) * (znum - 2))));
    out[3] = log10(sqrt(sc / (double)((ynum - 1) * (xnum - 1) * (znum - 1))));
}

/* Linear operator of the 3-D problem, as the outer Krylov solver sees it:
 * with RX..RC zero on entry, vx..pr hold x; after the BC pass the momentum
 * and continuity parts of A x are the negated residuals (interior nodes,
 * zeros elsewhere), and the continuity entry of cell (1,1,1) is replaced by
 * the pressure gauge p(1,1,1) / etan(1,1,1) (the multigrid's own gauge,
 * PressureUpdate.jl:5-21, written as an equation). */
void ebo_mg3_apply(ebo_mg3level *l)
{
    const int64_t xnum = l->xnum, ynum = l->ynum, znum = l->znum;
    ebo_mg3_set_bc(l);
    ebo_mg3_residuals(l);
    const int64_t nx = (ynum + 1) * xnum * (znum + 1), ny = ynum * (xnum + 1) * (znum + 1),
                  nz = (ynum + 1) * (xnum + 1) * znum, np = (ynum - 1) * (xnum - 1) * (znum - 1);
    for (int64_t q = 0; q < nx; ++q) l->rx[q] = -l->rx[q];
    for (int64_t q = 0; q < ny; ++q) l->ry[q] = -l->ry[q];
    for (int64_t q = 0; q < nz; ++q) l->rz[q] = -l->rz[q];
    for (int64_t q = 0; q < np; ++q) l->rc[q] = -l->rc[q];
    l->rc[0] = l->pr[0] / l->etan[0];
}
