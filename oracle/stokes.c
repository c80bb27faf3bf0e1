/*
 * Oracle (test infrastructure only): the assembled Stokes-continuity system.
 *
 *   ebo_stokes_rhs       K7: viscoelastic terms + RX1/RY1/RC1
 *                        (solver/rhs/StokesViscoelasticTerms.jl:26-108,
 *                         solver/rhs/StokesContinuityRhs.jl:30-190,
 *                         utils/Viscoelastic.jl:26-36)
 *   ebo_stokes_build     K8: COO triplets and RHS in the reference's order
 *                        (solver/build_system/StokesBuildManager.jl:661-706 and
 *                         every discretize_*_equation stencil file)
 *   ebo_stokes_unpack    K10: solution vector -> vx1, vy1, pr1 with velocity BCs
 *                        (solver/system_solver/SystemSolver.jl:141-271)
 *   ebo_coo_residual     K11: |L*x - R| summed in COO triplet order
 *                        (solver/residuals/StokesResiduals.jl:63-82)
 *
 * Paths below are relative to src/stokes_continuity_solver/solver/build_system.
 * Expressions keep the reference's operand order; build with -ffp-contract=off.
 */
#include "eb_stokes_oracle.h"
#include <math.h>

typedef struct {
    int64_t *Li, *Lj;
    double *Lv;
    int64_t inz;  /* next free 1-based slot */
} coo_t;

static int64_t push(coo_t *c, int64_t row, int64_t col, double v)
{
    int64_t k = c->inz;
    c->Li[k - 1] = row;
    c->Lj[k - 1] = col;
    c->Lv[k - 1] = v;
    c->inz = k + 1;
    return k;
}

#define LV(k) (c->Lv[(k) - 1])
#define RR(k) (R[(k) - 1])

/* src/utils/Domain.jl:37-175 */
static int out_zone_x(int64_t i, int64_t j, const int64_t *z)
{
    return (j != z[0] || i < z[1] || i > z[2]);
}
static int out_zone_y(int64_t i, int64_t j, const int64_t *z)
{
    return (j != z[4] || i < z[5] || i > z[6]);
}
static int internal_continuity(int64_t i, int64_t j, int64_t xnum, int64_t ynum, int64_t mode)
{
    if (mode == 0) return (j > 1 || i > 1);
    if (mode == 1) return (1 < i && i < ynum - 1);
    if (mode == 2) return (1 < j && j < xnum - 1);
    return 0;
}

/* discretize_x_stokes_equation/StencilForVxStokesUnknown.jl:35-50 with
 * internal/InternalCells.jl:46-73 (node order) and boundary/BoundaryCells.jl:64-106 */
static void x_stokes(coo_t *c, double *R, const ebo_stokes_in *s, int64_t i, int64_t j, int64_t ivx)
{
    const int64_t xnum = s->xnum, ynum = s->ynum, ny = ynum, nx = xnum;
    const double *xstp = s->xstp, *ystp = s->ystp, *xstpc = s->xstpc, *ystpc = s->ystpc;
    const double *etan = s->etan, *etas = s->etas;
    const int64_t hs = 3 * (ynum - 1);
    const double pscale = s->pscale;
#define ETAN(a, b) EBO_M(etan, ny - 1, a, b)
#define ETAS(a, b) EBO_M(etas, ny, a, b)
#define BX(arr, a, col) EBO_M(arr, nx + 1, a, col)
#define BY(arr, a, col) EBO_M(arr, ny + 1, a, col)
    if (j < xnum - 1 && out_zone_x(i, j, s->bintern_zone)) {
        /* CentralVelocityXNode.jl */
        int64_t inz_c = push(c, ivx, ivx,
            -2.0 * (ETAN(i, j + 1) / EBO_V(xstp, j + 1) + ETAN(i, j) / EBO_V(xstp, j)) / EBO_V(xstpc, j + 1) -
            (ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) + ETAS(i, j + 1) / EBO_V(ystpc, i)) / EBO_V(ystp, i));
        RR(ivx) = EBO_M(s->RX, ny + 1, i + 1, j + 1);
        /* LeftVelocityXNode.jl */
        if (j != 1) {
            push(c, ivx, ivx - hs, 2.0 * ETAN(i, j) / EBO_V(xstp, j) / EBO_V(xstpc, j + 1));
        } else {
            LV(inz_c) = LV(inz_c) + BY(s->bleftx, i + 1, 2) * 2.0 * ETAN(i, j) / EBO_V(xstp, j) / EBO_V(xstpc, j + 1);
            RR(ivx) = RR(ivx) - BY(s->bleftx, i + 1, 1) * 2.0 * ETAN(i, j) / EBO_V(xstp, j) / EBO_V(xstpc, j + 1);
        }
        /* RightVelocityXNode.jl */
        if (!(j >= xnum - 2)) {
            push(c, ivx, ivx + hs, 2.0 * ETAN(i, j + 1) / EBO_V(xstp, j + 1) / EBO_V(xstpc, j + 1));
        } else {
            LV(inz_c) = LV(inz_c) + BY(s->brightx, i + 1, 2) * 2.0 * ETAN(i, j + 1) / EBO_V(xstp, j + 1) / EBO_V(xstpc, j + 1);
            RR(ivx) = RR(ivx) - BY(s->brightx, i + 1, 1) * 2.0 * ETAN(i, j + 1) / EBO_V(xstp, j + 1) / EBO_V(xstpc, j + 1);
        }
        /* UpperVelocityXNode.jl */
        if (i != 1) {
            push(c, ivx, ivx - 3, ETAS(i, j + 1) / EBO_V(ystpc, i) / EBO_V(ystp, i));
        } else {
            LV(inz_c) = LV(inz_c) + BX(s->btopx, j + 1, 2) * ETAS(i, j + 1) / EBO_V(ystpc, i) / EBO_V(ystp, i);
            RR(ivx) = RR(ivx) - BX(s->btopx, j + 1, 1) * ETAS(i, j + 1) / EBO_V(ystpc, i) / EBO_V(ystp, i);
        }
        /* BottomVelocityXNode.jl */
        if (!(i >= ynum - 1)) {
            push(c, ivx, ivx + 3, ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(ystp, i));
        } else {
            LV(inz_c) = LV(inz_c) + BX(s->bbottomx, j + 1, 2) * ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(ystp, i);
            RR(ivx) = RR(ivx) - BX(s->bbottomx, j + 1, 1) * ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(ystp, i);
        }
        /* UpperLeftVelocityYNode.jl */
        int64_t inz_UL = -1, inz_LL = -1, inz_UR = -1, inz_LR = -1;
        if (i != 1) {
            inz_UL = push(c, ivx, ivx - 2, ETAS(i, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i));
        } else {
            inz_LL = push(c, ivx, ivx + 1, BX(s->btopy, j + 1, 2) * ETAS(i, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i));
            RR(ivx) = RR(ivx) - BX(s->btopy, j + 1, 1) * ETAS(i, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i);
        }
        /* UpperRightVelocityYNode.jl */
        if (i != 1) {
            inz_UR = push(c, ivx, ivx - 2 + hs, -ETAS(i, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i));
        } else {
            inz_LR = push(c, ivx, ivx + 1 + hs, -BX(s->btopy, j + 2, 2) * ETAS(i, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i));
            RR(ivx) = RR(ivx) + BX(s->btopy, j + 2, 1) * ETAS(i, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i);
        }
        /* BottomLeftVelocityYNode.jl */
        if (!(i >= ynum - 1)) {
            if (i != 1) push(c, ivx, ivx + 1, -ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i));
            else LV(inz_LL) = LV(inz_LL) - ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i);
        } else {
            LV(inz_UL) = LV(inz_UL) - BX(s->bbottomy, j + 1, 2) * ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i);
            RR(ivx) = RR(ivx) + BX(s->bbottomy, j + 1, 1) * ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i);
        }
        /* BottomRightVelocityYNode.jl */
        if (!(i >= ynum - 1)) {
            if (i != 1) push(c, ivx, ivx + 1 + hs, ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i));
            else LV(inz_LR) = LV(inz_LR) + ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i);
        } else {
            LV(inz_UR) = LV(inz_UR) + BX(s->bbottomy, j + 2, 2) * ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i);
            RR(ivx) = RR(ivx) - BX(s->bbottomy, j + 2, 1) * ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(ystp, i);
        }
        /* LeftAndRightPressureNodes.jl */
        push(c, ivx, ivx + 2, pscale / EBO_V(xstpc, j + 1));
        push(c, ivx, ivx + 2 + hs, -pscale / EBO_V(xstpc, j + 1));
    } else {
        /* boundary/BoundaryCells.jl:64-106 (Kbond row) */
        push(c, ivx, ivx, 2.0 * pscale / (s->xstpavg + s->ystpavg));
        if (out_zone_x(i, j, s->bintern_zone)) RR(ivx) = 0.0;
        else RR(ivx) = 2.0 * pscale / (s->xstpavg + s->ystpavg) * s->bintern_velocity[3];
    }
}

/* discretize_y_stokes_equation: StencilForVyStokesUnknown.jl, internal/InternalCells.jl
 * (node order and drunken-sailor terms :104-168), boundary/BoundaryCells.jl */
static void y_stokes(coo_t *c, double *R, const ebo_stokes_in *s, int64_t i, int64_t j, int64_t ivy)
{
    const int64_t xnum = s->xnum, ynum = s->ynum, ny = ynum, nx = xnum;
    const double *xstp = s->xstp, *ystp = s->ystp, *xstpc = s->xstpc, *ystpc = s->ystpc;
    const double *etan = s->etan, *etas = s->etas;
    const int64_t hs = 3 * (ynum - 1);
    const double pscale = s->pscale;
    (void)nx;
    if (i < ynum - 1 && out_zone_y(i, j, s->bintern_zone)) {
        double dy_g = 0.0, dx_g = 0.0;
        if (s->iuse_interface_stabilization == 1) {
#define RHO(a, b) EBO_M(s->rho1_vy, ny, a, b)
            double drhody = ((RHO(i + 2, j + 1) - RHO(i + 1, j + 1)) / EBO_V(ystp, i + 1) +
                             (RHO(i + 1, j + 1) - RHO(i, j + 1)) / EBO_V(ystp, i)) / 2.0;
            dy_g = drhody * s->gravity_y * s->timestep;
            double drhodx = ((RHO(i + 1, j + 2) - RHO(i + 1, j + 1)) / EBO_V(xstpc, j + 1) +
                             (RHO(i + 1, j + 1) - RHO(i + 1, j)) / EBO_V(xstpc, j)) / 2.0;
            dx_g = drhodx * s->gravity_y * s->timestep;
        }
        /* CentralVelocityYNode.jl */
        int64_t inz_c = push(c, ivy, ivy,
            -2.0 * (ETAN(i + 1, j) / EBO_V(ystp, i + 1) + ETAN(i, j) / EBO_V(ystp, i)) / EBO_V(ystpc, i + 1) -
            (ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) + ETAS(i + 1, j) / EBO_V(xstpc, j)) / EBO_V(xstp, j) - dy_g);
        RR(ivy) = EBO_M(s->RY, ny, i + 1, j + 1);
        /* UpperVelocityYNode.jl */
        if (i != 1) {
            push(c, ivy, ivy - 3, 2.0 * ETAN(i, j) / EBO_V(ystp, i) / EBO_V(ystpc, i + 1));
        } else {
            LV(inz_c) = LV(inz_c) + BX(s->btopy, j + 1, 2) * 2.0 * ETAN(i, j) / EBO_V(ystp, i) / EBO_V(ystpc, i + 1);
            RR(ivy) = RR(ivy) - BX(s->btopy, j + 1, 1) * 2.0 * ETAN(i, j) / EBO_V(ystp, i) / EBO_V(ystpc, i + 1);
        }
        /* BottomVelocityYNode.jl */
        if (!(i >= ynum - 2)) {
            push(c, ivy, ivy + 3, 2.0 * ETAN(i + 1, j) / EBO_V(ystp, i + 1) / EBO_V(ystpc, i + 1));
        } else {
            LV(inz_c) = LV(inz_c) + BX(s->bbottomy, j + 1, 2) * 2.0 * ETAN(i + 1, j) / EBO_V(ystp, i + 1) / EBO_V(ystpc, i + 1);
            RR(ivy) = RR(ivy) - BX(s->bbottomy, j + 1, 1) * 2.0 * ETAN(i + 1, j) / EBO_V(ystp, i + 1) / EBO_V(ystpc, i + 1);
        }
        /* LeftVelocityYNode.jl */
        if (j != 1) {
            push(c, ivy, ivy - hs, ETAS(i + 1, j) / EBO_V(xstpc, j) / EBO_V(xstp, j));
        } else {
            LV(inz_c) = LV(inz_c) + BY(s->blefty, i + 1, 2) * ETAS(i + 1, j) / EBO_V(xstpc, j) / EBO_V(xstp, j);
            RR(ivy) = RR(ivy) - BY(s->blefty, i + 1, 1) * ETAS(i + 1, j) / EBO_V(xstpc, j) / EBO_V(xstp, j);
        }
        /* RightVelocityYNode.jl */
        if (!(j >= xnum - 1)) {
            push(c, ivy, ivy + hs, ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(xstp, j));
        } else {
            LV(inz_c) = LV(inz_c) + BY(s->brighty, i + 1, 2) * ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(xstp, j);
            RR(ivy) = RR(ivy) - BY(s->brighty, i + 1, 1) * ETAS(i + 1, j + 1) / EBO_V(xstpc, j + 1) / EBO_V(xstp, j);
        }
        int64_t inz_UL = -1, inz_UR = -1, inz_LL = -1, inz_LR = -1;
        /* UpperLeftVelocityXNode.jl */
        if (j != 1) {
            inz_UL = push(c, ivy, ivy - 1 - hs, ETAS(i + 1, j) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0);
        } else {
            inz_UR = push(c, ivy, ivy - 1, BY(s->bleftx, i + 1, 2) * (ETAS(i + 1, j) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0));
            inz_UL = 0;
            RR(ivy) = RR(ivy) - BY(s->bleftx, i + 1, 1) * (ETAS(i + 1, j) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0);
        }
        /* BottomLeftVelocityXNode.jl */
        if (j != 1) {
            inz_LL = push(c, ivy, ivy + 2 - hs, -ETAS(i + 1, j) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0);
            inz_LR = 0;
        } else {
            inz_LR = push(c, ivy, ivy + 2, BY(s->bleftx, i + 2, 2) * (-ETAS(i + 1, j) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0));
            inz_LL = 0;
            RR(ivy) = RR(ivy) - BY(s->bleftx, i + 2, 1) * (-ETAS(i + 1, j) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0);
        }
        /* UpperRightVelocityXNode.jl */
        if (!(j >= xnum - 1)) {
            if (j != 1) push(c, ivy, ivy - 1, -ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0);
            else LV(inz_UR) = LV(inz_UR) - ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0;
        } else {
            LV(inz_UL) = LV(inz_UL) + BY(s->brightx, i + 1, 2) * (-ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0);
            RR(ivy) = RR(ivy) - BY(s->brightx, i + 1, 1) * (-ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0);
        }
        /* BottomRightVelocityXNode.jl */
        if (!(j >= xnum - 1)) {
            if (j != 1) push(c, ivy, ivy + 2, ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0);
            else LV(inz_LR) = LV(inz_LR) + ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0;
        } else {
            LV(inz_LL) = LV(inz_LL) + BY(s->brightx, i + 2, 2) * (ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0);
            RR(ivy) = RR(ivy) - BY(s->brightx, i + 2, 1) * (ETAS(i + 1, j + 1) / EBO_V(ystpc, i + 1) / EBO_V(xstp, j) - dx_g / 4.0);
        }
        /* UpperAndLowerPressureNodes.jl */
        push(c, ivy, ivy + 1, pscale / EBO_V(ystpc, i + 1));
        push(c, ivy, ivy + 4, -pscale / EBO_V(ystpc, i + 1));
    } else {
        push(c, ivy, ivy, 2.0 * pscale / (s->xstpavg + s->ystpavg));
        if (out_zone_y(i, j, s->bintern_zone)) RR(ivy) = 0.0;
        else RR(ivy) = 2.0 * pscale / (s->xstpavg + s->ystpavg) * s->bintern_velocity[7];
    }
}

/* discretize_continuity_equation: StencilForPContinuityUnknown.jl:31-53,
 * internal/InternalCells.jl + stencil_nodes, boundary/BoundaryCells.jl:67-203 */
static void continuity(coo_t *c, double *R, const ebo_stokes_in *s, int64_t i, int64_t j, int64_t ipr)
{
    const int64_t xnum = s->xnum, ynum = s->ynum, ny = ynum, nx = xnum;
    const double *xstp = s->xstp, *ystp = s->ystp;
    const int64_t hs = 3 * (ynum - 1);
    const double pscale = s->pscale;
    if (internal_continuity(i, j, xnum, ynum, s->pressure_bc_mode)) {
        RR(ipr) = EBO_M(s->RC, ny - 1, i, j) * pscale;
        if (j != 1) {  /* LeftVelocityXNode.jl */
            int64_t k = push(c, ipr, ipr - 2 - hs, -pscale / EBO_V(xstp, j));
            if (j >= xnum - 1) {
                LV(k) = LV(k) + BY(s->brightx, i + 1, 2) * pscale / EBO_V(xstp, j);
                RR(ipr) = RR(ipr) - BY(s->brightx, i + 1, 1) * pscale / EBO_V(xstp, j);
            }
        }
        if (!(j >= xnum - 1)) {  /* RightVelocityXNode.jl */
            int64_t k = push(c, ipr, ipr - 2, pscale / EBO_V(xstp, j));
            if (j == 1) {
                LV(k) = LV(k) - BY(s->bleftx, i + 1, 2) * pscale / EBO_V(xstp, j);
                RR(ipr) = RR(ipr) + BY(s->bleftx, i + 1, 1) * pscale / EBO_V(xstp, j);
            }
        }
        if (i != 1) {  /* UpperVelocityYNode.jl */
            int64_t k = push(c, ipr, ipr - 4, -pscale / EBO_V(ystp, i));
            if (i >= ynum - 1) {
                LV(k) = LV(k) + BX(s->bbottomy, j + 1, 2) * pscale / EBO_V(ystp, i);
                RR(ipr) = RR(ipr) - BX(s->bbottomy, j + 1, 1) * pscale / EBO_V(ystp, i);
            }
        }
        if (!(i >= ynum - 1)) {  /* BottomVelocityYNode.jl */
            int64_t k = push(c, ipr, ipr - 1, pscale / EBO_V(ystp, i));
            if (i == 1) {
                LV(k) = LV(k) - BX(s->btopy, j + 1, 2) * pscale / EBO_V(ystp, i);
                RR(ipr) = RR(ipr) + BX(s->btopy, j + 1, 1) * pscale / EBO_V(ystp, i);
            }
        }
    } else {
        push(c, ipr, ipr, 2.0 * pscale / (s->xstpavg + s->ystpavg));
        const double pbc = 2.0 * s->pressure_bc / (s->xstpavg + s->ystpavg);
        if (s->pressure_bc_mode == 0) RR(ipr) = pbc;
        else if (s->pressure_bc_mode == 1) RR(ipr) = (i == 1) ? pbc : 0.0;
        else RR(ipr) = (j == 1) ? pbc : 0.0;
    }
    (void)nx;
}

int64_t ebo_stokes_build(const ebo_stokes_in *s, int64_t *Li, int64_t *Lj, double *Lv, double *R)
{
    coo_t c = {Li, Lj, Lv, 1};
    const int64_t ynum = s->ynum, xnum = s->xnum;
    const int64_t N = 3 * (xnum - 1) * (ynum - 1);
    for (int64_t k = 0; k < N; ++k) R[k] = 0.0;
    /* StokesBuildManager.jl:666-671: i outer, j inner; x, y, continuity per cell */
    for (int64_t i = 1; i <= ynum - 1; ++i) {
        for (int64_t j = 1; j <= xnum - 1; ++j) {
            int64_t cell = (j - 1) * (ynum - 1) + i;   /* GlobalIndices.jl:13-27 */
            int64_t ivx = cell * 3 - 2;
            x_stokes(&c, R, s, i, j, ivx);
            y_stokes(&c, R, s, i, j, ivx + 1);
            continuity(&c, R, s, i, j, ivx + 2);
        }
    }
    return c.inz - 1;
}

/* SystemSolver.jl:108-115 */
double ebo_pressure_scaling_coeff(double etan11, double xstpavg, double ystpavg)
{
    return 2.0 * etan11 / (xstpavg + ystpavg);
}

/* StokesViscoelasticTerms.jl:26-108 and StokesContinuityRhs.jl:30-190 */
void ebo_stokes_rhs(const ebo_rhs_in *r, double *etas0, double *sxy0, double *etan0,
                    double *sxx0, double *RX1, double *RY1, double *RC1)
{
    const int64_t xnum = r->xnum, ynum = r->ynum, ny = ynum;
    const double dt = r->timestep;
    for (int64_t j = 1; j <= xnum; ++j)
        for (int64_t i = 1; i <= ynum; ++i) {
            double eta = EBO_M(r->etas1, ny, i, j), mu = EBO_M(r->mus1, ny, i, j);
            double f = eta / (eta + dt * mu);
            EBO_M(etas0, ny, i, j) = eta * (1.0 - f);
            EBO_M(sxy0, ny, i, j) = EBO_M(r->sxy1, ny, i, j) * f;
        }
    for (int64_t j = 1; j <= xnum - 1; ++j)
        for (int64_t i = 1; i <= ynum - 1; ++i) {
            double eta = EBO_M(r->etan1, ny - 1, i, j), mu = EBO_M(r->mun1, ny - 1, i, j);
            double f = eta / (eta + dt * mu);
            EBO_M(etan0, ny - 1, i, j) = eta * (1.0 - f);
            EBO_M(sxx0, ny - 1, i, j) = EBO_M(r->sxx1, ny - 1, i, j) * f;
        }
    for (int64_t k = 0; k < (ynum + 1) * xnum; ++k) RX1[k] = 0.0;
    for (int64_t k = 0; k < ynum * (xnum + 1); ++k) RY1[k] = 0.0;
    for (int64_t k = 0; k < (ynum - 1) * (xnum - 1); ++k) RC1[k] = 0.0;
    /* viscoelastic_rhs_y_stokes! */
    for (int64_t j = 2; j <= xnum; ++j)
        for (int64_t i = 2; i <= ynum; ++i) {
            if (i == ynum) continue;
            double v;
            if (r->iuse_interface_stabilization == 1)
                v = -r->gravity_y * EBO_M(r->rho1_vy, ny, i, j);
            else
                v = -r->gravity_y * (EBO_M(r->rho1, ny, i, j) + EBO_M(r->rho1, ny, i, j - 1)) / 2.0;
            v += (EBO_M(sxx0, ny - 1, i, j - 1) - EBO_M(sxx0, ny - 1, i - 1, j - 1)) / EBO_V(r->ystp_vx, i);
            v -= (EBO_M(sxy0, ny, i, j) - EBO_M(sxy0, ny, i, j - 1)) / EBO_V(r->xstp_b, j - 1);
            EBO_M(RY1, ny, i, j) = v;
        }
    /* viscoelastic_rhs_x_stokes! */
    for (int64_t j = 2; j <= xnum; ++j)
        for (int64_t i = 2; i <= ynum; ++i) {
            if (j == xnum) continue;
            double v = -r->gravity_x * (EBO_M(r->rho1, ny, i, j) + EBO_M(r->rho1, ny, i - 1, j)) / 2.0;
            v -= (EBO_M(sxx0, ny - 1, i - 1, j) - EBO_M(sxx0, ny - 1, i - 1, j - 1)) / EBO_V(r->xstp_vy, j);
            v -= (EBO_M(sxy0, ny, i, j) - EBO_M(sxy0, ny, i - 1, j)) / EBO_V(r->ystp_b, i - 1);
            EBO_M(RX1, ny + 1, i, j) = v;
        }
    /* viscoelastic_rhs_continuity! */
    if (r->dilatation && r->eii_plastic_pressure)
        for (int64_t j = 1; j <= xnum - 1; ++j)
            for (int64_t i = 1; i <= ynum - 1; ++i) {
                double ang = EBO_M(r->dilatation, ny - 1, i, j) * M_PI / 180.0;
                EBO_M(RC1, ny - 1, i, j) = 2.0 * sin(ang) * EBO_M(r->eii_plastic_pressure, ny - 1, i, j);
            }
}

/* SystemSolver.jl:141-271: reload_solu_vel! then apply_vel_bc! */
void ebo_stokes_unpack(const ebo_stokes_in *s, const double *S, double *vx1, double *vy1, double *pr1)
{
    const int64_t xnum = s->xnum, ynum = s->ynum, ny = ynum, nx = xnum;
    for (int64_t k = 0; k < (ynum + 1) * xnum; ++k) vx1[k] = 0.0;
    for (int64_t k = 0; k < ynum * (xnum + 1); ++k) vy1[k] = 0.0;
    for (int64_t k = 0; k < (ynum - 1) * (xnum - 1); ++k) pr1[k] = 0.0;
    for (int64_t j = 1; j <= xnum - 1; ++j)
        for (int64_t i = 1; i <= ynum - 1; ++i) {
            int64_t cell = (j - 1) * (ynum - 1) + i, ivx = 3 * cell - 2;
            if (j < xnum - 1) EBO_M(vx1, ny + 1, i + 1, j + 1) = S[ivx - 1];
            if (i < ynum - 1) EBO_M(vy1, ny, i + 1, j + 1) = S[ivx];
            EBO_M(pr1, ny - 1, i, j) = S[ivx + 1] * s->pscale;
        }
    for (int64_t j = 1; j <= xnum; ++j) {
        EBO_M(vx1, ny + 1, 1, j) = BX(s->btopx, j, 1) + BX(s->btopx, j, 2) * EBO_M(vx1, ny + 1, 2, j);
        EBO_M(vx1, ny + 1, ynum + 1, j) = BX(s->bbottomx, j, 1) + BX(s->bbottomx, j, 2) * EBO_M(vx1, ny + 1, ynum, j);
    }
    for (int64_t j = 1; j <= xnum + 1; ++j) {
        EBO_M(vy1, ny, 1, j) = BX(s->btopy, j, 1) + BX(s->btopy, j, 2) * EBO_M(vy1, ny, 2, j);
        EBO_M(vy1, ny, ynum, j) = BX(s->bbottomy, j, 1) + BX(s->bbottomy, j, 2) * EBO_M(vy1, ny, ynum - 1, j);
    }
    for (int64_t i = 1; i <= ynum + 1; ++i) {
        EBO_M(vx1, ny + 1, i, 1) = BY(s->bleftx, i, 1) + BY(s->bleftx, i, 2) * EBO_M(vx1, ny + 1, i, 2);
        EBO_M(vx1, ny + 1, i, xnum) = BY(s->brightx, i, 1) + BY(s->brightx, i, 2) * EBO_M(vx1, ny + 1, i, xnum - 1);
    }
    for (int64_t i = 1; i <= ynum; ++i) {
        EBO_M(vy1, ny, i, 1) = BY(s->blefty, i, 1) + BY(s->blefty, i, 2) * EBO_M(vy1, ny, i, 2);
        EBO_M(vy1, ny, i, xnum + 1) = BY(s->brighty, i, 1) + BY(s->brighty, i, 2) * EBO_M(vy1, ny, i, xnum);
    }
    (void)nx;
}

/* StokesResiduals.jl:63-82 */
void ebo_coo_residual(int64_t N, int64_t nnz, const int64_t *Li, const int64_t *Lj,
                      const double *Lv, const double *x, const double *R, double *res)
{
    for (int64_t k = 0; k < N; ++k) res[k] = 0.0;
    for (int64_t k = 0; k < nnz; ++k) res[Li[k] - 1] += Lv[k] * x[Lj[k] - 1];
    for (int64_t k = 0; k < N; ++k) res[k] = fabs(res[k] - R[k]);
}

/* COO SpMV y = L*x in triplet order (the signed counterpart of the above). */
void ebo_coo_spmv(int64_t N, int64_t nnz, const int64_t *Li, const int64_t *Lj,
                  const double *Lv, const double *x, double *y)
{
    for (int64_t k = 0; k < N; ++k) y[k] = 0.0;
    for (int64_t k = 0; k < nnz; ++k) y[Li[k] - 1] += Lv[k] * x[Lj[k] - 1];
}

/* compute_stokes_residuals! (solver/residuals/StokesResiduals.jl:94-137) with
 * XStokes.calculate_residual! (residuals/utils/XStokes.jl:66-117),
 * YStokes.calculate_residual! (YStokes.jl:59-112) and
 * Continuity.calculate_residual! (Continuity.jl:53-84): per-equation grid
 * residuals of (vx1, vy1, pr1); nodes in an internal velocity zone keep 0. */
void ebo_stokes_grid_residuals(const ebo_stokes_in *s, const double *vx1, const double *vy1,
                               const double *pr1, double *resx, double *resy, double *resc)
{
    const int64_t xnum = s->xnum, ynum = s->ynum;
    const int64_t lx = ynum + 1, ly = ynum, lp = ynum - 1, le = ynum, ln = ynum - 1;
    const int64_t *z = s->bintern_zone;
    for (int64_t k = 0; k < lx * xnum; ++k) resx[k] = 0.0;
    for (int64_t k = 0; k < ly * (xnum + 1); ++k) resy[k] = 0.0;
    for (int64_t k = 0; k < lp * (xnum - 1); ++k) resc[k] = 0.0;
    for (int64_t j = 1; j <= xnum + 1; ++j)
        for (int64_t i = 1; i <= ynum + 1; ++i) {
            /* XStokes.jl:66-117 */
            if (j < xnum + 1 && (j != z[0] || i < z[1] || i > z[2])) {
                if (i == 1 || i == ynum + 1 || j == 1 || j == xnum) {
                    EBO_M(resx, lx, i, j) = 0.0;
                } else {
                    double r = EBO_M(s->RX, lx, i, j) -
                               (2.0 * (EBO_M(s->etan, ln, i - 1, j) * (EBO_M(vx1, lx, i, j + 1) - EBO_M(vx1, lx, i, j)) /
                                           EBO_V(s->xstp, j) -
                                       EBO_M(s->etan, ln, i - 1, j - 1) * (EBO_M(vx1, lx, i, j) - EBO_M(vx1, lx, i, j - 1)) /
                                           EBO_V(s->xstp, j - 1)) -
                                (EBO_M(pr1, lp, i - 1, j) - EBO_M(pr1, lp, i - 1, j - 1))) /
                                   EBO_V(s->xstpc, j);
                    r = r - (EBO_M(s->etas, le, i, j) * ((EBO_M(vx1, lx, i + 1, j) - EBO_M(vx1, lx, i, j)) / EBO_V(s->ystpc, i) +
                                                        (EBO_M(vy1, ly, i, j + 1) - EBO_M(vy1, ly, i, j)) / EBO_V(s->xstpc, j)) -
                             EBO_M(s->etas, le, i - 1, j) * ((EBO_M(vx1, lx, i, j) - EBO_M(vx1, lx, i - 1, j)) / EBO_V(s->ystpc, i - 1) +
                                                            (EBO_M(vy1, ly, i - 1, j + 1) - EBO_M(vy1, ly, i - 1, j)) / EBO_V(s->xstpc, j))) /
                                EBO_V(s->ystp, i - 1);
                    EBO_M(resx, lx, i, j) = r;
                }
            }
            /* YStokes.jl:59-112 */
            if (i < ynum + 1 && (j != z[4] || i < z[5] || i > z[6])) {
                if (i == 1 || i == ynum || j == 1 || j == xnum + 1) {
                    EBO_M(resy, ly, i, j) = 0.0;
                } else {
                    double r = EBO_M(s->RY, ly, i, j) -
                               (2.0 * (EBO_M(s->etan, ln, i, j - 1) * (EBO_M(vy1, ly, i + 1, j) - EBO_M(vy1, ly, i, j)) /
                                           EBO_V(s->ystp, i) -
                                       EBO_M(s->etan, ln, i - 1, j - 1) * (EBO_M(vy1, ly, i, j) - EBO_M(vy1, ly, i - 1, j)) /
                                           EBO_V(s->ystp, i - 1)) -
                                (EBO_M(pr1, lp, i, j - 1) - EBO_M(pr1, lp, i - 1, j - 1))) /
                                   EBO_V(s->ystpc, i);
                    r = r - (EBO_M(s->etas, le, i, j) * ((EBO_M(vy1, ly, i, j + 1) - EBO_M(vy1, ly, i, j)) / EBO_V(s->xstpc, j) +
                                                        (EBO_M(vx1, lx, i + 1, j) - EBO_M(vx1, lx, i, j)) / EBO_V(s->ystpc, i)) -
                             EBO_M(s->etas, le, i, j - 1) * ((EBO_M(vy1, ly, i, j) - EBO_M(vy1, ly, i, j - 1)) / EBO_V(s->xstpc, j - 1) +
                                                            (EBO_M(vx1, lx, i + 1, j - 1) - EBO_M(vx1, lx, i, j - 1)) / EBO_V(s->ystpc, i))) /
                                EBO_V(s->xstp, j - 1);
                    EBO_M(resy, ly, i, j) = r;
                }
            }
            /* Continuity.jl:53-84 */
            if (i < ynum && j < xnum)
                EBO_M(resc, lp, i, j) = EBO_M(s->RC, lp, i, j) -
                                        ((EBO_M(vx1, lx, i + 1, j + 1) - EBO_M(vx1, lx, i + 1, j)) / EBO_V(s->xstp, j) +
                                         (EBO_M(vy1, ly, i + 1, j + 1) - EBO_M(vy1, ly, i, j + 1)) / EBO_V(s->ystp, i));
        }
}
