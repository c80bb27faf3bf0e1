/*
 * Oracle (test infrastructure only): the reference's experimental 2-D
 * geometric multigrid (src/stokes_continuity_solver/multigrid/), restated.
 *
 * Sweep order.  The reference smoother (smoother/solver/solve/
 * SolveStokes2dOpt.jl:9-76) visits the cells of one colour with
 * Threads.@threads over columns and updates vx, vy and p of a cell in turn.  The
 * staggered stencil couples same-colour diagonal cells (vx(i,j) reads
 * vy(i-1,j+1) and p(i-1,j-1)), so its result depends on the thread schedule.
 * This restatement fixes a deterministic, data-parallel order: per colour,
 * all vx of that colour, then all vy, then all p.  The formulas of every update
 * are the reference's (Residuals.jl:145-591 2-D variants, SolveStokes2dOpt.jl
 * update_vx!/update_vy!/update_pressure!).
 */
#include "eb_mg_oracle.h"
#include <math.h>
#include <stdlib.h>

#define VX(l, i, j) EBO_M((l)->vx, (l)->ynum + 1, i, j)
#define VY(l, i, j) EBO_M((l)->vy, (l)->ynum, i, j)
#define PR(l, i, j) EBO_M((l)->pr, (l)->ynum - 1, i, j)
#define ETAS(l, i, j) EBO_M((l)->etas, (l)->ynum, i, j)
#define ETAN(l, i, j) EBO_M((l)->etan, (l)->ynum - 1, i, j)

/* multigrid/utils/Domain.jl:117-156 */
static int on_vx_boundary(int64_t i, int64_t j, int64_t ynum, int64_t xnum)
{
    return i == 1 || i == ynum + 1 || j == 1 || j == xnum;
}
static int on_vy_boundary(int64_t i, int64_t j, int64_t ynum, int64_t xnum)
{
    return i == 1 || i == ynum || j == 1 || j == xnum + 1;
}

/* Residuals.jl: calculate_vx_residual (2-D) */
double ebo_mg_vx_residual(const ebo_mglevel *l, int64_t i, int64_t j, double *coef)
{
    const double dxL = EBO_V(l->xstp_b, j - 1), dxR = EBO_V(l->xstp_b, j), dxC = EBO_V(l->xstp_vy, j);
    const double dyU = EBO_V(l->ystp_vx, i - 1), dyD = EBO_V(l->ystp_vx, i), dyC = EBO_V(l->ystp_b, i - 1);
    const double nR = ETAN(l, i - 1, j), nL = ETAN(l, i - 1, j - 1);
    const double sU = ETAS(l, i - 1, j), sD = ETAS(l, i, j);
    const double pL = PR(l, i - 1, j - 1), pR = PR(l, i - 1, j);
    const double vxL = VX(l, i, j - 1), vxC = VX(l, i, j), vxR = VX(l, i, j + 1);
    const double vxU = VX(l, i - 1, j), vxD = VX(l, i + 1, j);
    const double vyUL = VY(l, i - 1, j), vyUR = VY(l, i - 1, j + 1);
    const double vyLL = VY(l, i, j), vyLR = VY(l, i, j + 1);
    const double dp_dx = (pR - pL) / dxC;
    const double dsxx = 2.0 * (nR * (vxR - vxC) / dxR - nL * (vxC - vxL) / dxL) / dxC;
    const double dsxy = (sD * ((vxD - vxC) / dyD + (vyLR - vyLL) / dxC) -
                         sU * ((vxC - vxU) / dyU + (vyUR - vyUL) / dxC)) / dyC;
    const double dR = EBO_M(l->RX, l->ynum + 1, i, j) + dp_dx - dsxx - dsxy;
    if (coef) *coef = -2.0 * (nR / dxR / dxC + nL / dxL / dxC) - (sD / dyD / dyC + sU / dyU / dyC);
    return dR;
}

/* Residuals.jl: calculate_vy_residual (2-D) */
double ebo_mg_vy_residual(const ebo_mglevel *l, int64_t i, int64_t j, double *coef)
{
    const double dxL = EBO_V(l->xstp_b, j - 1), dxR = EBO_V(l->xstp_vy, j), dxC = EBO_V(l->xstp_vy, j - 1);
    const double dyU = EBO_V(l->ystp_b, i - 1), dyD = EBO_V(l->ystp_b, i), dyC = EBO_V(l->ystp_vx, i);
    const double pD = PR(l, i, j - 1), pU = PR(l, i - 1, j - 1);
    const double nD = ETAN(l, i, j - 1), nU = ETAN(l, i - 1, j - 1);
    const double sR = ETAS(l, i, j), sL = ETAS(l, i, j - 1);
    const double vyU = VY(l, i - 1, j), vyC = VY(l, i, j), vyD = VY(l, i + 1, j);
    const double vyR = VY(l, i, j + 1), vyL = VY(l, i, j - 1);
    const double vxLR = VX(l, i + 1, j), vxUR = VX(l, i, j);
    const double vxLL = VX(l, i + 1, j - 1), vxUL = VX(l, i, j - 1);
    const double dp_dy = (pD - pU) / dyC;
    const double dsyy = 2.0 * (nD * (vyD - vyC) / dyD - nU * (vyC - vyU) / dyU) / dyC;
    const double dsyx = (sR * ((vyR - vyC) / dxR + (vxLR - vxUR) / dyC) -
                         sL * ((vyC - vyL) / dxL + (vxLL - vxUL) / dyC)) / dxC;
    const double dR = EBO_M(l->RY, l->ynum, i, j) + dp_dy - dsyy - dsyx;
    if (coef) *coef = -2.0 * (nD / dyD / dyC + nU / dyU / dyC) - (sR / dxR / dxC + sL / dxL / dxC);
    return dR;
}

/* Residuals.jl: calculate_pressure_residual (2-D) */
double ebo_mg_pr_residual(const ebo_mglevel *l, int64_t i, int64_t j)
{
    const double dxC = EBO_V(l->xstp_b, j), dyC = EBO_V(l->ystp_b, i);
    return EBO_M(l->RC, l->ynum - 1, i, j) -
           ((VX(l, i + 1, j + 1) - VX(l, i + 1, j)) / dxC + (VY(l, i + 1, j + 1) - VY(l, i, j + 1)) / dyC);
}

/* boundary_conditions/BoundaryConditions.jl:22-53 with
 * core/VxBoundaryConditions.jl and VyBoundaryConditions.jl (bc_type
 * :free_slip).  Generalised by the tangential mirror coefficients
 * d_top/d_bottom (vx) and d_left/d_right (vy): +1 free slip, -1 no slip; the
 * normal components are zero.  Visiting order: columns j, rows i. */
void ebo_mg_set_bc(ebo_mglevel *l, double d_top, double d_bottom, double d_left, double d_right)
{
    const int64_t xnum = l->xnum, ynum = l->ynum;
    for (int64_t j = 1; j <= xnum + 1; ++j)
        for (int64_t i = 1; i <= ynum + 1; ++i) {
            if (j < xnum + 1 && on_vx_boundary(i, j, ynum, xnum)) {
                VX(l, i, j) = 0.0;
                if (i == 1) VX(l, i, j) = d_top * VX(l, i + 1, j);
                if (i == ynum + 1) VX(l, i, j) = d_bottom * VX(l, i - 1, j);
            }
            if (i < ynum + 1 && on_vy_boundary(i, j, ynum, xnum)) {
                VY(l, i, j) = 0.0;
                if (j == 1) VY(l, i, j) = d_left * VY(l, i, j + 1);
                if (j == xnum + 1) VY(l, i, j) = d_right * VY(l, i, j - 1);
            }
        }
}

/* One smoothing sweep (both colours, phased order described above). */
void ebo_mg_sweep(ebo_mglevel *l, double theta_s, double theta_c)
{
    const int64_t xnum = l->xnum, ynum = l->ynum;
    for (int color = 0; color <= 1; ++color) {
        for (int64_t j = 1; j <= xnum; ++j)
            for (int64_t i = 1; i <= ynum + 1; ++i) {
                if (((i + j) & 1) != color || on_vx_boundary(i, j, ynum, xnum)) continue;
                double coef, dR = ebo_mg_vx_residual(l, i, j, &coef);
                VX(l, i, j) += dR / coef * theta_s;
            }
        for (int64_t j = 1; j <= xnum + 1; ++j)
            for (int64_t i = 1; i <= ynum; ++i) {
                if (((i + j) & 1) != color || on_vy_boundary(i, j, ynum, xnum)) continue;
                double coef, dR = ebo_mg_vy_residual(l, i, j, &coef);
                VY(l, i, j) += dR / coef * theta_s;
            }
        for (int64_t j = 1; j <= xnum - 1; ++j)
            for (int64_t i = 1; i <= ynum - 1; ++i) {
                if (((i + j) & 1) != color) continue;
                double dR = ebo_mg_pr_residual(l, i, j);
                PR(l, i, j) += ETAN(l, i, j) * dR * theta_c;
            }
    }
}

/* pressure_update/PressureUpdate.jl:24-37 */
void ebo_mg_normalize_pressure(ebo_mglevel *l, double prnorm)
{
    const int64_t n = (l->ynum - 1) * (l->xnum - 1);
    const double dp = prnorm - l->pr[0];
    for (int64_t k = 0; k < n; ++k) l->pr[k] += dp;
}

/* Residuals.jl:101-143 (compute_residuals! 2-D): interior residuals, zeros on
 * boundary nodes. */
void ebo_mg_residuals(ebo_mglevel *l)
{
    const int64_t xnum = l->xnum, ynum = l->ynum;
    for (int64_t k = 0; k < (ynum + 1) * xnum; ++k) l->rx[k] = 0.0;
    for (int64_t k = 0; k < ynum * (xnum + 1); ++k) l->ry[k] = 0.0;
    for (int64_t k = 0; k < (ynum - 1) * (xnum - 1); ++k) l->rc[k] = 0.0;
    for (int64_t j = 1; j <= xnum; ++j)
        for (int64_t i = 1; i <= ynum; ++i) {
            if (j < xnum && !on_vx_boundary(i, j, ynum, xnum))
                EBO_M(l->rx, ynum + 1, i, j) = ebo_mg_vx_residual(l, i, j, NULL);
            if (!on_vy_boundary(i, j, ynum, xnum))
                EBO_M(l->ry, ynum, i, j) = ebo_mg_vy_residual(l, i, j, NULL);
            if (i < ynum && j < xnum)
                EBO_M(l->rc, ynum - 1, i, j) = ebo_mg_pr_residual(l, i, j);
        }
}

/* smoother/Smoother.jl:66-99 (stokes_continuity2d_viscous_smoother!) with
 * StokesSolverGaussSeidel.jl:33-54: nsweeps x [BC refresh; sweep], pressure
 * normalisation, residuals. */
void ebo_mg_smoother(ebo_mglevel *l, int64_t nsweeps, double theta_s, double theta_c,
                     double prnorm, const double *bc_d)
{
    for (int64_t s = 0; s < nsweeps; ++s) {
        ebo_mg_set_bc(l, bc_d[0], bc_d[1], bc_d[2], bc_d[3]);
        ebo_mg_sweep(l, theta_s, theta_c);
    }
    ebo_mg_normalize_pressure(l, prnorm);
    ebo_mg_residuals(l);
}

/* grid_mapping/GridMappingManager.jl:289-338, one separable axis: index of the
 * left node of each point of axis a on axis b and the normalised distance. */
void ebo_mg_axis_map(int64_t na, const double *ga, int64_t nb, const double *gb,
                     int32_t *idx, double *d)
{
    for (int64_t k = 1; k <= na; ++k) {
        double xa = EBO_V(ga, k);
        int32_t ib = ebo_index_of_left_node(xa, gb, nb);
        double x0 = EBO_V(gb, ib);
        double w = EBO_V(gb, ib + 1) - x0;
        EBO_V(idx, k) = ib;
        EBO_V(d, k) = (xa - x0) / w;
    }
}

/* BiLinearInterpolation.jl:21-51 add_to_numerator_and_denominator_2d! */
static void add_num_den(int64_t ldc, double *num, double *den, int32_t iy, int32_t jx,
                        double dy, double dx, double val)
{
    const double omx = 1.0 - dx, omy = 1.0 - dy;
    const double w00 = omx * omy, w01 = omx * dy, w10 = dx * omy, w11 = dx * dy;
    EBO_M(num, ldc, iy, jx) += w00 * val;
    EBO_M(den, ldc, iy, jx) += w00;
    EBO_M(num, ldc, iy + 1, jx) += w01 * val;
    EBO_M(den, ldc, iy + 1, jx) += w01;
    EBO_M(num, ldc, iy, jx + 1) += w10 * val;
    EBO_M(den, ldc, iy, jx + 1) += w10;
    EBO_M(num, ldc, iy + 1, jx + 1) += w11 * val;
    EBO_M(den, ldc, iy + 1, jx + 1) += w11;
}

/* restriction/Restriction.jl:377-505 (2-D).  m holds the fine->coarse axis
 * maps: vx (y on gridy_vx, x on gridx_b), vy (y on gridy_b, x on gridx_vy),
 * pr (pressure grids). wts are scratch arrays of the coarse shapes. */
void ebo_mg_restrict(const ebo_mglevel *f, ebo_mglevel *c, const ebo_mgmap *m,
                     double *wtx, double *wty, double *wtc)
{
    const int64_t xf = f->xnum, yf = f->ynum, xc = c->xnum, yc = c->ynum;
    for (int64_t k = 0; k < (yc + 1) * xc; ++k) { c->RX[k] = 0.0; wtx[k] = 0.0; }
    for (int64_t k = 0; k < yc * (xc + 1); ++k) { c->RY[k] = 0.0; wty[k] = 0.0; }
    for (int64_t k = 0; k < (yc - 1) * (xc - 1); ++k) { c->RC[k] = 0.0; wtc[k] = 0.0; }
    for (int64_t j = 2; j <= xf; ++j)
        for (int64_t i = 2; i <= yf; ++i) {
            if (j < xf)
                add_num_den(yc + 1, c->RX, wtx, EBO_V(m->vx_iy, i), EBO_V(m->vx_jx, j),
                            EBO_V(m->vx_dy, i), EBO_V(m->vx_dx, j), EBO_M(f->rx, yf + 1, i, j));
            if (i < yf)
                add_num_den(yc, c->RY, wty, EBO_V(m->vy_iy, i), EBO_V(m->vy_jx, j),
                            EBO_V(m->vy_dy, i), EBO_V(m->vy_dx, j), EBO_M(f->ry, yf, i, j));
            if (i < yf && j < xf) {
                double v = ETAN(f, i, j) * EBO_M(f->rc, yf - 1, i, j);  /* etan_resc */
                add_num_den(yc - 1, c->RC, wtc, EBO_V(m->pr_iy, i), EBO_V(m->pr_jx, j),
                            EBO_V(m->pr_dy, i), EBO_V(m->pr_dx, j), v);
            }
        }
    for (int64_t jc = 1; jc <= xc + 1; ++jc)
        for (int64_t ic = 1; ic <= yc + 1; ++ic) {
            if (jc < xc + 1) {
                double *r = &EBO_M(c->RX, yc + 1, ic, jc);
                double w = EBO_M(wtx, yc + 1, ic, jc);
                if (w != 0 && ic > 1 && ic < yc + 1 && jc > 1 && jc < xc) *r = *r / w;
                else *r = 0.0;
            }
            if (ic < yc + 1) {
                double *r = &EBO_M(c->RY, yc, ic, jc);
                double w = EBO_M(wty, yc, ic, jc);
                if (w != 0 && ic > 1 && ic < yc && jc > 1 && jc < xc + 1) *r = *r / w;
                else *r = 0.0;
            }
            if (ic < yc && jc < xc) {
                double *r = &EBO_M(c->RC, yc - 1, ic, jc);
                double w = EBO_M(wtc, yc - 1, ic, jc);
                if (w != 0) *r = *r / w / ETAN(c, ic, jc);
                else *r = 0.0;
            }
        }
}

/* BiLinearInterpolation.jl:53-78 coarse2fine_bilinear_interpolation! */
static double c2f(int64_t ldc, const double *a, int32_t iy, int32_t jx, double dy, double dx)
{
    const double omx = 1.0 - dx, omy = 1.0 - dy;
    const double w00 = omx * omy, w01 = omx * dy, w10 = dx * omy, w11 = dx * dy;
    double v = 0.0;
    v += w00 * EBO_M(a, ldc, iy, jx);
    v += w01 * EBO_M(a, ldc, iy + 1, jx);
    v += w10 * EBO_M(a, ldc, iy, jx + 1);
    v += w11 * EBO_M(a, ldc, iy + 1, jx + 1);
    return v;
}

/* prolongation/Prolongation.jl:150-185 and vcycle/core/SmoothAndProlongate.jl:
 * fine += relax * bilinear(coarse) on every fine node (ghosts included). */
void ebo_mg_prolong(const ebo_mglevel *c, ebo_mglevel *f, const ebo_mgmap *m,
                    double relax_v, double relax_p)
{
    const int64_t xf = f->xnum, yf = f->ynum, yc = c->ynum;
    for (int64_t j = 1; j <= xf + 1; ++j)
        for (int64_t i = 1; i <= yf + 1; ++i) {
            if (j < xf + 1) {
                double dv = c2f(yc + 1, c->vx, EBO_V(m->vx_iy, i), EBO_V(m->vx_jx, j),
                                EBO_V(m->vx_dy, i), EBO_V(m->vx_dx, j));
                VX(f, i, j) += dv * relax_v;
            }
            if (i < yf + 1) {
                double dv = c2f(yc, c->vy, EBO_V(m->vy_iy, i), EBO_V(m->vy_jx, j),
                                EBO_V(m->vy_dy, i), EBO_V(m->vy_dx, j));
                VY(f, i, j) += dv * relax_v;
            }
            if (i < yf && j < xf) {
                double dp = c2f(yc - 1, c->pr, EBO_V(m->pr_iy, i), EBO_V(m->pr_jx, j),
                                EBO_V(m->pr_dy, i), EBO_V(m->pr_dx, j));
                PR(f, i, j) += dp * relax_p;
            }
        }
}

/* interpolation/ViscosityRestriction.jl:68-107, 185-298 (2-D): fine nodes are
 * visited i outer, j inner; coarse = sum(w*eta)/sum(w) where sum(w) != 0. */
void ebo_mg_viscosity_restrict(const ebo_mglevel *f, ebo_mglevel *c, const ebo_mgmap *m,
                               double *wts, double *wtn)
{
    const int64_t xf = f->xnum, yf = f->ynum, xc = c->xnum, yc = c->ynum;
    for (int64_t k = 0; k < yc * xc; ++k) { c->etas[k] = 0.0; wts[k] = 0.0; }
    for (int64_t k = 0; k < (yc - 1) * (xc - 1); ++k) { c->etan[k] = 0.0; wtn[k] = 0.0; }
    for (int64_t i = 1; i <= yf; ++i)
        for (int64_t j = 1; j <= xf; ++j) {
            add_num_den(yc, c->etas, wts, EBO_V(m->b_iy, i), EBO_V(m->b_jx, j),
                        EBO_V(m->b_dy, i), EBO_V(m->b_dx, j), ETAS(f, i, j));
            if (i < yf && j < xf)
                add_num_den(yc - 1, c->etan, wtn, EBO_V(m->pr_iy, i), EBO_V(m->pr_jx, j),
                            EBO_V(m->pr_dy, i), EBO_V(m->pr_dx, j), ETAN(f, i, j));
        }
    for (int64_t jc = 1; jc <= xc; ++jc)
        for (int64_t ic = 1; ic <= yc; ++ic) {
            double w = EBO_M(wts, yc, ic, jc);
            if (w != 0) ETAS(c, ic, jc) = ETAS(c, ic, jc) / w;
            if (ic < yc && jc < xc) {
                double wn = EBO_M(wtn, yc - 1, ic, jc);
                if (wn != 0) ETAN(c, ic, jc) = ETAN(c, ic, jc) / wn;
            }
        }
}

/* viscosity_scaling/ViscosityScaling.jl:62-80 (2-D scale_viscosity_for_all_levels!):
 * eta = eta_min_cur * (eta_orig / eta_min_o)^coef, coef = log(max_cur/min_cur)/log(max_o/min_o). */
void ebo_mg_scale_viscosity(int64_t n, const double *eta_orig, double *eta, double eta_min_o,
                            double eta_max_o, double eta_min_cur, double eta_max_cur)
{
    const double coef = log(eta_max_cur / eta_min_cur) / log(eta_max_o / eta_min_o);
    const double inv_min_o = 1.0 / eta_min_o;
    for (int64_t k = 0; k < n; ++k) eta[k] = eta_min_cur * pow(eta_orig[k] * inv_min_o, coef);
}

/* utils/CalculateMeanResiduals.jl:130-172 (2-D calculate_scaled_and_mean_residuals!):
 * out[0..2] = log10(sqrt(mean((r / scale)^2))) of the interior x- and y-Stokes
 * residuals (scale stokesscale) and of the continuity residuals
 * (continscale), summed serially in the order j, i. */
void ebo_mg_mean_residuals(const ebo_mglevel *l, double stokesscale, double continscale, double *out)
{
    const int64_t xnum = l->xnum, ynum = l->ynum;
    double sx = 0.0, sy = 0.0, sc = 0.0;
    for (int64_t j = 1; j <= xnum; ++j)
        for (int64_t i = 1; i <= ynum; ++i) {
            if (i > 1 && j > 1 && j < xnum) {
                const double r = EBO_M(l->rx, ynum + 1, i, j) / stokesscale;
                sx += r * r;
            }
            if (i > 1 && j > 1 && i < ynum) {
                const double r = EBO_M(l->ry, ynum, i, j) / stokesscale;
                sy += r * r;
            }
            if (j < xnum && i < ynum) {
                const double r = EBO_M(l->rc, ynum - 1, i, j) / continscale;
                sc += r * r;
            }
        }
    out[0] = log10(sqrt(sx / (double)((ynum - 1) * (xnum - 2))));
    out[1] = log10(sqrt(sy / (double)((ynum - 2) * (xnum - 1))));
    out[2] = log10(sqrt(sc / (double)((ynum - 1) * (xnum - 1))));
}
