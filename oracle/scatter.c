/*
 * Oracle (test infrastructure only): marker -> grid interpolation of the Stokes
 * transport properties, restated from the reference:
 *   indices   interpolation/mapping/MarkerGridMapping.jl:99-173
 *   distances MarkerGridMapping.jl:185-265
 *   weights   interpolation/initialize/utils/MarkerGridNodeWeights.jl:68-150
 *   sums      initialize/utils/BilinearDenominator.jl:34-182 (denominators),
 *             src/interpolation/utils/BilinearNumerator.jl:205-454 (numerators),
 *             rules interpolation/weights/WeightFuncs.jl:67-190
 *   averages  src/interpolation/utils/BilinearAverage.jl:57-178,276-322 with the
 *             backup fallback of backup/BackupTransportArrays.jl:33-104
 * Each sum runs serially over markers in index order, as the reference does.
 */
#include "eb_scatter_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void add_incl(double *a, int64_t ld, int32_t iy, int32_t ix, double ul, double ll,
                     double ur, double lr, double m)
{
    EBO_M(a, ld, iy, ix) += ul * m * 1.0;
    EBO_M(a, ld, iy + 1, ix) += ll * m * 1.0;
    EBO_M(a, ld, iy, ix + 1) += ur * m * 1.0;
    EBO_M(a, ld, iy + 1, ix + 1) += lr * m * 1.0;
}

static void add_excl(double *a, int64_t ld, int32_t iy, int32_t ix, double ul, double ll,
                     double ur, double lr, double dx, double dy, double m)
{
    if (dx <= 0.5 && dy <= 0.5) EBO_M(a, ld, iy, ix) += ul * m * 1.0;
    if (dx <= 0.5 && 0.5 <= dy) EBO_M(a, ld, iy + 1, ix) += ll * m * 1.0;
    if (dx >= 0.5 && 0.5 >= dy) EBO_M(a, ld, iy, ix + 1) += ur * m * 1.0;
    if (dx >= 0.5 && dy >= 0.5) EBO_M(a, ld, iy + 1, ix + 1) += lr * m * 1.0;
}

static void add_central(double *a, int64_t ld, int32_t iy, int32_t ix, double wc, double m)
{
    EBO_M(a, ld, iy, ix) += wc * m * 1.0;
}

/* BilinearAverage.jl:276-322; old = value before clearing (the suffix-0 backup) */
static void avg_arith(double *a, const double *old, const double *w, int64_t n)
{
    for (int64_t k = 0; k < n; ++k) a[k] = (w[k] != 0) ? a[k] / w[k] : old[k];
}
static void avg_harm(double *a, const double *old, const double *w, int64_t n)
{
    for (int64_t k = 0; k < n; ++k) a[k] = (w[k] != 0) ? 1.0 / (a[k] / w[k]) : old[k];
}

int ebo_scatter_stokes(const ebo_scatter_in *in, ebo_scatter_out *o)
{
    const int64_t M = in->marknum, nx = in->xnum, ny = in->ynum;
    const int64_t nb = nx * ny, np = (nx - 1) * (ny - 1), nv = ny * (nx + 1);
    /* per-marker mapping (MarkerGridMapping.jl, MarkerGridNodeWeights.jl) */
    int32_t *xn = o->xn, *yn = o->yn;
    int32_t *xnv = malloc(sizeof(int32_t) * (M ? M : 1));
    double *dx = malloc(sizeof(double) * (M ? M : 1)), *dy = malloc(sizeof(double) * (M ? M : 1));
    double *dxv = malloc(sizeof(double) * (M ? M : 1));
    for (int64_t m = 0; m < M; ++m) {
        if (in->flags[m] != 1) continue;
        xn[m] = ebo_index_of_left_node(in->x[m], in->gridx_b, nx);
        yn[m] = ebo_index_of_left_node(in->y[m], in->gridy_b, ny);
        xnv[m] = ebo_index_of_left_node(in->x[m], in->gridx_vy, nx + 1);
        dx[m] = (in->x[m] - EBO_V(in->gridx_b, xn[m])) / EBO_V(in->xstp_b, xn[m]);
        dy[m] = (in->y[m] - EBO_V(in->gridy_b, yn[m])) / EBO_V(in->ystp_b, yn[m]);
        dxv[m] = (in->x[m] - EBO_V(in->gridx_vy, xnv[m])) / EBO_V(in->xstp_vy, xnv[m]);
    }
    /* backups (BackupTransportArrays.jl:33-104), then clear */
    double *outs[15] = {o->rho1, o->rho1_vy, o->etas1, o->etan1, o->eta_flow, o->mus1, o->mun1,
                        o->sxy1, o->sxx1, o->plastics, o->plasticn, o->cohesion, o->fric_deg,
                        o->dilatation, o->meltfrac};
    const int64_t sizes[15] = {nb, nv, nb, np, nb, nb, np, nb, np, nb, np, nb, nb, np, np};
    double *bak[15];
    for (int f = 0; f < 15; ++f) {
        bak[f] = malloc(sizeof(double) * sizes[f]);
        memcpy(bak[f], outs[f], sizeof(double) * sizes[f]);
        memset(outs[f], 0, sizeof(double) * sizes[f]);
    }
    memset(o->wtnodes, 0, sizeof(double) * nb);
    memset(o->wtetas, 0, sizeof(double) * nb);
    memset(o->wtetan, 0, sizeof(double) * np);
    memset(o->wtnodes_vy, 0, sizeof(double) * nv);
    int err = 0;
    for (int64_t m = 0; m < M; ++m) {
        if (in->flags[m] != 1) continue;
        const double ddx = dx[m], ddy = dy[m];
        const double ul = (1.0 - ddx) * (1.0 - ddy), ll = (1.0 - ddx) * ddy;
        const double ur = ddx * (1.0 - ddy), lr = ddx * ddy;
        const double wc = (1.0 - fabs(0.5 - ddx)) * (1.0 - fabs(0.5 - ddy));
        const int32_t ix = xn[m], iy = yn[m];
        /* denominators: BilinearDenominator.jl:34-114 */
        add_incl(o->wtnodes, ny, iy, ix, ul, ll, ur, lr, 1.0);
        add_excl(o->wtetas, ny, iy, ix, ul, ll, ur, lr, ddx, ddy, 1.0);
        add_central(o->wtetan, ny - 1, iy, ix, wc, 1.0);
        /* numerators: BilinearNumerator.jl:259-388 */
        const int16_t mt = in->matid[m];
        const double mshear = EBO_V(in->mat_mu, mt);
        if (mshear == 0.0) err = 1;   /* check_shear_modulus (:155-160) */
        const double mmu_inv = 1 / mshear;
        const double mfric = asin(in->fric[m]) * 180.0 / M_PI;
        const double plastyn = (double)in->pfailure[m];
        const double mdil = (double)in->dilatation[m];
        add_incl(o->rho1, ny, iy, ix, ul, ll, ur, lr, in->rho[m]);
        add_excl(o->cohesion, ny, iy, ix, ul, ll, ur, lr, ddx, ddy, in->cohesion[m]);
        add_excl(o->fric_deg, ny, iy, ix, ul, ll, ur, lr, ddx, ddy, mfric);
        add_excl(o->eta_flow, ny, iy, ix, ul, ll, ur, lr, ddx, ddy, in->eta_flow[m]);
        add_excl(o->etas1, ny, iy, ix, ul, ll, ur, lr, ddx, ddy, in->eta[m]);
        add_central(o->etan1, ny - 1, iy, ix, wc, in->eta[m]);
        add_excl(o->mus1, ny, iy, ix, ul, ll, ur, lr, ddx, ddy, mmu_inv);
        add_central(o->mun1, ny - 1, iy, ix, wc, mmu_inv);
        add_excl(o->sxy1, ny, iy, ix, ul, ll, ur, lr, ddx, ddy, in->sxy[m]);
        add_excl(o->plastics, ny, iy, ix, ul, ll, ur, lr, ddx, ddy, plastyn);
        add_central(o->plasticn, ny - 1, iy, ix, wc, plastyn);
        add_central(o->dilatation, ny - 1, iy, ix, wc, mdil);
        add_central(o->meltfrac, ny - 1, iy, ix, wc, in->meltfrac[m]);
        add_central(o->sxx1, ny - 1, iy, ix, wc, in->sxx[m]);
    }
    /* vy grid: BilinearDenominator.jl:116-182 and BilinearNumerator.jl:408-454 */
    for (int64_t m = 0; m < M; ++m) {
        if (in->flags[m] != 1) continue;
        const double ddx = dxv[m], ddy = dy[m];
        const double ul = (1.0 - ddx) * (1.0 - ddy), ll = (1.0 - ddx) * ddy;
        const double ur = ddx * (1.0 - ddy), lr = ddx * ddy;
        add_incl(o->wtnodes_vy, ny, yn[m], xnv[m], ul, ll, ur, lr, 1.0);
        add_incl(o->rho1_vy, ny, yn[m], xnv[m], ul, ll, ur, lr, in->rho[m]);
    }
    /* averages: BilinearAverage.jl:57-178 */
    avg_arith(o->rho1, bak[0], o->wtnodes, nb);
    avg_arith(o->rho1_vy, bak[1], o->wtnodes_vy, nv);
    avg_harm(o->mus1, bak[5], o->wtetas, nb);
    avg_arith(o->etas1, bak[2], o->wtetas, nb);
    avg_arith(o->sxy1, bak[7], o->wtetas, nb);
    avg_arith(o->plastics, bak[9], o->wtetas, nb);
    avg_arith(o->cohesion, bak[11], o->wtetas, nb);
    avg_arith(o->fric_deg, bak[12], o->wtetas, nb);
    avg_arith(o->eta_flow, bak[4], o->wtetas, nb);
    avg_harm(o->mun1, bak[6], o->wtetan, np);
    avg_arith(o->etan1, bak[3], o->wtetan, np);
    avg_arith(o->sxx1, bak[8], o->wtetan, np);
    avg_arith(o->plasticn, bak[10], o->wtetan, np);
    avg_arith(o->dilatation, bak[13], o->wtetan, np);
    avg_arith(o->meltfrac, bak[14], o->wtetan, np);
    for (int f = 0; f < 15; ++f) free(bak[f]);
    free(xnv);
    free(dx);
    free(dy);
    free(dxv);
    return err;
}
