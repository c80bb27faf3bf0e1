// Viscoelastic terms and right-hand sides of the Stokes-continuity system.
//
// Replaces StokesViscoelasticTerms.calc_viscoelastic_terms!
// (solver/rhs/StokesViscoelasticTerms.jl:26-108) and
// StokesContinuityRhs.viscoelastic_rhs! (solver/rhs/StokesContinuityRhs.jl:30-190),
// which the reference calls back to back (StokesContinuitySolver.jl:189-190):
// the viscoelastic factor eta / (eta + dt mu) (Viscoelastic.jl:26-37) on the
// basic and pressure nodes, then gravity and old-stress terms on the vx and vy
// nodes and the plastic volumetric term on the pressure nodes.  Every
// expression keeps the reference's operand order, so the results are the
// serial loops' bits (the sine of the dilatation angle aside).
#include "ebg_common.hpp"

#include <cmath>

namespace ebg {

namespace {

struct VeArgs {
    int xnum, ynum;
    const double *xstp_b, *ystp_b, *xstp_vy, *ystp_vx;
    ebg_viscoelastic_in in;
    ebg_viscoelastic_out out;
};

// etas0, sxy0 on basic nodes and etan0, sxx0 on pressure nodes
__global__ void ve_terms_kernel(VeArgs a)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = (int64_t)a.xnum * a.ynum, np = (int64_t)(a.xnum - 1) * (a.ynum - 1);
    const double dt = a.in.timestep;
    if (t < nb) {
        const double eta = a.in.etas1[t];
        const double f = eta / (eta + dt * a.in.mus1[t]);
        a.out.etas0[t] = eta * (1.0 - f);
        a.out.sxy0[t] = a.in.sxy1[t] * f;
    } else if (t < nb + np) {
        const int64_t c = t - nb;
        const double eta = a.in.etan1[c];
        const double f = eta / (eta + dt * a.in.mun1[c]);
        a.out.etan0[c] = eta * (1.0 - f);
        a.out.sxx0[c] = a.in.sxx1[c] * f;
    }
}

// RX1 (ynum+1, xnum), RY1 (ynum, xnum+1) and RC1 (ynum-1, xnum-1); entries
// the reference loops skip are zero (ArrayUtils.setzeros! first)
__global__ void ve_rhs_kernel(VeArgs a)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    const int64_t nrx = (int64_t)(ny + 1) * nx, nry = (int64_t)ny * (nx + 1), nrc = (int64_t)(ny - 1) * (nx - 1);
    const double* rho1 = a.in.rho1;
    const double* sxx0 = a.out.sxx0;
    const double* sxy0 = a.out.sxy0;
#define B(arr, i, j) (arr)[(size_t)((j) - 1) * ny + ((i) - 1)]
#define P(arr, i, j) (arr)[(size_t)((j) - 1) * (ny - 1) + ((i) - 1)]
    if (t < nrx) {
        const int i = (int)(t % (ny + 1)) + 1, j = (int)(t / (ny + 1)) + 1;
        double v = 0.0;
        if (i >= 2 && i <= ny && j >= 2 && j <= nx - 1) {  // viscoelastic_rhs_x_stokes!
            v = -a.in.gravity_x * (B(rho1, i, j) + B(rho1, i - 1, j)) / 2;
            v -= (P(sxx0, i - 1, j) - P(sxx0, i - 1, j - 1)) / a.xstp_vy[j - 1];
            v -= (B(sxy0, i, j) - B(sxy0, i - 1, j)) / a.ystp_b[i - 2];
        }
        a.out.RX1[t] = v;
    } else if (t < nrx + nry) {
        const int64_t u = t - nrx;
        const int i = (int)(u % ny) + 1, j = (int)(u / ny) + 1;
        double v = 0.0;
        if (i >= 2 && i <= ny - 1 && j >= 2 && j <= nx) {  // viscoelastic_rhs_y_stokes!
            if (a.in.iuse_interface_stabilization == 1)
                v = -a.in.gravity_y * a.in.rho1_vy[u];
            else
                v = -a.in.gravity_y * (B(rho1, i, j) + B(rho1, i, j - 1)) / 2;
            v += (P(sxx0, i, j - 1) - P(sxx0, i - 1, j - 1)) / a.ystp_vx[i - 1];
            v -= (B(sxy0, i, j) - B(sxy0, i, j - 1)) / a.xstp_b[j - 2];
        }
        a.out.RY1[u] = v;
    } else if (t < nrx + nry + nrc) {
        const int64_t c = t - nrx - nry;
        double v = 0.0;
        if (a.in.dilatation_grid) {  // viscoelastic_rhs_continuity!, calc_plastic_volumetric_effect
            const double ang = a.in.dilatation_grid[c] * M_PI / 180.0;
            v = 2.0 * sin(ang) * a.in.eii_plastic_pressure[c];
        }
        a.out.RC1[c] = v;
    }
#undef B
#undef P
}

unsigned nblocks(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

}  // namespace ebg

using namespace ebg;

extern "C" int ebg_stokes_viscoelastic_rhs(ebg_ctx* ctx, const ebg_viscoelastic_in* in, ebg_viscoelastic_out* out)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_stokes_viscoelastic_rhs: no grid");
    EBG_CHECK(in && out, EBG_ERR_ARG, "ebg_stokes_viscoelastic_rhs: NULL arguments");
    EBG_CHECK(in->etas1 && in->mus1 && in->sxy1 && in->rho1 && in->etan1 && in->mun1 && in->sxx1, EBG_ERR_ARG,
              "ebg_stokes_viscoelastic_rhs: NULL input field");
    EBG_CHECK(out->etas0 && out->sxy0 && out->etan0 && out->sxx0 && out->RX1 && out->RY1 && out->RC1, EBG_ERR_ARG,
              "ebg_stokes_viscoelastic_rhs: NULL output field");
    EBG_CHECK(in->iuse_interface_stabilization != 1 || in->rho1_vy, EBG_ERR_ARG,
              "ebg_stokes_viscoelastic_rhs: interface stabilization needs rho1_vy");
    EBG_CHECK((in->dilatation_grid == nullptr) == (in->eii_plastic_pressure == nullptr), EBG_ERR_ARG,
              "ebg_stokes_viscoelastic_rhs: dilatation_grid and eii_plastic_pressure go together");
    const Grid2D& g = ctx->grid;
    VeArgs a{};
    a.xnum = (int)g.xnum;
    a.ynum = (int)g.ynum;
    a.xstp_b = g.xstp_b;
    a.ystp_b = g.ystp_b;
    a.xstp_vy = g.xstp_vy;
    a.ystp_vx = g.ystp_vx;
    a.in = *in;
    a.out = *out;
    const int64_t nb = g.xnum * g.ynum, np = (g.xnum - 1) * (g.ynum - 1);
    const int64_t nrhs = (g.ynum + 1) * g.xnum + g.ynum * (g.xnum + 1) + np;
    hipLaunchKernelGGL(ve_terms_kernel, dim3(nblocks(nb + np)), dim3(256), 0, ctx->stream, a);
    hipLaunchKernelGGL(ve_rhs_kernel, dim3(nblocks(nrhs)), dim3(256), 0, ctx->stream, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}
