// Grid strain rate, forecast viscoelastic stress and stress change after a
// Stokes solve.
//
// Replaces GridStrainRate.calculate_normal_strain_rate!, calculate_shear_strain_rate!
// and calculate_second_invariant! (stokes_continuity_solver/utils/strain_rate/
// GridStrainRate.jl:33-159), GridStress.forecast_viscoelastic_grid_stress! and
// calculate_deviatoric_grid_stress_change! (stokes_continuity_solver/utils/GridStress.jl).
// One pass writes the strain rates, forecast stresses and stress changes on
// basic nodes and pressure cells; a second pass forms the invariant, which
// needs the four shear strain rates around each cell.  Expressions keep the
// reference's operand order (x^2.0 is x*x), so results are the serial loops'
// bits.
#include "ebg_common.hpp"

#include <cmath>

namespace ebg {

namespace {

struct StrainArgs {
    int xnum, ynum;
    const double *xstp_b, *ystp_b, *xstp_vy, *ystp_vx;
    ebg_strain_stress_in in;
    ebg_strain_stress_out out;
};

__global__ void strain_stress_kernel(StrainArgs a)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    const int64_t nb = (int64_t)nx * ny, np = (int64_t)(nx - 1) * (ny - 1);
    const double dt = a.in.timestep;
#define VX(i, j) a.in.vx1[(size_t)((j) - 1) * (ny + 1) + ((i) - 1)]
#define VY(i, j) a.in.vy1[(size_t)((j) - 1) * ny + ((i) - 1)]
    if (t < nb) {  // basic node: shear strain rate, forecast shear stress, change
        const int i = (int)(t % ny) + 1, j = (int)(t / ny) + 1;
        const double exy = 0.5 * ((VX(i + 1, j) - VX(i, j)) / a.ystp_vx[i - 1] + (VY(i, j + 1) - VY(i, j)) / a.xstp_vy[j - 1]);
        a.out.exy[t] = exy;
        const double eta = a.in.etas1[t];
        const double f = eta / (eta + dt * a.in.mus1[t]);
        const double s2 = (1.0 - f) * 2.0 * eta * exy + f * a.in.sxy1[t];
        a.out.sxy2[t] = s2;
        a.out.dsxy[t] = s2 - a.in.sxy1[t];
    } else if (t < nb + np) {  // pressure cell: normal strain rate, forecast normal stress, change
        const int64_t c = t - nb;
        const int i = (int)(c % (ny - 1)) + 1, j = (int)(c / (ny - 1)) + 1;
        const double exx = 0.5 * ((VX(i + 1, j + 1) - VX(i + 1, j)) / a.xstp_b[j - 1] -
                                  (VY(i + 1, j + 1) - VY(i, j + 1)) / a.ystp_b[i - 1]);
        a.out.exx[c] = exx;
        const double eta = a.in.etan1[c];
        const double f = eta / (eta + dt * a.in.mun1[c]);
        const double s2 = (1.0 - f) * 2.0 * eta * exx + f * a.in.sxx1[c];
        a.out.sxx2[c] = s2;
        a.out.dsxx[c] = s2 - a.in.sxx1[c];
    }
#undef VX
#undef VY
}

// eii on pressure cells from exx and the four surrounding exy
__global__ void strain_invariant_kernel(StrainArgs a)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    if (c >= (int64_t)(nx - 1) * (ny - 1)) return;
    const int i = (int)(c % (ny - 1)) + 2, j = (int)(c / (ny - 1)) + 2;  // basic node (i, j), cell (i-1, j-1)
#define EXY(ii, jj) a.out.exy[(size_t)((jj) - 1) * ny + ((ii) - 1)]
    const double e = a.out.exx[c];
    const double b = EXY(i - 1, j - 1), cc = EXY(i, j - 1), d = EXY(i - 1, j), f = EXY(i, j);
    a.out.eii[c] = sqrt(e * e + (b * b + cc * cc + d * d + f * f) / 4.0);
#undef EXY
}

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

}  // namespace ebg

using namespace ebg;

extern "C" int ebg_grid_strain_rate_and_stress(ebg_ctx* ctx, const ebg_strain_stress_in* in,
                                               ebg_strain_stress_out* out)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_grid_strain_rate_and_stress: no grid");
    EBG_CHECK(in && out, EBG_ERR_ARG, "ebg_grid_strain_rate_and_stress: NULL arguments");
    EBG_CHECK(in->vx1 && in->vy1 && in->etas1 && in->mus1 && in->sxy1 && in->etan1 && in->mun1 && in->sxx1,
              EBG_ERR_ARG, "ebg_grid_strain_rate_and_stress: NULL input field");
    EBG_CHECK(out->exx && out->exy && out->eii && out->sxy2 && out->sxx2 && out->dsxy && out->dsxx, EBG_ERR_ARG,
              "ebg_grid_strain_rate_and_stress: NULL output field");
    const Grid2D& g = ctx->grid;
    StrainArgs a{};
    a.xnum = (int)g.xnum;
    a.ynum = (int)g.ynum;
    a.xstp_b = g.xstp_b;
    a.ystp_b = g.ystp_b;
    a.xstp_vy = g.xstp_vy;
    a.ystp_vx = g.ystp_vx;
    a.in = *in;
    a.out = *out;
    const int64_t nb = g.xnum * g.ynum, np = (g.xnum - 1) * (g.ynum - 1);
    hipLaunchKernelGGL(strain_stress_kernel, dim3(nblk(nb + np)), dim3(256), 0, ctx->stream, a);
    hipLaunchKernelGGL(strain_invariant_kernel, dim3(nblk(np)), dim3(256), 0, ctx->stream, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}
