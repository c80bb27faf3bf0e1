/* TEST INFRASTRUCTURE ONLY: grid strain rate, forecast stress and stress
 * change, restated from stokes_continuity_solver/utils/strain_rate/
 * GridStrainRate.jl:33-159 and stokes_continuity_solver/utils/GridStress.jl. */
#include <math.h>

#include "eb_stokes_oracle.h"

void ebo_grid_strain_rate_and_stress(const ebo_strain_in *s, double *exx, double *exy, double *eii,
                                     double *sxy2, double *sxx2, double *dsxy, double *dsxx)
{
    const int64_t xnum = s->xnum, ynum = s->ynum, ny = ynum;
    const double dt = s->timestep;
    for (int64_t k = 0; k < (ynum - 1) * (xnum - 1); ++k) exx[k] = 0.0;
    for (int64_t j = 1; j <= xnum; ++j)   /* calculate_normal_strain_rate! */
        for (int64_t i = 1; i <= ynum; ++i)
            if (i < ynum && j < xnum)
                EBO_M(exx, ny - 1, i, j) = 0.5 * ((EBO_M(s->vx1, ny + 1, i + 1, j + 1) - EBO_M(s->vx1, ny + 1, i + 1, j)) /
                                                      EBO_V(s->xstp_b, j) -
                                                  (EBO_M(s->vy1, ny, i + 1, j + 1) - EBO_M(s->vy1, ny, i, j + 1)) /
                                                      EBO_V(s->ystp_b, i));
    for (int64_t j = 1; j <= xnum; ++j)   /* calculate_shear_strain_rate! */
        for (int64_t i = 1; i <= ynum; ++i)
            EBO_M(exy, ny, i, j) = 0.5 * ((EBO_M(s->vx1, ny + 1, i + 1, j) - EBO_M(s->vx1, ny + 1, i, j)) / EBO_V(s->ystp_vx, i) +
                                          (EBO_M(s->vy1, ny, i, j + 1) - EBO_M(s->vy1, ny, i, j)) / EBO_V(s->xstp_vy, j));
    for (int64_t k = 0; k < (ynum - 1) * (xnum - 1); ++k) eii[k] = 0.0;
    for (int64_t j = 1; j <= xnum; ++j)   /* calculate_second_invariant! */
        for (int64_t i = 1; i <= ynum; ++i)
            if (i > 1 && j > 1) {
                const double a = EBO_M(exx, ny - 1, i - 1, j - 1);
                const double b = EBO_M(exy, ny, i - 1, j - 1), c = EBO_M(exy, ny, i, j - 1);
                const double d = EBO_M(exy, ny, i - 1, j), e = EBO_M(exy, ny, i, j);
                EBO_M(eii, ny - 1, i - 1, j - 1) = sqrt(a * a + (b * b + c * c + d * d + e * e) / 4.0);
            }
    for (int64_t j = 1; j <= xnum; ++j)   /* forecast_viscoelastic_shear_stress! */
        for (int64_t i = 1; i <= ynum; ++i) {
            const double eta = EBO_M(s->etas1, ny, i, j);
            const double f = eta / (eta + dt * EBO_M(s->mus1, ny, i, j));
            EBO_M(sxy2, ny, i, j) = (1.0 - f) * 2.0 * eta * EBO_M(exy, ny, i, j) + f * EBO_M(s->sxy1, ny, i, j);
        }
    for (int64_t j = 1; j <= xnum - 1; ++j)   /* forecast_viscoelastic_normal_stress! */
        for (int64_t i = 1; i <= ynum - 1; ++i) {
            const double eta = EBO_M(s->etan1, ny - 1, i, j);
            const double f = eta / (eta + dt * EBO_M(s->mun1, ny - 1, i, j));
            EBO_M(sxx2, ny - 1, i, j) = (1.0 - f) * 2.0 * eta * EBO_M(exx, ny - 1, i, j) + f * EBO_M(s->sxx1, ny - 1, i, j);
        }
    for (int64_t k = 0; k < ynum * xnum; ++k) dsxy[k] = sxy2[k] - s->sxy1[k];   /* calculate_deviatoric_grid_stress_change! */
    for (int64_t k = 0; k < (ynum - 1) * (xnum - 1); ++k) dsxx[k] = sxx2[k] - s->sxx1[k];
}
