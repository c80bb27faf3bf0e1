// C ABI of the reference's experimental 2-D multigrid solver
// (stokes_continuity_solver/multigrid/, driver MultigridSolver.run_multigrid_solver
// at multigrid_solver/MultigridSolver.jl:113-135), FP64 on the device.
#include "ebg_common.hpp"
#include "ebg_mg.hpp"

using namespace ebg;

namespace {
struct MGRef {
    MGHierarchy<double> H;
};
MGRef* get(ebg_ctx* ctx) { return ctx ? (MGRef*)ctx->mg_ref : nullptr; }

double* field_ptr(MGLevelDev<double>& l, int which, size_t* n)
{
    const size_t ny = l.ynum, nx = l.xnum;
    switch (which) {
        case EBG_MG_ETAS: *n = ny * nx; return l.etas;
        case EBG_MG_ETAN: *n = (ny - 1) * (nx - 1); return l.etan;
        case EBG_MG_VX: *n = (ny + 1) * nx; return l.vx;
        case EBG_MG_VY: *n = ny * (nx + 1); return l.vy;
        case EBG_MG_PR: *n = (ny - 1) * (nx - 1); return l.pr;
        case EBG_MG_RX: *n = (ny + 1) * nx; return l.RX;
        case EBG_MG_RY: *n = ny * (nx + 1); return l.RY;
        case EBG_MG_RC: *n = (ny - 1) * (nx - 1); return l.RC;
        case EBG_MG_RESX: *n = (ny + 1) * nx; return l.rx;
        case EBG_MG_RESY: *n = ny * (nx + 1); return l.ry;
        case EBG_MG_RESC: *n = (ny - 1) * (nx - 1); return l.rc;
        default: *n = 0; return nullptr;
    }
}
}  // namespace

extern "C" {

int ebg_mgref_release(ebg_ctx* ctx)
{
    if (ctx && ctx->mg_ref) {
        delete (MGRef*)ctx->mg_ref;
        ctx->mg_ref = nullptr;
    }
    return EBG_OK;
}

int ebg_mg2d_create(ebg_ctx* ctx, int64_t xnum, int64_t ynum, double xsize, double ysize,
                    int nlevels, int* levels_out)
{
    EBG_CHECK(ctx, EBG_ERR_ARG, "ebg_mg2d_create: ctx is NULL");
    EBG_CHECK(xnum >= 4 && ynum >= 4 && xnum < (1 << 20) && ynum < (1 << 20), EBG_ERR_ARG,
              "ebg_mg2d_create: need 4 <= xnum, ynum < 2^20");
    EBG_CHECK(xsize > 0 && ysize > 0, EBG_ERR_ARG, "ebg_mg2d_create: sizes must be > 0");
    EBG_HIP(hipSetDevice(ctx->device));
    EBG_HIP(hipStreamSynchronize(ctx->stream));
    ebg_mgref_release(ctx);
    MGRef* m = new MGRef();
    int rc = mg_build(m->H, (int)xnum, (int)ynum, xsize, ysize, nlevels, nullptr);
    if (rc) {
        delete m;
        return rc;
    }
    ctx->mg_ref = m;
    if (levels_out) *levels_out = m->H.nlev();
    return EBG_OK;
}

int ebg_mg2d_level_shape(ebg_ctx* ctx, int level, int64_t* xnum, int64_t* ynum)
{
    MGRef* m = get(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg2d_level_shape: no hierarchy");
    EBG_CHECK(level >= 0 && level < m->H.nlev(), EBG_ERR_ARG, "ebg_mg2d_level_shape: bad level");
    *xnum = m->H.lev[level].xnum;
    *ynum = m->H.lev[level].ynum;
    return EBG_OK;
}

int ebg_mg2d_copy_field(ebg_ctx* ctx, int level, int which, double* dst, const double* src)
{
    MGRef* m = get(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg2d_copy_field: no hierarchy");
    EBG_CHECK(level >= 0 && level < m->H.nlev(), EBG_ERR_ARG, "ebg_mg2d_copy_field: bad level");
    EBG_CHECK((dst == nullptr) != (src == nullptr), EBG_ERR_ARG,
              "ebg_mg2d_copy_field: give exactly one of dst (read) or src (write)");
    size_t n = 0;
    double* f = field_ptr(m->H.lev[level], which, &n);
    EBG_CHECK(f, EBG_ERR_ARG, "ebg_mg2d_copy_field: unknown field id");
    if (dst) EBG_HIP(hipMemcpyAsync(dst, f, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    else EBG_HIP(hipMemcpyAsync(f, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return EBG_OK;
}

int ebg_mg2d_restrict_viscosity(ebg_ctx* ctx)
{
    MGRef* m = get(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg2d_restrict_viscosity: no hierarchy");
    mg_restrict_viscosity(m->H, ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_mg2d_smoother(ebg_ctx* ctx, int level, int nsweeps, double theta_s, double theta_c,
                      double prnorm, const double* bc_d4)
{
    MGRef* m = get(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg2d_smoother: no hierarchy");
    EBG_CHECK(level >= 0 && level < m->H.nlev() && nsweeps >= 0, EBG_ERR_ARG,
              "ebg_mg2d_smoother: bad level or sweep count");
    const double one[4] = {1.0, 1.0, 1.0, 1.0};
    mg_smoother(m->H, level, nsweeps, theta_s, theta_c, prnorm, (const double*)nullptr,
                bc_d4 ? bc_d4 : one, ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_mg2d_smoother_tiled(ebg_ctx* ctx, int level, int nsweeps, int local_sweeps, int tile,
                            double theta_s, double theta_c, double prnorm, const double* bc_d4)
{
    MGRef* m = get(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg2d_smoother_tiled: no hierarchy");
    EBG_CHECK(level >= 0 && level < m->H.nlev() && nsweeps >= 0 && local_sweeps >= 1, EBG_ERR_ARG,
              "ebg_mg2d_smoother_tiled: bad level or sweep count");
    EBG_CHECK(tile == 0 || (tile >= 4 && tile <= 32), EBG_ERR_ARG,
              "ebg_mg2d_smoother_tiled: tile must be 0 (whole level) or in [4, 32]");
    int rc = mg_plan_tiles(m->H, tile == 0 ? 16 : tile, tile == 0, false);
    if (rc) return rc;
    EBG_CHECK(tile != 0 || m->H.tiles[level].single, EBG_ERR_ARG,
              "ebg_mg2d_smoother_tiled: level does not fit one LDS tile");
    const double one[4] = {1.0, 1.0, 1.0, 1.0};
    mg_smoother_tiled(m->H, level, nsweeps, local_sweeps, theta_s, theta_c, prnorm,
                      (const double*)nullptr, bc_d4 ? bc_d4 : one, ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_mg2d_vcycles(ebg_ctx* ctx, int ncycles, int finest_sweeps, int multiplier, int cap,
                     double theta_s, double theta_c, double relax_v, double relax_p,
                     double prnorm, const double* bc_d4)
{
    MGRef* m = get(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg2d_vcycles: no hierarchy");
    EBG_CHECK(ncycles >= 0 && finest_sweeps >= 0 && multiplier >= 1, EBG_ERR_ARG,
              "ebg_mg2d_vcycles: bad cycle parameters");
    const double one[4] = {1.0, 1.0, 1.0, 1.0};
    auto sm = mg_smoothing_counts(m->H.nlev(), finest_sweeps, multiplier, cap);
    for (int c = 0; c < ncycles; ++c)
        mg_vcycle(m->H, sm, theta_s, theta_c, relax_v, relax_p, prnorm, (const double*)nullptr,
                  bc_d4 ? bc_d4 : one, ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

}  // extern "C"
