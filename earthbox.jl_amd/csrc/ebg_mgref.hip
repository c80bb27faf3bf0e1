// C ABI of the reference's experimental 2-D multigrid solver
// (stokes_continuity_solver/multigrid/, driver MultigridSolver.run_multigrid_solver
// at multigrid_solver/MultigridSolver.jl:113-135), FP64 on the device.
#include "ebg_common.hpp"
#include "ebg_krylov.hpp"
#include "ebg_mg.hpp"

#include <rocblas/rocblas.h>

#include <array>
#include <cmath>
#include <vector>

using namespace ebg;

namespace {
struct MGRef {
    MGHierarchy<double> H;
    // original (unscaled) etas and etan per level for the viscosity
    // continuation and the reduction partials (3 * MG_RUN_BLOCKS); allocated
    // by the first ebg_mg2d_run
    std::vector<std::array<double*, 2>> orig;
    double* partial = nullptr;
    // the global level 0 of the multimulti driver: vx, vy, pr, RX, RY, RC,
    // rx, ry, rc on the finest grid (allocated by the first multimulti run)
    std::array<double*, 9> lev0{};
    ~MGRef()
    {
        for (auto& a : orig)
            for (double* q : a) (void)hipFree(q);
        if (partial) (void)hipFree(partial);
        for (double* q : lev0)
            if (q) (void)hipFree(q);
    }
};

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

__global__ void mm_add_kernel(double* __restrict__ a, const double* __restrict__ b, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] += b[t];
}

// per-block sums of the squared scaled interior residuals of the finest level
// (CalculateMeanResiduals.jl:130-172): part[3*b + {x, y, c}]
__global__ void __launch_bounds__(256) mean_residual2d_partial_kernel(MGLevelDev<double> l, double ss, double cs,
                                                                      double* __restrict__ part)
{
    __shared__ double sh[3][256];
    const int X = l.xnum, Y = l.ynum;
    const int64_t total = (int64_t)Y * X;
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(t % Y), j = 1 + (int)(t / Y);
        if (i > 1 && j > 1 && j < X) {
            const double v = l.rx[(size_t)(j - 1) * (Y + 1) + (i - 1)] / ss;
            s[0] += v * v;
        }
        if (i > 1 && j > 1 && i < Y) {
            const double v = l.ry[(size_t)(j - 1) * Y + (i - 1)] / ss;
            s[1] += v * v;
        }
        if (j < X && i < Y) {
            const double v = l.rc[(size_t)(j - 1) * (Y - 1) + (i - 1)] / cs;
            s[2] += v * v;
        }
    }
    for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}
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
    EBG_CHECK(tile != 0 || m->H.tiles[level].whole_ok, EBG_ERR_ARG,
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

// Exact solve of one level's equations for its RX, RY, RC (the dense direct
// solve the Stokes preconditioner uses on its coarsest level): assembles the
// level's operator, inverts it, and writes vx, vy, pr with p(1,1) = 0 and
// the ghost pass applied.
int ebg_mg2d_direct_solve(ebg_ctx* ctx, int level, const double* bc_d4)
{
    MGRef* m = get(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg2d_direct_solve: no hierarchy");
    EBG_CHECK(level >= 0 && level < m->H.nlev(), EBG_ERR_ARG, "ebg_mg2d_direct_solve: bad level");
    const MGLevelDev<double>& l = m->H.lev[level];
    const int64_t n = mg_level_unknowns(l.xnum, l.ynum);
    EBG_CHECK(n <= 20000, EBG_ERR_ARG, "ebg_mg2d_direct_solve: level has more than 20000 unknowns");
    const double one[4] = {1.0, 1.0, 1.0, 1.0};
    const double* bcd = bc_d4 ? bc_d4 : one;
    hipStream_t s = ctx->stream;
    DevBuf A, W, ipiv, info, inv, vec;
    int rc;
    if ((rc = A.ensure((size_t)n * n * sizeof(double))) || (rc = W.ensure((size_t)n * n * sizeof(double))) ||
        (rc = ipiv.ensure((size_t)n * sizeof(int))) ||
        (rc = info.ensure(sizeof(int))) || (rc = inv.ensure((size_t)n * n * sizeof(double))) ||
        (rc = vec.ensure((size_t)2 * n * sizeof(double))))
        return rc;
    rocblas_handle h = nullptr;
    EBG_CHECK(rocblas_create_handle(&h) == rocblas_status_success, EBG_ERR_HIP,
              "ebg_mg2d_direct_solve: rocblas_create_handle failed");
    rc = mg_coarse_setup(m->H, level, bcd, A.as<double>(), W.as<double>(), ipiv.as<int>(), info.as<int>(),
                         inv.as<double>(), h, s);
    if (rc == EBG_OK) {
        // solve on a view of the hierarchy that ends at `level`
        MGHierarchy<double>& H = m->H;
        std::vector<MGLevelDev<double>> saved(H.lev.begin() + level + 1, H.lev.end());
        H.lev.resize(level + 1);
        H.coarse_inv = inv.as<double>();
        H.coarse_b = vec.as<double>();
        H.coarse_x = vec.as<double>() + n;
        H.coarse_n = (int)n;
        mg_coarse_solve(H, bcd, s);
        H.coarse_inv = nullptr;
        H.lev.insert(H.lev.end(), saved.begin(), saved.end());
        EBG_HIP(hipStreamSynchronize(s));
    } else {
        (void)hipStreamSynchronize(s);
    }
    (void)rocblas_destroy_handle(h);
    return rc;
}

// MultigridSolver.run_multigrid_solver for MultigridData2d
// (MultigridSolver.jl:113-135): viscosity restriction, saved originals,
// viscosity continuation from constant viscosity (MultigridDataManager.jl:321-373,
// ViscosityScaling.jl:26-39, 62-79), V-cycles (MultigridVCycle.jl:197-267)
// with mean principal residuals (CalculateMeanResiduals.jl:130-172).
// use_multimulti = false: a jump every nvcycles_viscosity_jump cycles
// (ViscosityScaling.jl:112-131, no residual trigger in 2-D) and the test of
// MultigridVCycle.jl:414-434, 454-468.  use_multimulti = true: level 0 (the
// global solution, original viscosity and right-hand sides) and level 1's
// correction problem (Smoother.jl:117-150, ViscosityScaling.jl:170-205,
// 207-337) with the global-residual test (MultigridVCycle.jl:424-431, 488-504).
static int mg2d_run_impl(ebg_ctx* ctx, const ebg_mg_run_params* prm, const double* bc_d4, bool multimulti,
                         int* cycles_done, int* converged, double* history, int* global_updates,
                         double* global_history, int global_capacity)
{
    MGRef* m = get(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg2d_run: no hierarchy");
    EBG_CHECK(prm, EBG_ERR_ARG, "ebg_mg2d_run: params is NULL");
    const ebg_mg_run_params& p = *prm;
    EBG_CHECK(p.nvcycles >= 0 && p.finest_sweeps >= 0 && p.sweep_multiplier >= 1 && p.nviscosity_jumps > 0 &&
                  p.stokesscale != 0 && p.continscale != 0,
              EBG_ERR_ARG, "ebg_mg2d_run: bad parameters");
    hipStream_t s = ctx->stream;
    MGHierarchy<double>& H = m->H;
    const int nl = H.nlev();
    auto nvisc = [&](int k, int f) -> int64_t {
        const int64_t y = H.lev[k].ynum, x = H.lev[k].xnum;
        return f == 0 ? y * x : (y - 1) * (x - 1);
    };
    auto field = [&](int k, int f) { return f == 0 ? H.lev[k].etas : H.lev[k].etan; };
    if (!m->partial) {   // first run (or a failed first allocation): allocate everything
        for (auto& a : m->orig)
            for (double*& q : a) {
                if (q) (void)hipFree(q);
                q = nullptr;
            }
        m->orig.assign(nl, {nullptr, nullptr});
        for (int k = 0; k < nl; ++k)
            for (int f = 0; f < 2; ++f) EBG_HIP(hipMalloc(&m->orig[k][f], (size_t)nvisc(k, f) * sizeof(double)));
        EBG_HIP(hipMalloc(&m->partial, (size_t)MG_RUN_BLOCKS * 3 * sizeof(double)));
    }
    const double one[4] = {1.0, 1.0, 1.0, 1.0};
    const double* bcd = bc_d4 ? bc_d4 : one;
    MGLevelDev<double>& l1 = H.lev[0];
    const int64_t ny = l1.ynum, nx = l1.xnum;
    const int64_t nvx = (ny + 1) * nx, nvy = ny * (nx + 1), npr = (ny - 1) * (nx - 1);
    const int64_t sizes[3] = {nvx, nvy, npr};
    // level 0 view: its own solution and right-hand sides and residual arrays
    MGLevelDev<double> l0 = l1;
    auto global_residual = [&]() {
        // calculate_residuals_without_smoothing, then level 1's right-hand sides
        hipLaunchKernelGGL(mg_bc_kernel<double>, dim3(nblk(mg_bc_threads(l0.xnum, l0.ynum))), dim3(256), 0, s, l0,
                           bcd[0], bcd[1], bcd[2], bcd[3]);
        hipLaunchKernelGGL(mg_pressure_shift_kernel<double>, dim3(1), dim3(64), 0, s, l0.pr, (const double*)nullptr,
                           p.pressure_bc, H.scratch);
        hipLaunchKernelGGL(mg_pressure_add_kernel<double>, dim3(nblk(npr)), dim3(256), 0, s, l0.pr, (int)npr,
                           (const double*)H.scratch);
        hipLaunchKernelGGL(mg_residual_kernel<double>, dim3(nblk((ny + 1) * (nx + 1))), dim3(256), 0, s, l0);
        const double* src[3] = {l0.rx, l0.ry, l0.rc};
        double* dst[3] = {l1.RX, l1.RY, l1.RC};
        for (int q = 0; q < 3; ++q)
            (void)hipMemcpyAsync(dst[q], src[q], sizes[q] * sizeof(double), hipMemcpyDeviceToDevice, s);
    };
    if (multimulti) {
        if (!m->lev0[0]) {
            const int64_t sz[9] = {nvx, nvy, npr, nvx, nvy, npr, nvx, nvy, npr};
            for (int q = 0; q < 9; ++q) EBG_HIP(hipMalloc(&m->lev0[q], (size_t)sz[q] * sizeof(double)));
        }
        l0.vx = m->lev0[0];
        l0.vy = m->lev0[1];
        l0.pr = m->lev0[2];
        l0.RX = m->lev0[3];
        l0.RY = m->lev0[4];
        l0.RC = m->lev0[5];
        l0.rx = m->lev0[6];
        l0.ry = m->lev0[7];
        l0.rc = m->lev0[8];
        l0.fix_vx = l0.fix_vy = l0.fix_p = nullptr;
        // the caller's initial guess and right-hand sides belong to level 0
        // (initialize_pressure_and_rhs_parts!, MultigridDataManager.jl:376-421)
        double* from[6] = {l1.vx, l1.vy, l1.pr, l1.RX, l1.RY, l1.RC};
        double* to[6] = {l0.vx, l0.vy, l0.pr, l0.RX, l0.RY, l0.RC};
        for (int q = 0; q < 6; ++q)
            EBG_HIP(hipMemcpyAsync(to[q], from[q], sizes[q % 3] * sizeof(double), hipMemcpyDeviceToDevice, s));
        mg_zero_solution(H, 0, s);
        global_residual();   // update_rhs_parts_on_level1_using_global_level0_residuals
        EBG_HIP(hipGetLastError());
    }
    int rc;
    double vmin = INFINITY, vmax = -INFINITY;
    for (int f = 0; f < 2; ++f)
        if ((rc = mg_device_range(field(0, f), nvisc(0, f), m->partial, s, vmin, vmax))) return rc;
    EBG_CHECK(vmin > 0.0 && std::isfinite(vmax), EBG_ERR_ARG, "ebg_mg2d_run: viscosity must be positive and finite");
    ViscRamp ramp;
    ramp.init(vmin, vmax, p.nviscosity_jumps);
    const double max_start = vmin;   // set_initial_viscosity_scaling_parameters
    mg_restrict_viscosity(H, s);
    double omin = INFINITY, omax = -INFINITY;
    for (int k = 0; k < nl; ++k)
        for (int f = 0; f < 2; ++f) {
            EBG_HIP(hipMemcpyAsync(m->orig[k][f], field(k, f), (size_t)nvisc(k, f) * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
            if ((rc = mg_device_range(m->orig[k][f], nvisc(k, f), m->partial, s, omin, omax))) return rc;
        }
    // level 0 keeps the original viscosity (update_level0_viscosity_using_original!)
    l0.etas = m->orig[0][0];
    l0.etan = m->orig[0][1];
    // a negative restricted viscosity under a fractional power is the
    // reference's DomainError
    auto scale_all = [&]() -> int {
        const double coef = ramp.coef();
        EBG_CHECK(omin >= 0.0 || coef == std::floor(coef), EBG_ERR_ARG,
                  "ebg_mg2d_run: restricted viscosity is negative on a coarse level (non-integer power)");
        for (int k = 0; k < nl; ++k)
            for (int f = 0; f < 2; ++f)
                mg_scale_viscosity(m->orig[k][f], field(k, f), nvisc(k, f), ramp.min_cur, 1.0 / ramp.min_o, coef, s);
        return EBG_OK;
    };
    if (ramp.below_full() && (rc = scale_all())) return rc;   // rescale_viscosity!
    const auto sm = mg_smoothing_counts(nl, p.finest_sweeps, p.sweep_multiplier, p.sweep_cap);
    const double cnt[3] = {(double)((ny - 1) * (nx - 2)), (double)((ny - 2) * (nx - 1)), (double)((ny - 1) * (nx - 1))};
    double part[3 * MG_RUN_BLOCKS];
    auto mean_residuals = [&](const MGLevelDev<double>& l, double* res) -> int {
        hipLaunchKernelGGL(mean_residual2d_partial_kernel, dim3(MG_RUN_BLOCKS), dim3(256), 0, s, l, p.stokesscale,
                           p.continscale, m->partial);
        EBG_HIP(hipMemcpyAsync(part, m->partial, sizeof(part), hipMemcpyDeviceToHost, s));
        EBG_HIP(hipStreamSynchronize(s));
        for (int q = 0; q < 3; ++q) {
            double sum = 0.0;
            for (int b = 0; b < MG_RUN_BLOCKS; ++b) sum += part[3 * b + q];
            res[q] = std::log10(std::sqrt(sum / cnt[q]));
        }
        return EBG_OK;
    };
    int jump = 1, overstep = 0, iglobal = 1, done = 0, conv = 0;
    double glast[3] = {0.0, 0.0, 0.0};
    for (int ivcycle = 1; ivcycle <= p.nvcycles; ++ivcycle) {
        mg_vcycle(H, sm, p.theta_s, p.theta_c, p.relax_v, p.relax_p, p.pressure_bc, (const double*)nullptr, bcd, s);
        double res[3];
        if ((rc = mean_residuals(l1, res))) return rc;
        if (history)
            for (int q = 0; q < 3; ++q) history[3 * (ivcycle - 1) + q] = res[q];
        done = ivcycle;
        ++jump;
        if (!multimulti) {   // rescale_viscosity_for_levels! (2-D)
            if (jump > p.nvcycles_viscosity_jump && ramp.below_full()) {
                jump = 1;
                ramp.jump();
                if ((rc = scale_all())) return rc;
            }
            if (mg_max_principal_residual(res, 3, ivcycle) < p.convergence_criterion &&
                ivcycle > p.nvcycles_viscosity_jump * p.nviscosity_jumps) {
                conv = 1;
                break;
            }
            continue;
        }
        // rescale_viscosity_for_levels_using_multimulti! (ViscosityScaling.jl:170-205)
        if (jump > p.nvcycles_viscosity_jump) {
            jump = 1;
            ramp.max_cur = ramp.max_cur * ramp.max_factor;
            if (ramp.max_cur >= ramp.max_o) ++overstep;
            if (ramp.max_cur > ramp.max_o) ramp.max_cur = ramp.max_o;
            if (overstep > 1) {
                ramp.max_cur = max_start;
                overstep = 0;
                double* l1s[3] = {l1.vx, l1.vy, l1.pr};
                double* l0s[3] = {l0.vx, l0.vy, l0.pr};
                for (int q = 0; q < 3; ++q)
                    hipLaunchKernelGGL(mm_add_kernel, dim3(nblk(sizes[q])), dim3(256), 0, s, l0s[q],
                                       (const double*)l1s[q], sizes[q]);
                global_residual();
                if ((rc = mean_residuals(l0, glast))) return rc;
                if (global_history && iglobal <= global_capacity)
                    for (int q = 0; q < 3; ++q) global_history[3 * (iglobal - 1) + q] = glast[q];
                mg_zero_solution(H, 0, s);
                ++iglobal;
            }
            if ((rc = scale_all())) return rc;
        }
        double gmax = 1e38;   // calculate_max_global_residual_2d
        if (iglobal - 1 > 1) {
            gmax = 0.0;
            for (int q = 0; q < 3; ++q) {
                const double v = std::fabs(std::pow(10.0, glast[q]));
                if (std::isnan(v)) { gmax = NAN; break; }
                gmax = std::max(gmax, v);
            }
        }
        if (gmax < p.convergence_criterion) {
            conv = 1;
            break;
        }
    }
    if (multimulti) {   // the global solution is level 0's (print_final_solution_stats)
        double* l1s[3] = {l1.vx, l1.vy, l1.pr};
        const double* l0s[3] = {l0.vx, l0.vy, l0.pr};
        for (int q = 0; q < 3; ++q)
            EBG_HIP(hipMemcpyAsync(l1s[q], l0s[q], sizes[q] * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    EBG_HIP(hipGetLastError());
    if (cycles_done) *cycles_done = done;
    if (converged) *converged = conv;
    if (global_updates) *global_updates = iglobal - 1;
    return EBG_OK;
}

int ebg_mg2d_run(ebg_ctx* ctx, const ebg_mg_run_params* prm, const double* bc_d4, int* cycles_done, int* converged,
                 double* history)
{
    return mg2d_run_impl(ctx, prm, bc_d4, false, cycles_done, converged, history, nullptr, nullptr, 0);
}

int ebg_mg2d_run_multimulti(ebg_ctx* ctx, const ebg_mg_run_params* prm, const double* bc_d4, int* cycles_done,
                            int* converged, double* history, int* global_updates, double* global_history,
                            int global_capacity)
{
    EBG_CHECK(global_capacity >= 0, EBG_ERR_ARG, "ebg_mg2d_run_multimulti: global_capacity < 0");
    return mg2d_run_impl(ctx, prm, bc_d4, true, cycles_done, converged, history, global_updates, global_history,
                         global_capacity);
}

}  // extern "C"
