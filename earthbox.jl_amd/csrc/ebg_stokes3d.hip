// 3-D Stokes-continuity solve on the system of the reference's experimental
// 3-D multigrid (multigrid/smoother/residuals/Residuals.jl:10-560 on a uniform
// level grid with the free-slip ghost passes of BoundaryConditions.jl:92-174).
//
// The reference drives that system only with V-cycles (MultigridVCycle.jl:
// 79-135).  Here GMRES(m) in FP64 solves it, right-preconditioned by one
// V-cycle of the same multigrid in FP32 (mixed precision) or FP64 (plain):
// flexible (every preconditioned vector kept) or, to fit twice the restart
// length in HBM, with only the basis kept and the V-cycle applied once more
// to the combined correction at each restart.  Stiff inclusions leave a few
// small eigenvalues (their near-rigid motions) that short restarts cannot
// resolve; restart lengths up to 240 are supported for that reason.
// Unknowns: interior velocity nodes and every pressure cell, stored in the
// level layout [vx | vy | vz | pr] (ghost slots carried along, zero in every
// Krylov basis vector).  Equations: the interior momentum and continuity
// residuals, with the continuity equation of cell (1,1,1) replaced by the
// multigrid's pressure gauge p(1,1,1) = prnorm (PressureUpdate.jl:5-21),
// scaled by 1/etan(1,1,1).
#include "ebg_common.hpp"
#include "ebg_krylov.hpp"
#include "ebg_mg.hpp"
#include "ebg_mg3d.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>

namespace ebg {

static inline unsigned nbk(int64_t n) { return (unsigned)((n + 255) / 256); }

// b = rhs on the unknowns (ghost slots zero), gauge entry prnorm/etan(1,1,1)
__global__ void __launch_bounds__(256) stokes3d_pack_rhs_kernel(MG3LevelDev<double> l, const double* __restrict__ rhs,
                                                                double* __restrict__ b, double prnorm)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int X = l.xnum, Y = l.ynum, Z = l.znum;
    const int ni = Y + 1, nj = X + 1;
    if (t >= (int64_t)ni * nj * (Z + 1)) return;
    const int i = 1 + (int)(t % ni);
    const int64_t r = t / ni;
    const int j = 1 + (int)(r % nj), k = 1 + (int)(r / nj);
    const size_t ox = 0, oy = l.nvx(), oz = oy + l.nvy(), op = oz + l.nvz();
    if (j <= X) {
        const size_t q = ox + i3(Y + 1, X, i, j, k);
        b[q] = (i >= 2 && i <= Y && j >= 2 && j <= X - 1 && k >= 2 && k <= Z) ? rhs[q] : 0.0;
    }
    if (i <= Y) {
        const size_t q = oy + i3(Y, X + 1, i, j, k);
        b[q] = (i >= 2 && i <= Y - 1 && j >= 2 && j <= X && k >= 2 && k <= Z) ? rhs[q] : 0.0;
    }
    if (k <= Z) {
        const size_t q = oz + i3(Y + 1, X + 1, i, j, k);
        b[q] = (i >= 2 && i <= Y && j >= 2 && j <= X && k >= 2 && k <= Z - 1) ? rhs[q] : 0.0;
    }
    if (i <= Y - 1 && j <= X - 1 && k <= Z - 1) {
        const size_t q = op + i3(Y - 1, X - 1, i, j, k);
        b[q] = (i == 1 && j == 1 && k == 1) ? prnorm / l.etan[0] : rhs[q];
    }
}

template <typename S, typename D>
__global__ void __launch_bounds__(256) convert3_kernel(const S* __restrict__ a, D* __restrict__ b, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (D)a[t];
}

__global__ void __launch_bounds__(256) sub3_kernel(const double* __restrict__ b, const double* __restrict__ w,
                                                   double* __restrict__ r, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) r[t] = b[t] - w[t];
}

// pressure part of x shifted so that p(1,1,1) = prnorm exactly
__global__ void __launch_bounds__(256) gauge_shift_kernel(double* __restrict__ pr, int64_t n, double prnorm,
                                                          double p0)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) pr[t] = t == 0 ? prnorm : pr[t] + (prnorm - p0);
}

struct Stokes3State {
    int xnum = 0, ynum = 0, znum = 0;
    double xsize = 0, ysize = 0, zsize = 0;
    size_t nvx = 0, nvy = 0, nvz = 0, npr = 0, n = 0;
    DevBuf etan, etaxy, etaxz, etayz;  // system viscosity (FP64)
    std::unique_ptr<MG3Hierarchy<double>> Hd;
    std::unique_ptr<MG3Hierarchy<float>> Hf;
    int lev_d = -1, lev_f = -1;        // nlevels the hierarchies were built with
    bool visc_d = false, visc_f = false;
    DevBuf b, x, r, w, V, Z, partial, hdev, ydev, nrm;
};

static Stokes3State* state3(ebg_ctx* ctx) { return ctx ? (Stokes3State*)ctx->stokes3d : nullptr; }

int stokes3d_release(ebg_ctx* ctx)
{
    if (ctx && ctx->stokes3d) {
        delete (Stokes3State*)ctx->stokes3d;
        ctx->stokes3d = nullptr;
    }
    return EBG_OK;
}

// FP64 hierarchy (also the operator's level 0); rebuilt when nlevels changes
static int ensure_d(ebg_ctx* ctx, Stokes3State& st, int nlevels)
{
    if (st.Hd && st.lev_d == nlevels) return EBG_OK;
    EBG_HIP(hipStreamSynchronize(ctx->stream));
    st.Hd = std::make_unique<MG3Hierarchy<double>>();
    int rc = mg3_build(*st.Hd, st.xnum, st.ynum, st.znum, st.xsize, st.ysize, st.zsize, nlevels, true);
    if (rc) {
        st.Hd.reset();
        return rc;
    }
    st.lev_d = nlevels;
    st.visc_d = false;
    const MG3LevelDev<double>& l = st.Hd->lev[0];
    hipStream_t s = ctx->stream;
    EBG_HIP(hipMemcpyAsync(l.etan, st.etan.p, st.npr * 8, hipMemcpyDeviceToDevice, s));
    EBG_HIP(hipMemcpyAsync(l.etaxy, st.etaxy.p, st.etaxy.bytes, hipMemcpyDeviceToDevice, s));
    EBG_HIP(hipMemcpyAsync(l.etaxz, st.etaxz.p, st.etaxz.bytes, hipMemcpyDeviceToDevice, s));
    EBG_HIP(hipMemcpyAsync(l.etayz, st.etayz.p, st.etayz.bytes, hipMemcpyDeviceToDevice, s));
    return EBG_OK;
}

static int ensure_f(ebg_ctx* ctx, Stokes3State& st, int nlevels)
{
    hipStream_t s = ctx->stream;
    if (!(st.Hf && st.lev_f == nlevels)) {
        EBG_HIP(hipStreamSynchronize(s));
        st.Hf = std::make_unique<MG3Hierarchy<float>>();
        int rc = mg3_build(*st.Hf, st.xnum, st.ynum, st.znum, st.xsize, st.ysize, st.zsize, nlevels, true);
        if (rc) {
            st.Hf.reset();
            return rc;
        }
        st.lev_f = nlevels;
        st.visc_f = false;
    }
    if (!st.visc_f) {
        const MG3LevelDev<float>& l = st.Hf->lev[0];
        const struct { const DevBuf* a; float* d; } fields[4] = {
            {&st.etan, l.etan}, {&st.etaxy, l.etaxy}, {&st.etaxz, l.etaxz}, {&st.etayz, l.etayz}};
        for (const auto& f : fields) {
            const int64_t n = (int64_t)(f.a->bytes / 8);
            hipLaunchKernelGGL((convert3_kernel<double, float>), dim3(nbk(n)), dim3(256), 0, s,
                               (const double*)f.a->p, f.d, n);
        }
        mg3_restrict_viscosity(*st.Hf, s);
        st.visc_f = true;
    }
    return EBG_OK;
}

// operator view: x and y in the [vx | vy | vz | pr] layout
static MG3LevelDev<double> op_view(Stokes3State& st, double* x, double* y)
{
    MG3LevelDev<double> o = st.Hd->lev[0];
    o.vx = x;
    o.vy = x + st.nvx;
    o.vz = x + st.nvx + st.nvy;
    o.pr = x + st.nvx + st.nvy + st.nvz;
    o.rx = y;
    o.ry = y + st.nvx;
    o.rz = y + st.nvx + st.nvy;
    o.rc = y + st.nvx + st.nvy + st.nvz;
    return o;
}

template <typename T>
static void precondition3(Stokes3State& st, MG3Hierarchy<T>& H, const ebg_solve_opts& o, const std::vector<int>& sm,
                          const double* v, double* z, hipStream_t s)
{
    const MG3LevelDev<T>& l = H.lev[0];
    const size_t off[4] = {0, st.nvx, st.nvx + st.nvy, st.nvx + st.nvy + st.nvz};
    const size_t len[4] = {st.nvx, st.nvy, st.nvz, st.npr};
    T* rhs[4] = {l.RX, l.RY, l.RZ, l.RC};
    T* sol[4] = {l.vx, l.vy, l.vz, l.pr};
    for (int c = 0; c < 4; ++c)
        hipLaunchKernelGGL((convert3_kernel<double, T>), dim3(nbk((int64_t)len[c])), dim3(256), 0, s, v + off[c],
                           rhs[c], (int64_t)len[c]);
    mg3_zero_solution(H, 0, s);
    mg3_vcycle(H, sm, (T)o.theta_s, (T)o.theta_c, T(1), T(1), T(0), s);
    for (int c = 0; c < 4; ++c)
        hipLaunchKernelGGL((convert3_kernel<T, double>), dim3(nbk((int64_t)len[c])), dim3(256), 0, s, sol[c],
                           z + off[c], (int64_t)len[c]);
}

// longest restart cycle of the 3-D solver: the Krylov basis lives in HBM,
// (m+1) vectors of the level layout (545 MB each at 257^3 in FP64), plus m
// preconditioned vectors in the flexible mode
constexpr int MAX_RESTART3 = 240;
// out[q] = <w, V[q]> for q < k: DOT_CHUNK basis vectors per pass inside one
// launch, then one fixed-order reduction per column
static void multidot3(Stokes3State& st, const double* w, const double* V, int64_t ldv, int k, int64_t n, double* out,
                      hipStream_t s)
{
    hipLaunchKernelGGL(multidot_chunks_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, w, V, ldv, k, n,
                       st.partial.as<double>(), (const int*)nullptr);
    hipLaunchKernelGGL(reduce_columns_kernel, dim3((k + 3) / 4), dim3(256), 0, s, st.partial.as<double>(), RED_BLOCKS,
                       k, out, 0, (const int*)nullptr);
}

static int prepare3(ebg_ctx* ctx, Stokes3State& st, const ebg_solve_opts& o, std::vector<int>& sm)
{
    int rc;
    if ((rc = ensure_d(ctx, st, o.nlevels))) return rc;
    if (o.precision == 1) {
        if ((rc = ensure_f(ctx, st, o.nlevels))) return rc;
    } else if (!st.visc_d) {
        mg3_restrict_viscosity(*st.Hd, ctx->stream);
        st.visc_d = true;
    }
    const int nl = o.precision == 1 ? st.Hf->nlev() : st.Hd->nlev();
    sm = mg_smoothing_counts(nl, o.finest_sweeps, o.sweep_multiplier, o.sweep_cap);
    return EBG_OK;
}

static void apply_m(Stokes3State& st, const ebg_solve_opts& o, const std::vector<int>& sm, const double* v, double* z,
                    hipStream_t s)
{
    if (o.precision == 1) precondition3(st, *st.Hf, o, sm, v, z, s);
    else precondition3(st, *st.Hd, o, sm, v, z, s);
}

static int fgmres3(ebg_ctx* ctx, Stokes3State& st, const ebg_solve_opts& o, const double* rhs, double prnorm,
                   double* xout, ebg_solve_stats* stats)
{
    hipStream_t s = ctx->stream;
    const int64_t n = (int64_t)st.n;
    const int m = o.restart;
    const bool flexible = o.flexible != 0;
    int rc;
    if ((rc = st.b.ensure(n * 8)) || (rc = st.r.ensure(n * 8)) || (rc = st.w.ensure(n * 8)) ||
        (rc = st.V.ensure((size_t)(m + 1) * n * 8)) || (rc = st.Z.ensure((size_t)(flexible ? m : 1) * n * 8)) ||
        (rc = st.partial.ensure((size_t)RED_BLOCKS * (MAX_RESTART3 + 1) * 8)) ||
        (rc = st.hdev.ensure((size_t)(MAX_RESTART3 + 1) * 8 * 2)) ||
        (rc = st.ydev.ensure((size_t)(MAX_RESTART3 + 1) * 8)) || (rc = st.nrm.ensure(64)))
        return rc;
    std::vector<int> sm;
    if ((rc = prepare3(ctx, st, o, sm))) return rc;
    double* b = st.b.as<double>();
    double* x = xout;
    double* r = st.r.as<double>();
    double* w = st.w.as<double>();
    double* V = st.V.as<double>();
    double* Z = st.Z.as<double>();
    double* hd = st.hdev.as<double>();
    double* xpr = x + st.nvx + st.nvy + st.nvz;
    {
        const MG3LevelDev<double>& l0 = st.Hd->lev[0];
        const int64_t nt = (int64_t)(l0.ynum + 1) * (l0.xnum + 1) * (l0.znum + 1);
        hipLaunchKernelGGL(stokes3d_pack_rhs_kernel, dim3(nbk(nt)), dim3(256), 0, s, l0, rhs, b, prnorm);
    }
    if (o.use_initial_guess) {
        double p0 = 0;
        EBG_HIP(hipMemcpyAsync(&p0, xpr, 8, hipMemcpyDeviceToHost, s));
        EBG_HIP(hipStreamSynchronize(s));
        hipLaunchKernelGGL(gauge_shift_kernel, dim3(nbk((int64_t)st.npr)), dim3(256), 0, s, xpr, (int64_t)st.npr,
                           prnorm, p0);
    } else {
        hipLaunchKernelGGL(fill_kernel<double>, dim3(nbk(n)), dim3(256), 0, s, x, n, 0.0);
        hipLaunchKernelGGL(fill_kernel<double>, dim3(nbk((int64_t)st.npr)), dim3(256), 0, s, xpr, (int64_t)st.npr,
                           prnorm);
    }
    auto norm2 = [&](const double* v, double* out) { multidot3(st, v, v, 0, 1, n, out, s); };
    const int HOFF = MAX_RESTART3 + 1;  // second Gram-Schmidt pass coefficients
    std::vector<double> hbuf(2 * (size_t)HOFF);
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1);
    double bnorm2 = 0;
    norm2(b, st.nrm.as<double>());
    EBG_HIP(hipMemcpyAsync(&bnorm2, st.nrm.p, 8, hipMemcpyDeviceToHost, s));
    EBG_HIP(hipStreamSynchronize(s));
    const double bnorm = std::sqrt(bnorm2);
    if (!(bnorm > 0) || !std::isfinite(bnorm)) {
        if (stats) { stats->iterations = 0; stats->relres = 0; stats->relres_initial = 0; stats->learned_evaluated = 0; }
        mg3_set_bc(op_view(st, x, w), s);
        if (bnorm == 0) return EBG_OK;
        set_error("ebg_stokes3d_solve: right-hand side is not finite");
        return EBG_ERR_NUMERIC;
    }
    int total = 0, restarts = 0;
    double relres = 1.0, relres0 = -1.0;
    while (true) {
        mg3_apply(op_view(st, x, w), s);
        hipLaunchKernelGGL(sub3_kernel, dim3(nbk(n)), dim3(256), 0, s, b, w, r, n);
        norm2(r, st.nrm.as<double>());
        double beta2;
        EBG_HIP(hipMemcpyAsync(&beta2, st.nrm.p, 8, hipMemcpyDeviceToHost, s));
        EBG_HIP(hipStreamSynchronize(s));
        const double beta = std::sqrt(beta2);
        relres = beta / bnorm;
        if (relres0 < 0) relres0 = relres;
        if (!std::isfinite(relres)) {
            set_error("ebg_stokes3d_solve: residual became non-finite");
            return EBG_ERR_NUMERIC;
        }
        if (relres <= o.rtol || total >= o.max_iterations) break;
        EBG_HIP(hipMemcpyAsync(st.nrm.p, &beta, 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(scale_kernel, dim3(nbk(n)), dim3(256), 0, s, V, r, st.nrm.as<double>(), n);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        for (; k < m && total < o.max_iterations; ++k) {
            double* vk = V + (int64_t)k * n;
            double* zk = flexible ? Z + (int64_t)k * n : Z;
            apply_m(st, o, sm, vk, zk, s);
            mg3_apply(op_view(st, zk, w), s);
            // classical Gram-Schmidt; a second pass only when the first one
            // cancelled more than half of the norm ("twice is enough")
            norm2(w, st.nrm.as<double>() + 1);
            multidot3(st, w, V, n, k + 1, n, hd, s);
            hipLaunchKernelGGL(update_kernel, dim3(nbk(n)), dim3(256), 0, s, w, V, n, hd, k + 1, n);
            norm2(w, st.nrm.as<double>());
            double nn[2];
            EBG_HIP(hipMemcpyAsync(hbuf.data(), hd, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, s));
            EBG_HIP(hipMemcpyAsync(nn, st.nrm.p, 16, hipMemcpyDeviceToHost, s));
            EBG_HIP(hipStreamSynchronize(s));
            double wn2 = nn[0];
            for (int q = 0; q <= k; ++q) H[(size_t)k * (m + 1) + q] = hbuf[q];
            if (!(wn2 > REORTH_RATIO2 * nn[1])) {
                multidot3(st, w, V, n, k + 1, n, hd + HOFF, s);
                hipLaunchKernelGGL(update_kernel, dim3(nbk(n)), dim3(256), 0, s, w, V, n, hd + HOFF, k + 1, n);
                norm2(w, st.nrm.as<double>());
                EBG_HIP(hipMemcpyAsync(hbuf.data() + HOFF, hd + HOFF, sizeof(double) * (k + 1),
                                       hipMemcpyDeviceToHost, s));
                EBG_HIP(hipMemcpyAsync(&wn2, st.nrm.p, 8, hipMemcpyDeviceToHost, s));
                EBG_HIP(hipStreamSynchronize(s));
                for (int q = 0; q <= k; ++q) H[(size_t)k * (m + 1) + q] += hbuf[HOFF + q];
            }
            const double hn = std::sqrt(wn2);
            H[(size_t)k * (m + 1) + k + 1] = hn;
            if (!std::isfinite(hn)) {
                set_error("ebg_stokes3d_solve: preconditioned Krylov vector is non-finite "
                          "(multigrid preconditioner diverged)");
                return EBG_ERR_NUMERIC;
            }
            if (hn > 0) {
                EBG_HIP(hipMemcpyAsync(st.nrm.p, &hn, 8, hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(scale_kernel, dim3(nbk(n)), dim3(256), 0, s, V + (int64_t)(k + 1) * n, w,
                                   st.nrm.as<double>(), n);
            }
            double* col = &H[(size_t)k * (m + 1)];
            for (int q = 0; q < k; ++q) {
                const double t = cs[q] * col[q] + sn[q] * col[q + 1];
                col[q + 1] = -sn[q] * col[q] + cs[q] * col[q + 1];
                col[q] = t;
            }
            const double d = std::hypot(col[k], col[k + 1]);
            cs[k] = col[k] / d;
            sn[k] = col[k + 1] / d;
            col[k] = d;
            col[k + 1] = 0.0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            ++total;
            if (std::fabs(g[k + 1]) / bnorm <= o.rtol || hn == 0) {
                ++k;
                break;
            }
        }
        std::vector<double> y(k);
        for (int q = k - 1; q >= 0; --q) {
            double v = g[q];
            for (int p = q + 1; p < k; ++p) v -= H[(size_t)p * (m + 1) + q] * y[p];
            y[q] = v / H[(size_t)q * (m + 1) + q];
        }
        EBG_HIP(hipMemcpyAsync(st.ydev.p, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        if (flexible) {
            hipLaunchKernelGGL(combine_kernel, dim3(nbk(n)), dim3(256), 0, s, x, Z, n, st.ydev.as<double>(), k, n);
        } else {
            // x += M (V y): the preconditioner applied once to the combination
            static const double one = 1.0;
            hipLaunchKernelGGL(fill_kernel<double>, dim3(nbk(n)), dim3(256), 0, s, r, n, 0.0);
            hipLaunchKernelGGL(combine_kernel, dim3(nbk(n)), dim3(256), 0, s, r, V, n, st.ydev.as<double>(), k, n);
            apply_m(st, o, sm, r, Z, s);
            EBG_HIP(hipMemcpyAsync(st.ydev.p, &one, 8, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(combine_kernel, dim3(nbk(n)), dim3(256), 0, s, x, Z, n, st.ydev.as<double>(), 1, n);
        }
        ++restarts;
    }
    // leave the ghost nodes of the solution consistent with its interior
    mg3_set_bc(op_view(st, x, w), s);
    if (stats) {
        stats->iterations = total;
        stats->restarts = restarts;
        stats->relres = relres;
        stats->relres_initial = relres0;
        stats->learned_accepted = 0;
        stats->learned_evaluated = 0;
    }
    return relres <= o.rtol ? EBG_OK : EBG_NOT_CONVERGED;
}

}  // namespace ebg

using namespace ebg;

extern "C" {

int ebg_stokes3d_set_system(ebg_ctx* ctx, int64_t xnum, int64_t ynum, int64_t znum, double xsize, double ysize,
                            double zsize, const double* etan, const double* etaxy, const double* etaxz,
                            const double* etayz)
{
    EBG_CHECK(ctx, EBG_ERR_ARG, "ebg_stokes3d_set_system: ctx is NULL");
    EBG_CHECK(xnum >= 4 && ynum >= 4 && znum >= 4, EBG_ERR_ARG, "ebg_stokes3d_set_system: need xnum, ynum, znum >= 4");
    EBG_CHECK((double)(ynum + 1) * (double)(xnum + 1) * (double)(znum + 1) < 2.0e9, EBG_ERR_ARG,
              "ebg_stokes3d_set_system: level arrays must hold fewer than 2^31 entries");
    EBG_CHECK(xsize > 0 && ysize > 0 && zsize > 0, EBG_ERR_ARG, "ebg_stokes3d_set_system: sizes must be > 0");
    EBG_CHECK(etan && etaxy && etaxz && etayz, EBG_ERR_ARG, "ebg_stokes3d_set_system: NULL viscosity array");
    EBG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    EBG_HIP(hipStreamSynchronize(s));
    Stokes3State* st = state3(ctx);
    const bool same = st && st->xnum == xnum && st->ynum == ynum && st->znum == znum && st->xsize == xsize &&
                      st->ysize == ysize && st->zsize == zsize;
    if (!same) {
        stokes3d_release(ctx);
        st = new Stokes3State();
        ctx->stokes3d = st;
        st->xnum = (int)xnum;
        st->ynum = (int)ynum;
        st->znum = (int)znum;
        st->xsize = xsize;
        st->ysize = ysize;
        st->zsize = zsize;
        st->nvx = (size_t)(ynum + 1) * xnum * (znum + 1);
        st->nvy = (size_t)ynum * (xnum + 1) * (znum + 1);
        st->nvz = (size_t)(ynum + 1) * (xnum + 1) * znum;
        st->npr = (size_t)(ynum - 1) * (xnum - 1) * (znum - 1);
        st->n = st->nvx + st->nvy + st->nvz + st->npr;
    }
    const size_t nxy = (size_t)ynum * xnum * (znum - 1), nxz = (size_t)(ynum - 1) * xnum * znum,
                 nyz = (size_t)ynum * (xnum - 1) * znum;
    int rc;
    if ((rc = st->etan.ensure(st->npr * 8)) || (rc = st->etaxy.ensure(nxy * 8)) || (rc = st->etaxz.ensure(nxz * 8)) ||
        (rc = st->etayz.ensure(nyz * 8)))
        return rc;
    st->etan.bytes = st->npr * 8;
    st->etaxy.bytes = nxy * 8;
    st->etaxz.bytes = nxz * 8;
    st->etayz.bytes = nyz * 8;
    EBG_HIP(hipMemcpyAsync(st->etan.p, etan, st->npr * 8, hipMemcpyDeviceToDevice, s));
    EBG_HIP(hipMemcpyAsync(st->etaxy.p, etaxy, nxy * 8, hipMemcpyDeviceToDevice, s));
    EBG_HIP(hipMemcpyAsync(st->etaxz.p, etaxz, nxz * 8, hipMemcpyDeviceToDevice, s));
    EBG_HIP(hipMemcpyAsync(st->etayz.p, etayz, nyz * 8, hipMemcpyDeviceToDevice, s));
    if (st->Hd) {
        const MG3LevelDev<double>& l = st->Hd->lev[0];
        EBG_HIP(hipMemcpyAsync(l.etan, etan, st->npr * 8, hipMemcpyDeviceToDevice, s));
        EBG_HIP(hipMemcpyAsync(l.etaxy, etaxy, nxy * 8, hipMemcpyDeviceToDevice, s));
        EBG_HIP(hipMemcpyAsync(l.etaxz, etaxz, nxz * 8, hipMemcpyDeviceToDevice, s));
        EBG_HIP(hipMemcpyAsync(l.etayz, etayz, nyz * 8, hipMemcpyDeviceToDevice, s));
    }
    st->visc_d = false;
    st->visc_f = false;
    return EBG_OK;
}

int ebg_stokes3d_vector_length(ebg_ctx* ctx, int64_t* n)
{
    Stokes3State* st = state3(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_stokes3d_vector_length: no system (ebg_stokes3d_set_system)");
    EBG_CHECK(n, EBG_ERR_ARG, "ebg_stokes3d_vector_length: n is NULL");
    *n = (int64_t)st->n;
    return EBG_OK;
}

int ebg_stokes3d_apply(ebg_ctx* ctx, double* x, double* y)
{
    Stokes3State* st = state3(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_stokes3d_apply: no system (ebg_stokes3d_set_system)");
    EBG_CHECK(x && y && x != y, EBG_ERR_ARG, "ebg_stokes3d_apply: need distinct x and y");
    int rc = ensure_d(ctx, *st, st->lev_d < 0 ? 0 : st->lev_d);
    if (rc) return rc;
    mg3_apply(op_view(*st, x, y), ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_stokes3d_precondition(ebg_ctx* ctx, const double* r, double* z, const ebg_solve_opts* opts)
{
    Stokes3State* st = state3(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_stokes3d_precondition: no system (ebg_stokes3d_set_system)");
    EBG_CHECK(r && z && opts, EBG_ERR_ARG, "ebg_stokes3d_precondition: NULL argument");
    std::vector<int> sm;
    int rc = prepare3(ctx, *st, *opts, sm);
    if (rc) return rc;
    apply_m(*st, *opts, sm, r, z, ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_stokes3d_solve(ebg_ctx* ctx, const double* rhs, double prnorm, double* x, const ebg_solve_opts* opts,
                       ebg_solve_stats* stats)
{
    Stokes3State* st = state3(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_stokes3d_solve: no system (ebg_stokes3d_set_system)");
    EBG_CHECK(rhs && x, EBG_ERR_ARG, "ebg_stokes3d_solve: NULL rhs or x");
    ebg_solve_opts o;
    if (opts) o = *opts;
    else ebg_stokes_solve_opts_default(&o);
    EBG_CHECK(o.restart >= 1 && o.restart <= MAX_RESTART3, EBG_ERR_ARG,
              "ebg_stokes3d_solve: restart must be in 1..240");
    EBG_CHECK(o.precision == 0 || o.precision == 1, EBG_ERR_ARG, "ebg_stokes3d_solve: precision must be 0 or 1");
    EBG_CHECK(o.finest_sweeps >= 0 && o.sweep_multiplier >= 1, EBG_ERR_ARG, "ebg_stokes3d_solve: bad sweep counts");
    EBG_HIP(hipSetDevice(ctx->device));
    const auto t0 = std::chrono::steady_clock::now();
    int rc = fgmres3(ctx, *st, o, rhs, prnorm, x, stats);
    if (rc < 0) return rc;
    EBG_HIP(hipStreamSynchronize(ctx->stream));
    if (stats) {
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        stats->status = rc;
    }
    return rc;
}

}  // extern "C"
