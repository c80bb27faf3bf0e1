// Host side of the multigrid hierarchy: level planning, uniform level grids,
// fine->coarse maps with gather ranges, device allocation and V-cycle driver.
#include "ebg_common.hpp"
#include "ebg_mg.hpp"
#include "ebg_mg_tile.hpp"

#include <algorithm>
#include <cmath>

namespace ebg {

// multigrid/level_manager/LevelManager.jl:441-457
std::vector<std::pair<int, int>> mg_level_sizes(int xnum, int ynum, int max_levels)
{
    std::vector<std::pair<int, int>> s{{xnum, ynum}};
    while (std::min(s.back().first, s.back().second) > 4) {
        const int xn = (s.back().first + 1) / 2, yn = (s.back().second + 1) / 2;
        if (std::min(xn, yn) < 4 || (int)s.size() >= max_levels) break;
        s.push_back({xn, yn});
    }
    return s;
}

// multigrid_data/MultigridDataManager.jl:223-241
std::vector<int> mg_smoothing_counts(int nlev, int finest, int multiplier, int cap)
{
    std::vector<int> s(nlev);
    s[0] = finest;
    for (int i = 1; i < nlev; ++i) s[i] = s[i - 1] * multiplier;
    if (cap > 0)
        for (int i = 1; i < nlev; ++i) s[i] = std::min(s[i], cap);
    return s;
}

void basic_axis(int n, double step, std::vector<double>& g, std::vector<double>& stp)
{
    g.assign(n, 0.0);
    for (int i = 1; i < n; ++i) g[i] = g[i - 1] + step;
    stp.assign(n - 1, 0.0);
    for (int i = 0; i < n - 1; ++i) stp[i] = g[i + 1] - g[i];
}
void staggered_axis(int n, const std::vector<double>& g, const std::vector<double>& stp,
                    std::vector<double>& gs, std::vector<double>& stps)
{
    gs.assign(n + 1, 0.0);
    gs[0] = g[0] - stp[0] / 2.0;
    gs[n] = g[n - 1] + stp[n - 2] / 2.0;
    for (int i = 1; i < n; ++i) gs[i] = (g[i] + g[i - 1]) / 2.0;
    stps.assign(n, 0.0);
    stps[0] = stp[0];
    stps[n - 1] = stp[n - 2];
    for (int i = 1; i < n - 1; ++i) stps[i] = (g[i + 1] - g[i - 1]) / 2.0;
}
void pressure_axis(int n, const std::vector<double>& g, std::vector<double>& gp)
{
    std::vector<double> stpp(n - 2);
    for (int i = 0; i < n - 2; ++i) stpp[i] = (g[i + 2] - g[i]) / 2.0;
    gp.assign(n - 1, 0.0);
    gp[0] = (g[0] + g[1]) / 2.0;
    for (int i = 1; i < n - 1; ++i) gp[i] = gp[i - 1] + stpp[i - 1];
}

// Uniform level grid as staggered_grid/types_manager/uniform/core/Initialization.jl:27-43
// and the staggered-grid tools build it.
HostLevelGrid make_uniform_level_grid(int xnum, int ynum, double xsize, double ysize)
{
    HostLevelGrid h;
    h.xnum = xnum;
    h.ynum = ynum;
    basic_axis(xnum, xsize / (double)(xnum - 1), h.gridx_b, h.xstp_b);
    basic_axis(ynum, ysize / (double)(ynum - 1), h.gridy_b, h.ystp_b);
    staggered_axis(xnum, h.gridx_b, h.xstp_b, h.gridx_vy, h.xstp_vy);
    staggered_axis(ynum, h.gridy_b, h.ystp_b, h.gridy_vx, h.ystp_vx);
    pressure_axis(xnum, h.gridx_b, h.gridx_pr);
    pressure_axis(ynum, h.gridy_b, h.gridy_pr);
    return h;
}

template <typename T>
static int dev_upload(MGHierarchy<T>& H, const void* src, size_t bytes, void** out)
{
    void* p = nullptr;
    EBG_HIP(hipMalloc(&p, bytes ? bytes : 4));
    H.allocations.push_back(p);
    if (bytes) EBG_HIP(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    *out = p;
    return EBG_OK;
}

template <typename T>
static int dev_zeros(MGHierarchy<T>& H, size_t n, T** out)
{
    void* p = nullptr;
    EBG_HIP(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
    H.allocations.push_back(p);
    EBG_HIP(hipMemset(p, 0, sizeof(T) * (n ? n : 1)));
    *out = (T*)p;
    return EBG_OK;
}

template <typename T>
static int upload_axis(MGHierarchy<T>& H, const HostAxisMap<T>& m, AxisMapDev<T>& d)
{
    int rc;
    void* p;
    if ((rc = dev_upload(H, m.idx.data(), m.idx.size() * sizeof(int), &p))) return rc;
    d.idx = (const int*)p;
    if ((rc = dev_upload(H, m.d.data(), m.d.size() * sizeof(T), &p))) return rc;
    d.d = (const T*)p;
    if ((rc = dev_upload(H, m.lo.data(), m.lo.size() * sizeof(int), &p))) return rc;
    d.lo = (const int*)p;
    if ((rc = dev_upload(H, m.hi.data(), m.hi.size() * sizeof(int), &p))) return rc;
    d.hi = (const int*)p;
    return EBG_OK;
}

template <typename T>
static int upload_vec(MGHierarchy<T>& H, const std::vector<double>& v, const T** out)
{
    std::vector<T> t(v.begin(), v.end());
    void* p;
    int rc = dev_upload(H, t.data(), t.size() * sizeof(T), &p);
    *out = (const T*)p;
    return rc;
}

template <typename T>
MGHierarchy<T>::~MGHierarchy()
{
    for (void* p : allocations) (void)hipFree(p);
}

template <typename T>
int mg_build(MGHierarchy<T>& H, int xnum, int ynum, double xsize, double ysize, int nlevels,
             const std::vector<HostLevelGrid>* override_grids, bool cw)
{
    auto sizes = mg_level_sizes(xnum, ynum, 10);
    if (nlevels > 0 && nlevels < (int)sizes.size()) sizes.resize(nlevels);
    H.hgrid.clear();
    for (size_t k = 0; k < sizes.size(); ++k) {
        if (k == 0 && override_grids && !override_grids->empty()) H.hgrid.push_back((*override_grids)[0]);
        else H.hgrid.push_back(make_uniform_level_grid(sizes[k].first, sizes[k].second, xsize, ysize));
    }
    const int nl = (int)H.hgrid.size();
    H.lev.resize(nl);
    int rc;
    for (int k = 0; k < nl; ++k) {
        const HostLevelGrid& g = H.hgrid[k];
        MGLevelDev<T>& l = H.lev[k];
        l.xnum = g.xnum;
        l.ynum = g.ynum;
        const size_t ny = g.ynum, nx = g.xnum;
        if ((rc = upload_vec(H, g.xstp_b, &l.xstp_b))) return rc;
        if ((rc = upload_vec(H, g.ystp_b, &l.ystp_b))) return rc;
        if ((rc = upload_vec(H, g.xstp_vy, &l.xstp_vy))) return rc;
        if ((rc = upload_vec(H, g.ystp_vx, &l.ystp_vx))) return rc;
        if ((rc = dev_zeros(H, ny * nx, &l.etas))) return rc;
        if ((rc = dev_zeros(H, (ny - 1) * (nx - 1), &l.etan))) return rc;
        if ((rc = dev_zeros(H, (ny + 1) * nx, &l.vx))) return rc;
        if ((rc = dev_zeros(H, ny * (nx + 1), &l.vy))) return rc;
        if ((rc = dev_zeros(H, (ny - 1) * (nx - 1), &l.pr))) return rc;
        if (k > 0) {
            if ((rc = dev_zeros(H, (ny + 1) * nx, &l.vx_alt))) return rc;
            if ((rc = dev_zeros(H, ny * (nx + 1), &l.vy_alt))) return rc;
        }
        if ((rc = dev_zeros(H, (ny - 1) * (nx - 1), &l.pr_alt))) return rc;
        if ((rc = dev_zeros(H, (ny + 1) * nx, &l.RX))) return rc;
        if ((rc = dev_zeros(H, ny * (nx + 1), &l.RY))) return rc;
        if ((rc = dev_zeros(H, (ny - 1) * (nx - 1), &l.RC))) return rc;
        if ((rc = dev_zeros(H, (ny + 1) * nx, &l.rx))) return rc;
        if ((rc = dev_zeros(H, ny * (nx + 1), &l.ry))) return rc;
        if ((rc = dev_zeros(H, (ny - 1) * (nx - 1), &l.rc))) return rc;
    }
    H.xfer.resize(nl > 0 ? nl - 1 : 0);
    for (int k = 0; k + 1 < nl; ++k) {
        const HostLevelGrid& f = H.hgrid[k];
        const HostLevelGrid& c = H.hgrid[k + 1];
        const int yf = f.ynum, xf = f.xnum;
        TransferDev<T>& t = H.xfer[k];
        // residual restriction loop ranges, Restriction.jl:444-463
        if ((rc = upload_axis(H, axis_map<T>(f.gridy_vx, c.gridy_vx, 2, yf, cw), t.vx_y))) return rc;
        if ((rc = upload_axis(H, axis_map<T>(f.gridx_b, c.gridx_b, 2, xf - 1, cw), t.vx_x))) return rc;
        if ((rc = upload_axis(H, axis_map<T>(f.gridy_b, c.gridy_b, 2, yf - 1, cw), t.vy_y))) return rc;
        if ((rc = upload_axis(H, axis_map<T>(f.gridx_vy, c.gridx_vy, 2, xf, cw), t.vy_x))) return rc;
        if ((rc = upload_axis(H, axis_map<T>(f.gridy_pr, c.gridy_pr, 2, yf - 1, cw), t.pr_y))) return rc;
        if ((rc = upload_axis(H, axis_map<T>(f.gridx_pr, c.gridx_pr, 2, xf - 1, cw), t.pr_x))) return rc;
        // viscosity restriction ranges, ViscosityRestriction.jl:202-215
        if ((rc = upload_axis(H, axis_map<T>(f.gridy_b, c.gridy_b, 1, yf, cw), t.b_y))) return rc;
        if ((rc = upload_axis(H, axis_map<T>(f.gridx_b, c.gridx_b, 1, xf, cw), t.b_x))) return rc;
        if ((rc = upload_axis(H, axis_map<T>(f.gridy_pr, c.gridy_pr, 1, yf - 1, cw), t.e_y))) return rc;
        if ((rc = upload_axis(H, axis_map<T>(f.gridx_pr, c.gridx_pr, 1, xf - 1, cw), t.e_x))) return rc;
    }
    if ((rc = dev_zeros(H, 4, &H.scratch))) return rc;
    return EBG_OK;
}

static inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

template <typename T>
void mg_restrict_viscosity(MGHierarchy<T>& H, hipStream_t s)
{
    for (int k = 0; k + 1 < H.nlev(); ++k) {
        const MGLevelDev<T>& f = H.lev[k];
        const MGLevelDev<T>& c = H.lev[k + 1];
        const TransferDev<T>& t = H.xfer[k];
        hipLaunchKernelGGL((mg_visc_restrict_kernel<T, 0>), dim3(nblk((int64_t)c.ynum * c.xnum)),
                           dim3(256), 0, s, f, c, t.b_y, t.b_x);
        hipLaunchKernelGGL((mg_visc_restrict_kernel<T, 1>),
                           dim3(nblk((int64_t)(c.ynum - 1) * (c.xnum - 1))), dim3(256), 0, s, f, c,
                           t.e_y, t.e_x);
    }
}

template <typename T>
void mg_zero_solution(MGHierarchy<T>& H, int k, hipStream_t s)
{
    const MGLevelDev<T>& l = H.lev[k];
    const int64_t nvx = (int64_t)(l.ynum + 1) * l.xnum, nvy = (int64_t)l.ynum * (l.xnum + 1),
                  np = (int64_t)(l.ynum - 1) * (l.xnum - 1);
    hipLaunchKernelGGL(fill_kernel<T>, dim3(nblk(nvx)), dim3(256), 0, s, l.vx, nvx, T(0));
    hipLaunchKernelGGL(fill_kernel<T>, dim3(nblk(nvy)), dim3(256), 0, s, l.vy, nvy, T(0));
    hipLaunchKernelGGL(fill_kernel<T>, dim3(nblk(np)), dim3(256), 0, s, l.pr, np, T(0));
}

// smoother/Smoother.jl:66-99 with StokesSolverGaussSeidel.jl:33-54
template <typename T>
void mg_smoother(MGHierarchy<T>& H, int k, int nsweeps, T theta_s, T theta_c, T prnorm,
                 const T* prnorm_dev, const T* bc_d, hipStream_t s)
{
    const bool normalize = (k > 0) || H.normalize_finest;
    const MGLevelDev<T>& l = H.lev[k];
    const int nbc = mg_bc_threads(l.xnum, l.ynum);
    const int64_t nvx = (int64_t)(l.ynum - 1) * (l.xnum - 2);
    const int64_t nvy = (int64_t)(l.ynum - 2) * (l.xnum - 1);
    const int64_t np = (int64_t)(l.ynum - 1) * (l.xnum - 1);
    for (int it = 0; it < nsweeps; ++it) {
        hipLaunchKernelGGL(mg_bc_kernel<T>, dim3(nblk(nbc)), dim3(256), 0, s, l, bc_d[0], bc_d[1],
                           bc_d[2], bc_d[3]);
        for (int color = 0; color <= 1; ++color) {
            hipLaunchKernelGGL((mg_sweep_kernel<T, 0>), dim3(nblk(nvx)), dim3(256), 0, s, l, color,
                               theta_s, theta_c);
            hipLaunchKernelGGL((mg_sweep_kernel<T, 1>), dim3(nblk(nvy)), dim3(256), 0, s, l, color,
                               theta_s, theta_c);
            hipLaunchKernelGGL((mg_sweep_kernel<T, 2>), dim3(nblk(np)), dim3(256), 0, s, l, color,
                               theta_s, theta_c);
        }
    }
    if (normalize) {
        hipLaunchKernelGGL(mg_pressure_shift_kernel<T>, dim3(1), dim3(64), 0, s, l.pr, prnorm_dev, prnorm,
                           H.scratch);
        hipLaunchKernelGGL(mg_pressure_add_kernel<T>, dim3(nblk(np)), dim3(256), 0, s, l.pr, (int)np,
                           H.scratch);
    }
    hipLaunchKernelGGL(mg_residual_kernel<T>, dim3(nblk((int64_t)(l.ynum + 1) * (l.xnum + 1))),
                       dim3(256), 0, s, l);
}

// vcycle/core/SmoothAndRestrict.jl:44-71 and SmoothAndProlongate.jl:60-115
template <typename T>
void mg_vcycle(MGHierarchy<T>& H, const std::vector<int>& sm, T theta_s, T theta_c, T relax_v,
               T relax_p, T prnorm, const T* prnorm_dev, const T* bc_d, hipStream_t s)
{
    const int nl = H.nlev();
    const bool direct = H.coarse_inv && nl > 1;   // exact solve replaces the coarsest smoothing
    for (int n = 0; n < nl; ++n) {
        if (direct && n == nl - 1) {
            mg_coarse_solve(H, bc_d, s);
            break;
        }
        if (n > 0) mg_zero_solution(H, n, s);
        mg_smoother(H, n, sm[n], theta_s, theta_c, prnorm, prnorm_dev, bc_d, s);
        if (n < nl - 1) {
            const MGLevelDev<T>& f = H.lev[n];
            const MGLevelDev<T>& c = H.lev[n + 1];
            const TransferDev<T>& t = H.xfer[n];
            hipLaunchKernelGGL((mg_restrict_kernel<T, 0>), dim3(nblk((int64_t)(c.ynum + 1) * c.xnum)),
                               dim3(256), 0, s, f, c, t.vx_y, t.vx_x);
            hipLaunchKernelGGL((mg_restrict_kernel<T, 1>), dim3(nblk((int64_t)c.ynum * (c.xnum + 1))),
                               dim3(256), 0, s, f, c, t.vy_y, t.vy_x);
            hipLaunchKernelGGL((mg_restrict_kernel<T, 2>),
                               dim3(nblk((int64_t)(c.ynum - 1) * (c.xnum - 1))), dim3(256), 0, s, f, c,
                               t.pr_y, t.pr_x);
        }
    }
    for (int n = nl - 1; n >= 0; --n) {
        if (n == 0) {
            mg_smoother(H, 0, sm[0], theta_s, theta_c, prnorm, prnorm_dev, bc_d, s);
        } else {
            if (!(direct && n == nl - 1))
                mg_smoother(H, n, sm[n], theta_s, theta_c, T(0), (const T*)nullptr, bc_d, s);
            const MGLevelDev<T>& c = H.lev[n];
            const MGLevelDev<T>& f = H.lev[n - 1];
            hipLaunchKernelGGL(mg_prolong_kernel<T>, dim3(nblk((int64_t)(f.ynum + 1) * (f.xnum + 1))),
                               dim3(256), 0, s, c, f, H.xfer[n - 1], relax_v, relax_p);
        }
    }
}

template <typename T>
int mg_plan_tiles(MGHierarchy<T>& H, int tile, bool allow_single, bool fast)
{
    int dev = 0, max_lds = 65536;
    EBG_HIP(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        max_lds = 65536;
    size_t need = 0;
    H.tiles.clear();
    for (int k = 0; k < H.nlev(); ++k) {
        const MGLevelDev<T>& l = H.lev[k];
        typename MGHierarchy<T>::TilePlan p;
        p.whole_lds = tile_lds_bytes<T>(l.ynum + 1, l.xnum + 1);
        p.whole_ok = allow_single && p.whole_lds <= (size_t)max_lds;
        p.ti = p.tj = tile;
        p.nti = (l.ynum + 1 + tile - 1) / tile;
        p.ntj = (l.xnum + 1 + tile - 1) / tile;
        p.lds = tile_lds_bytes<T>(tile, tile);
        need = std::max(need, p.whole_ok ? std::max(p.lds, p.whole_lds) : p.lds);
        H.tiles.push_back(p);
    }
    H.fast_div = fast;
    if (need > 65536) {
        const void* k = fast ? (const void*)mg_tile_smooth_kernel<T, true>
                             : (const void*)mg_tile_smooth_kernel<T, false>;
        EBG_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    }
    return EBG_OK;
}

template <typename T>
void mg_smoother_tiled(MGHierarchy<T>& H, int k, int nsweeps, int local_sweeps, T theta_s,
                       T theta_c, T prnorm, const T* prnorm_dev, const T* bc_d, hipStream_t s)
{
    const bool normalize = (k > 0) || H.normalize_finest;
    MGLevelDev<T>& l = H.lev[k];
    const auto& p = H.tiles[k];
    const TileSpec ts{p.ti, p.tj, p.nti, p.ntj, 4};
    const int np = (l.ynum - 1) * (l.xnum - 1);
    auto launch = [&](unsigned nblocks, unsigned nthreads, size_t lds, const TileSpec& spec,
                      const TileSmoothArgs& a, T* ovx, T* ovy, T* opr, T* shift) {
        if (H.fast_div)
            hipLaunchKernelGGL((mg_tile_smooth_kernel<T, true>), dim3(nblocks), dim3(nthreads), lds, s, l,
                               spec, a, theta_s, theta_c, bc_d[0], bc_d[1], bc_d[2], bc_d[3], prnorm,
                               prnorm_dev, ovx, ovy, opr, shift);
        else
            hipLaunchKernelGGL((mg_tile_smooth_kernel<T, false>), dim3(nblocks), dim3(nthreads), lds, s, l,
                               spec, a, theta_s, theta_c, bc_d[0], bc_d[1], bc_d[2], bc_d[3], prnorm,
                               prnorm_dev, ovx, ovy, opr, shift);
    };
    const bool additive = H.additive_tiles && k > 0 && l.vx_alt;
    // one workgroup for the whole level (all sweeps, gauge shift and residuals
    // in one launch) where the level fits its LDS; with additive tiles only
    // for levels of at most 2 x 2 tiles, above which the all-tile launches
    // are quicker than one workgroup's serial sweeps
    if (p.whole_ok && (!additive || p.nti * p.ntj <= 4)) {
        const TileSpec whole{l.ynum + 1, l.xnum + 1, 1, 1, 1};
        launch(1, 1024, p.whole_lds, whole, TileSmoothArgs{0, nsweeps, 1, normalize ? 1 : 0}, l.vx, l.vy,
               l.pr, (T*)nullptr);
        return;
    }
    const int nthreads = std::min(1024, ((p.ti * p.tj + 63) / 64) * 64);
    for (int done = 0; done < nsweeps; done += local_sweeps) {
        const int m = std::min(local_sweeps, nsweeps - done);
        // the gauge shift prnorm - p(1,1) is taken by the block owning node
        // (1,1) in the last pass (no later launch of the pass writes it)
        T* shift = (normalize && done + m >= nsweeps) ? H.scratch : (T*)nullptr;
        if (additive) {
            // every tile in one launch, reading the previous iterate (halos
            // included) and writing the other buffer: deterministic
            const TileSpec all{p.ti, p.tj, p.nti, p.ntj, 1};
            launch((unsigned)(p.nti * p.ntj), nthreads, p.lds, all, TileSmoothArgs{0, m, 0, 0}, l.vx_alt,
                   l.vy_alt, l.pr_alt, shift);
            std::swap(l.vx, l.vx_alt);
            std::swap(l.vy, l.vy_alt);
            std::swap(l.pr, l.pr_alt);
            continue;
        }
        for (int c = 0; c < 4; ++c) {
            const int ci = c & 1, cj = c >> 1;
            const int ni = (p.nti - ci + 1) / 2, nj = (p.ntj - cj + 1) / 2;
            if (ni <= 0 || nj <= 0) continue;
            launch((unsigned)(ni * nj), nthreads, p.lds, ts, TileSmoothArgs{c, m, 0, 0}, l.vx, l.vy, l.pr,
                   c == 0 ? shift : (T*)nullptr);
        }
    }
    const unsigned nres = nblk((int64_t)(l.ynum + 1) * (l.xnum + 1));
    if (normalize) {
        // shift and residuals in one pass, the shifted pressures in the
        // second buffer
        hipLaunchKernelGGL(mg_shift_residual_kernel<T>, dim3(nres), dim3(256), 0, s, l, (const T*)H.scratch,
                           l.pr_alt);
        std::swap(l.pr, l.pr_alt);
        return;
    }
    hipLaunchKernelGGL(mg_residual_kernel<T>, dim3(nres), dim3(256), 0, s, l);
}

// Same level order and transfers as mg_vcycle; the smoothing calls use the
// tiled smoother.
template <typename T>
void mg_vcycle_tiled(MGHierarchy<T>& H, const std::vector<int>& sm, int local_sweeps, T theta_s,
                     T theta_c, T relax_v, T relax_p, T prnorm, const T* prnorm_dev, const T* bc_d,
                     hipStream_t s)
{
    const int nl = H.nlev();
    const bool direct = H.coarse_inv && nl > 1;   // exact solve replaces the coarsest smoothing
    // levels below the finest start from zero: the fused restriction into a
    // level clears its solution (level 0 is cleared by the caller)
    for (int n = 0; n < nl; ++n) {
        if (direct && n == nl - 1) {
            mg_coarse_solve(H, bc_d, s, true);
            break;
        }
        mg_smoother_tiled(H, n, sm[n], local_sweeps, theta_s, theta_c, prnorm, prnorm_dev, bc_d, s);
        if (n < nl - 1) {
            const MGLevelDev<T>& f = H.lev[n];
            const MGLevelDev<T>& c = H.lev[n + 1];
            const int64_t nall = (int64_t)(c.ynum + 1) * c.xnum + (int64_t)c.ynum * (c.xnum + 1) +
                                 (int64_t)(c.ynum - 1) * (c.xnum - 1);
            hipLaunchKernelGGL(mg_restrict_zero_kernel<T>, dim3(nblk(nall)), dim3(256), 0, s, f, c, H.xfer[n]);
        }
    }
    for (int n = nl - 1; n >= 0; --n) {
        if (n == 0) {
            mg_smoother_tiled(H, 0, sm[0], local_sweeps, theta_s, theta_c, prnorm, prnorm_dev, bc_d, s);
        } else {
            if (!(direct && n == nl - 1))
                mg_smoother_tiled(H, n, sm[n], local_sweeps, theta_s, theta_c, T(0), (const T*)nullptr, bc_d, s);
            const MGLevelDev<T>& c = H.lev[n];
            const MGLevelDev<T>& f = H.lev[n - 1];
            hipLaunchKernelGGL(mg_prolong_kernel<T>, dim3(nblk((int64_t)(f.ynum + 1) * (f.xnum + 1))),
                               dim3(256), 0, s, c, f, H.xfer[n - 1], relax_v, relax_p);
        }
    }
}

#define EBG_INST(T)                                                                              \
    template struct MGHierarchy<T>;                                                              \
    template int mg_build<T>(MGHierarchy<T>&, int, int, double, double, int,                     \
                             const std::vector<HostLevelGrid>*, bool);                           \
    template void mg_restrict_viscosity<T>(MGHierarchy<T>&, hipStream_t);                        \
    template void mg_smoother<T>(MGHierarchy<T>&, int, int, T, T, T, const T*, const T*,         \
                                 hipStream_t);                                                   \
    template void mg_vcycle<T>(MGHierarchy<T>&, const std::vector<int>&, T, T, T, T, T, const T*, \
                               const T*, hipStream_t);                                           \
    template void mg_zero_solution<T>(MGHierarchy<T>&, int, hipStream_t);                        \
    template int mg_plan_tiles<T>(MGHierarchy<T>&, int, bool, bool);                             \
    template void mg_smoother_tiled<T>(MGHierarchy<T>&, int, int, int, T, T, T, const T*, const T*, \
                                       hipStream_t);                                             \
    template void mg_vcycle_tiled<T>(MGHierarchy<T>&, const std::vector<int>&, int, T, T, T, T, T, \
                                     const T*, const T*, hipStream_t);
EBG_INST(float)
EBG_INST(double)

void ViscRamp::init(double vmin, double vmax, double njumps)
{
    min_o = vmin;
    max_o = vmax;
    min_cur = vmin;
    max_cur = vmin;
    min_factor = 1.0;
    max_factor = std::pow(10.0, std::log10(vmax / vmin) / njumps);
}

void ViscRamp::jump()
{
    min_cur = min_cur * min_factor;
    if (min_cur < min_o) min_cur = min_o;
    max_cur = max_cur * max_factor;
    if (max_cur > max_o) max_cur = max_o;
}

double ViscRamp::coef() const { return std::log(max_cur / min_cur) / std::log(max_o / min_o); }

// per-block minimum and maximum of a[0..n)
__global__ void __launch_bounds__(256) mg_minmax_partial_kernel(const double* __restrict__ a, int64_t n,
                                                                double* __restrict__ part)
{
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        lo = fmin(lo, a[t]);
        hi = fmax(hi, a[t]);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + off]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smin[0];
        part[2 * blockIdx.x + 1] = smax[0];
    }
}

int mg_device_range(const double* a, int64_t n, double* partial, hipStream_t s, double& lo, double& hi)
{
    double host[2 * MG_RUN_BLOCKS];
    hipLaunchKernelGGL(mg_minmax_partial_kernel, dim3(MG_RUN_BLOCKS), dim3(256), 0, s, a, n, partial);
    EBG_HIP(hipMemcpyAsync(host, partial, sizeof(host), hipMemcpyDeviceToHost, s));
    EBG_HIP(hipStreamSynchronize(s));
    for (int b = 0; b < MG_RUN_BLOCKS; ++b) {
        lo = std::min(lo, host[2 * b]);
        hi = std::max(hi, host[2 * b + 1]);
    }
    return EBG_OK;
}

// ViscosityScaling.scale_viscosity_for_all_levels! (ViscosityScaling.jl:41-79), one array
__global__ void __launch_bounds__(256) mg_visc_scale_kernel(const double* __restrict__ orig, double* __restrict__ eta,
                                                            int64_t n, double min_cur, double inv_min_o, double coef)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) eta[t] = min_cur * pow(orig[t] * inv_min_o, coef);
}

void mg_scale_viscosity(const double* orig, double* eta, int64_t n, double min_cur, double inv_min_o, double coef,
                        hipStream_t s)
{
    hipLaunchKernelGGL(mg_visc_scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, orig, eta, n,
                       min_cur, inv_min_o, coef);
}

double mg_max_principal_residual(const double* res, int n, int ivcycle)
{
    if (ivcycle <= 2) return 1e38;
    double m = 0.0;
    for (int q = 0; q < n; ++q) {
        const double v = std::fabs(std::pow(10.0, res[q]));
        if (std::isnan(v)) return NAN;
        m = std::max(m, v);
    }
    return m;
}

}  // namespace ebg
