// Marker -> grid interpolation of the Stokes transport properties.
//
// Replaces Interpolation.initialize_bilinear_interpolation! and
// interpolate_marker_stokes_parameters_to_grid! (interpolation/Interpolation.jl:61-95):
// marker cell indices (MarkerGridMapping.jl:99-173), weights
// (MarkerGridNodeWeights.jl:68-150), weight sums (BilinearDenominator.jl:34-182,
// BilinearNumerator.jl:205-454 with the rules of WeightFuncs.jl:67-190) and the
// averages with backup fallback (BilinearAverage.jl:57-178, BackupTransportArrays.jl).
//
// GPU formulation: markers are binned by basic cell with a stable radix sort of
// (cell, marker index) pairs; each grid node then gathers the markers of its
// neighbouring cells, merging the cell lists by marker index, so every node
// sum is accumulated in exactly the reference's serial marker order.  No
// atomics, deterministic, bit-identical to the serial loops (asin of the
// friction angle aside, which goes through the device math library).
#include "ebg_common.hpp"
#include "ebg_device.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>

namespace ebg {

struct ScatterState {
    int64_t cap = 0;
    int ncell = 0;
    uint32_t *keys = nullptr, *keys_s = nullptr, *vals = nullptr, *vals_s = nullptr;
    int *cstart = nullptr, *cend = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    double* wbuf = nullptr;  // internal weight arrays when the caller passes none
    int64_t wbuf_n = 0;
    ~ScatterState()
    {
        void* ps[] = {keys, keys_s, vals, vals_s, cstart, cend, tmp, wbuf};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
};

struct AxisD {
    const double* g;
    int n;
    double x0, inv_h;
    int uniform;
};

struct ScatterArgs {
    int64_t n;
    int xnum, ynum, ncell;
    const double *x, *y;
    const int8_t* flags;
    double xmin, xmax, ymin, ymax;
    AxisD ax, ay, avx;  // basic x, basic y, vy-grid x
    const double *xstp_b, *ystp_b, *gridx_vy, *xstp_vy;
    ebg_marker_stokes_props p;
    ebg_stokes_grid_fields f;
    const uint32_t* vals_s;
    const int *cstart, *cend;
    uint32_t *keys, *vals;
    double *wtnodes, *wtetas, *wtetan, *wtnodes_vy;
};

__device__ __forceinline__ bool inside_marker(const ScatterArgs& a, int64_t m, double xm, double ym)
{
    if (a.p.inside_flags) return a.p.inside_flags[m] == 1;
    return a.xmin <= xm && xm <= a.xmax && a.ymin <= ym && ym <= a.ymax;
}

__global__ void scatter_keys_kernel(ScatterArgs a)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.n) return;
    const double xm = a.x[m], ym = a.y[m];
    uint32_t key = (uint32_t)a.ncell;
    if (inside_marker(a, m, xm, ym)) {
        const int xn = left_node_index(xm, a.ax.g, a.ax.n, a.ax.x0, a.ax.inv_h, a.ax.uniform);
        const int yn = left_node_index(ym, a.ay.g, a.ay.n, a.ay.x0, a.ay.inv_h, a.ay.uniform);
        key = (uint32_t)((xn - 1) * (a.ynum - 1) + (yn - 1));
        if (a.f.marker_xn) a.f.marker_xn[m] = xn;
        if (a.f.marker_yn) a.f.marker_yn[m] = yn;
    }
    a.keys[m] = key;
    a.vals[m] = (uint32_t)m;
}

__global__ void cell_ranges_kernel(const uint32_t* __restrict__ keys, int64_t n, int ncell,
                                   int* __restrict__ cstart, int* __restrict__ cend)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t c = keys[k];
    if ((int)c >= ncell) return;
    if (k == 0 || keys[k - 1] != c) cstart[c] = (int)k;
    if (k == n - 1 || keys[k + 1] != c) cend[c] = (int)(k + 1);
}

__global__ void fill_int_kernel(int* a, int n, int v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// Merge-walk over up to NC cell lists in increasing marker index.
template <int NC>
struct Merger {
    int cur[NC], end[NC];
    uint32_t head[NC];
    __device__ void init(const ScatterArgs& a, const int* cells)
    {
        for (int q = 0; q < NC; ++q) {
            if (cells[q] >= 0) {
                cur[q] = a.cstart[cells[q]];
                end[q] = a.cend[cells[q]];
            } else {
                cur[q] = 0;
                end[q] = 0;
            }
            head[q] = cur[q] < end[q] ? a.vals_s[cur[q]] : 0xffffffffu;
        }
    }
    // returns the list index of the next marker or -1 when all lists are empty
    __device__ int next(const ScatterArgs& a, uint32_t* m)
    {
        int best = -1;
        uint32_t bv = 0xffffffffu;
        for (int q = 0; q < NC; ++q)
            if (head[q] < bv) {
                bv = head[q];
                best = q;
            }
        if (best < 0) return -1;
        *m = bv;
        const int c = ++cur[best];
        head[best] = c < end[best] ? a.vals_s[c] : 0xffffffffu;
        return best;
    }
};

__device__ __forceinline__ double getp(const double* p, uint32_t m) { return p ? p[m] : 0.0; }

// Basic nodes: inclusive (rho, wtnodes) and exclusive (wtetas and 7 fields).
__global__ void __launch_bounds__(256) scatter_basic_kernel(ScatterArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    if (t >= ny * nx) return;
    const int i = t % ny + 1, j = t / ny + 1;
    // cells whose marker feeds node (i,j): role UL (i,j), LL (i-1,j), UR (i,j-1), LR (i-1,j-1)
    int cells[4];
    const int ci[4] = {i, i - 1, i, i - 1}, cj[4] = {j, j, j - 1, j - 1};
    for (int q = 0; q < 4; ++q)
        cells[q] = (ci[q] >= 1 && ci[q] <= ny - 1 && cj[q] >= 1 && cj[q] <= nx - 1)
                       ? (cj[q] - 1) * (ny - 1) + (ci[q] - 1)
                       : -1;
    Merger<4> mg;
    mg.init(a, cells);
    double w_in = 0.0, w_ex = 0.0, s_rho = 0.0, s_coh = 0.0, s_fric = 0.0, s_ef = 0.0, s_eta = 0.0,
           s_mu = 0.0, s_sxy = 0.0, s_pl = 0.0;
    uint32_t m;
    int q;
    while ((q = mg.next(a, &m)) >= 0) {
        const int yn = ci[q], xn = cj[q];
        const double dx = (a.x[m] - a.ax.g[xn - 1]) / a.xstp_b[xn - 1];
        const double dy = (a.y[m] - a.ay.g[yn - 1]) / a.ystp_b[yn - 1];
        double w;
        bool ex;
        if (q == 0) { w = (1.0 - dx) * (1.0 - dy); ex = (dx <= 0.5 && dy <= 0.5); }
        else if (q == 1) { w = (1.0 - dx) * dy; ex = (dx <= 0.5 && 0.5 <= dy); }
        else if (q == 2) { w = dx * (1.0 - dy); ex = (dx >= 0.5 && 0.5 >= dy); }
        else { w = dx * dy; ex = (dx >= 0.5 && dy >= 0.5); }
        w_in += w * 1.0 * 1.0;
        s_rho += w * getp(a.p.rho, m) * 1.0;
        if (ex) {
            w_ex += w * 1.0 * 1.0;
            s_coh += w * getp(a.p.cohesion, m) * 1.0;
            const double fr = a.p.fric ? asin(a.p.fric[m]) * 180.0 / M_PI : 0.0;
            s_fric += w * fr * 1.0;
            s_ef += w * getp(a.p.eta_flow, m) * 1.0;
            s_eta += w * getp(a.p.eta, m) * 1.0;
            const double mu = a.p.mat_mu ? a.p.mat_mu[a.p.matid[m] - 1] : 1.0;
            s_mu += w * (1 / mu) * 1.0;
            s_sxy += w * getp(a.p.sxy, m) * 1.0;
            s_pl += w * (a.p.pfailure ? (double)a.p.pfailure[m] : 0.0) * 1.0;
        }
    }
    a.wtnodes[t] = w_in;
    a.wtetas[t] = w_ex;
    // BilinearAverage.jl:276-322 (old value where no weight)
    auto ar = [&](double* out, double s, double w) { if (out) out[t] = (w != 0) ? s / w : out[t]; };
    ar(a.f.rho1, s_rho, w_in);
    if (a.f.mus1) a.f.mus1[t] = (w_ex != 0) ? 1.0 / (s_mu / w_ex) : a.f.mus1[t];
    ar(a.f.etas1, s_eta, w_ex);
    ar(a.f.sxy1, s_sxy, w_ex);
    ar(a.f.plastics, s_pl, w_ex);
    ar(a.f.cohesion_grid, s_coh, w_ex);
    ar(a.f.fric_degrees_grid, s_fric, w_ex);
    ar(a.f.eta_flow, s_ef, w_ex);
}

// Pressure nodes: central rule from the node's own cell.
__global__ void __launch_bounds__(256) scatter_pressure_kernel(ScatterArgs a)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    if (c >= (ny - 1) * (nx - 1)) return;
    const int i = c % (ny - 1) + 1, j = c / (ny - 1) + 1;
    double w_c = 0.0, s_eta = 0.0, s_mu = 0.0, s_pl = 0.0, s_dil = 0.0, s_melt = 0.0, s_sxx = 0.0;
    const double gx = a.ax.g[j - 1], hx = a.xstp_b[j - 1], gy = a.ay.g[i - 1], hy = a.ystp_b[i - 1];
    for (int k = a.cstart[c]; k < a.cend[c]; ++k) {
        const uint32_t m = a.vals_s[k];
        const double dx = (a.x[m] - gx) / hx;
        const double dy = (a.y[m] - gy) / hy;
        const double w = (1.0 - fabs(0.5 - dx)) * (1.0 - fabs(0.5 - dy));
        w_c += w * 1.0 * 1.0;
        s_eta += w * getp(a.p.eta, m) * 1.0;
        const double mu = a.p.mat_mu ? a.p.mat_mu[a.p.matid[m] - 1] : 1.0;
        s_mu += w * (1 / mu) * 1.0;
        s_pl += w * (a.p.pfailure ? (double)a.p.pfailure[m] : 0.0) * 1.0;
        s_dil += w * (a.p.dilatation ? (double)a.p.dilatation[m] : 0.0) * 1.0;
        s_melt += w * getp(a.p.meltfrac, m) * 1.0;
        s_sxx += w * getp(a.p.sxx, m) * 1.0;
    }
    a.wtetan[c] = w_c;
    auto ar = [&](double* out, double s) { if (out) out[c] = (w_c != 0) ? s / w_c : out[c]; };
    if (a.f.mun1) a.f.mun1[c] = (w_c != 0) ? 1.0 / (s_mu / w_c) : a.f.mun1[c];
    ar(a.f.etan1, s_eta);
    ar(a.f.sxx1, s_sxx);
    ar(a.f.plasticn, s_pl);
    ar(a.f.dilatation_grid, s_dil);
    ar(a.f.extractable_meltfrac_grid, s_melt);
}

// vy nodes: inclusive rule on the vy grid (x remapped on gridx_vy, y basic).
__global__ void __launch_bounds__(256) scatter_vy_kernel(ScatterArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    if (t >= ny * (nx + 1)) return;
    const int i = t % ny + 1, jv = t / ny + 1;
    int cells[6], cyi[6];
    int q = 0;
    for (int dj = -2; dj <= 0; ++dj)
        for (int di = -1; di <= 0; ++di) {
            const int ci = i + di, cj = jv + dj;
            cyi[q] = ci;
            cells[q++] = (ci >= 1 && ci <= ny - 1 && cj >= 1 && cj <= nx - 1) ? (cj - 1) * (ny - 1) + (ci - 1) : -1;
        }
    Merger<6> mg;
    mg.init(a, cells);
    double w_s = 0.0, s_rho = 0.0;
    uint32_t m;
    int l;
    while ((l = mg.next(a, &m)) >= 0) {
        const double xm = a.x[m], ym = a.y[m];
        const int xnv = left_node_index(xm, a.avx.g, a.avx.n, a.avx.x0, a.avx.inv_h, a.avx.uniform);
        if (xnv != jv && xnv != jv - 1) continue;
        const int yn = cyi[l];
        const double dx = (xm - a.gridx_vy[xnv - 1]) / a.xstp_vy[xnv - 1];
        const double dy = (ym - a.ay.g[yn - 1]) / a.ystp_b[yn - 1];
        double w;
        if (yn == i && xnv == jv) w = (1.0 - dx) * (1.0 - dy);
        else if (yn == i - 1 && xnv == jv) w = (1.0 - dx) * dy;
        else if (yn == i && xnv == jv - 1) w = dx * (1.0 - dy);
        else w = dx * dy;
        w_s += w * 1.0 * 1.0;
        s_rho += w * getp(a.p.rho, m) * 1.0;
    }
    a.wtnodes_vy[t] = w_s;
    if (a.f.rho1_vy) a.f.rho1_vy[t] = (w_s != 0) ? s_rho / w_s : a.f.rho1_vy[t];
}

static int axis_uniform_h(const std::vector<double>& g, double* x0, double* inv_h)
{
    const int64_t n = (int64_t)g.size();
    const double h = (g[n - 1] - g[0]) / (double)(n - 1);
    *x0 = g[0];
    *inv_h = 1.0 / h;
    if (!(h > 0)) return 0;
    for (int64_t k = 0; k < n; ++k)
        if (std::fabs(g[k] - (g[0] + (double)k * h)) > 0.25 * h) return 0;
    return 1;
}

static unsigned nbk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace ebg

using namespace ebg;

extern "C" {

int ebg_scatter_release(ebg_ctx* ctx)
{
    if (ctx && ctx->scatter) {
        delete (ScatterState*)ctx->scatter;
        ctx->scatter = nullptr;
    }
    return EBG_OK;
}

int ebg_interpolate_marker_stokes_parameters_to_grid(ebg_ctx* ctx, const ebg_marker_stokes_props* props,
                                                     ebg_stokes_grid_fields* fields)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_interpolate_marker_stokes_parameters_to_grid: no grid");
    EBG_CHECK(props && fields, EBG_ERR_ARG, "ebg_interpolate_marker_stokes_parameters_to_grid: NULL arguments");
    const Grid2D& g = ctx->grid;
    const int64_t n = props->marknum;
    EBG_CHECK(n >= 0 && n < (int64_t)0xffffffffLL, EBG_ERR_ARG, "scatter: marknum out of range");
    EBG_CHECK(n == 0 || (props->x && props->y), EBG_ERR_ARG, "scatter: NULL marker coordinates");
    EBG_CHECK((props->mat_mu == nullptr) == (props->matid == nullptr), EBG_ERR_ARG,
              "scatter: matid and mat_mu must be given together");
    const int nx = (int)g.xnum, ny = (int)g.ynum;
    const int ncell = (nx - 1) * (ny - 1);
    if (!ctx->scatter) ctx->scatter = new ScatterState();
    ScatterState& S = *(ScatterState*)ctx->scatter;
    hipStream_t s = ctx->stream;
    if (props->mat_mu) {
        // check_shear_modulus (BilinearNumerator.jl:155-160): zero modulus is an error
        EBG_CHECK(props->nmat >= 1, EBG_ERR_ARG, "scatter: nmat must be >= 1");
        std::vector<double> mu((size_t)props->nmat);
        EBG_HIP(hipMemcpyAsync(mu.data(), props->mat_mu, mu.size() * 8, hipMemcpyDeviceToHost, s));
        EBG_HIP(hipStreamSynchronize(s));
        for (size_t k = 0; k < mu.size(); ++k)
            EBG_CHECK(mu[k] != 0.0, EBG_ERR_ARG, "scatter: shear modulus of a material is zero");
    }
    if (S.cap < n || S.ncell != ncell) {
        ScatterState fresh;
        std::swap(S.keys, fresh.keys);
        std::swap(S.keys_s, fresh.keys_s);
        std::swap(S.vals, fresh.vals);
        std::swap(S.vals_s, fresh.vals_s);
        std::swap(S.cstart, fresh.cstart);
        std::swap(S.cend, fresh.cend);
        std::swap(S.tmp, fresh.tmp);
        S.tmp_bytes = 0;
        const size_t cap = (size_t)std::max<int64_t>(n, 1);
        EBG_HIP(hipMalloc(&S.keys, cap * 4));
        EBG_HIP(hipMalloc(&S.keys_s, cap * 4));
        EBG_HIP(hipMalloc(&S.vals, cap * 4));
        EBG_HIP(hipMalloc(&S.vals_s, cap * 4));
        EBG_HIP(hipMalloc(&S.cstart, (size_t)ncell * 4));
        EBG_HIP(hipMalloc(&S.cend, (size_t)ncell * 4));
        S.cap = (int64_t)cap;
        S.ncell = ncell;
    }
    const int64_t nb = (int64_t)nx * ny, np = (int64_t)ncell, nvy = (int64_t)ny * (nx + 1);
    double *wtn = fields->wtnodes, *wts = fields->wtetas, *wte = fields->wtetan, *wtv = fields->wtnodes_vy;
    if (!wtn || !wts || !wte || !wtv) {
        const int64_t need = 2 * nb + np + nvy;
        if (S.wbuf_n < need) {
            if (S.wbuf) (void)hipFree(S.wbuf);
            S.wbuf = nullptr;
            EBG_HIP(hipMalloc(&S.wbuf, need * 8));
            S.wbuf_n = need;
        }
        if (!wtn) wtn = S.wbuf;
        if (!wts) wts = S.wbuf + nb;
        if (!wte) wte = S.wbuf + 2 * nb;
        if (!wtv) wtv = S.wbuf + 2 * nb + np;
    }
    ScatterArgs a{};
    a.n = n;
    a.xnum = nx;
    a.ynum = ny;
    a.ncell = ncell;
    a.x = props->x;
    a.y = props->y;
    a.xmin = g.xmin;
    a.xmax = g.xmax;
    a.ymin = g.ymin;
    a.ymax = g.ymax;
    a.ax = AxisD{g.gridx_b, nx, g.x0, g.inv_hx, g.uniform_x};
    a.ay = AxisD{g.gridy_b, ny, g.y0, g.inv_hy, g.uniform_y};
    double vx0, vinv;
    const int vuni = axis_uniform_h(g.h_gridx_vy, &vx0, &vinv);
    a.avx = AxisD{g.gridx_vy, nx + 1, vx0, vinv, vuni};
    a.xstp_b = g.xstp_b;
    a.ystp_b = g.ystp_b;
    a.gridx_vy = g.gridx_vy;
    a.xstp_vy = g.xstp_vy;
    a.p = *props;
    a.f = *fields;
    a.keys = S.keys;
    a.vals = S.vals;
    a.vals_s = S.vals_s;
    a.cstart = S.cstart;
    a.cend = S.cend;
    a.wtnodes = wtn;
    a.wtetas = wts;
    a.wtetan = wte;
    a.wtnodes_vy = wtv;
    hipLaunchKernelGGL(fill_int_kernel, dim3(nbk(ncell)), dim3(256), 0, s, S.cstart, ncell, 0);
    hipLaunchKernelGGL(fill_int_kernel, dim3(nbk(ncell)), dim3(256), 0, s, S.cend, ncell, 0);
    if (n > 0) {
        hipLaunchKernelGGL(scatter_keys_kernel, dim3(nbk(n)), dim3(256), 0, s, a);
        unsigned end_bit = 1;
        while ((1ull << end_bit) <= (unsigned long long)ncell) ++end_bit;
        size_t bytes = 0;
        EBG_HIP(rocprim::radix_sort_pairs(nullptr, bytes, S.keys, S.keys_s, S.vals, S.vals_s,
                                          (unsigned int)n, 0, end_bit, s));
        if (bytes > S.tmp_bytes) {
            if (S.tmp) (void)hipFree(S.tmp);
            S.tmp = nullptr;
            EBG_HIP(hipMalloc(&S.tmp, bytes));
            S.tmp_bytes = bytes;
        }
        EBG_HIP(rocprim::radix_sort_pairs(S.tmp, bytes, S.keys, S.keys_s, S.vals, S.vals_s,
                                          (unsigned int)n, 0, end_bit, s));
        hipLaunchKernelGGL(cell_ranges_kernel, dim3(nbk(n)), dim3(256), 0, s, S.keys_s, n, ncell,
                           S.cstart, S.cend);
    }
    hipLaunchKernelGGL(scatter_basic_kernel, dim3(nbk(nb)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(scatter_pressure_kernel, dim3(nbk(np)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(scatter_vy_kernel, dim3(nbk(nvy)), dim3(256), 0, s, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

}  // extern "C"
