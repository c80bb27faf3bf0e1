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
// (cell, marker index) pairs.  One pass then gathers every marker once into
// per-cell lists laid out in 64-cell slices (entry t of 64 neighbouring cells
// in adjacent words), holding its normalized distances and the property
// values the requested outputs consume.  Each grid node walks the lists of
// its neighbouring cells, merging them by marker index, so every node sum is
// accumulated in exactly the reference's serial marker order.  No atomics,
// deterministic, bit-identical to the serial loops (asin of the friction
// angle aside, which goes through the device math library).
#include "ebg_common.hpp"
#include "ebg_device.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>

namespace ebg {

// Cell lists in a sliced layout: cells are taken in groups of 64 consecutive
// cell indices (one wavefront of node lanes touches consecutive cells), and
// group b owns rows rowbase[b] .. rowbase[b] + nrow[b] - 1, nrow[b] being the
// largest marker count of its cells.  Entry t of cell c (its t-th marker in
// increasing marker index) sits at (rowbase[c/64] + t) * 64 + c % 64, so the
// lanes of a wave reading entry t of neighbouring cells read adjacent words.
constexpr int CGROUP = 64;

struct ScatterState {
    int64_t cap = 0;
    int ncell = 0;
    uint32_t *keys = nullptr, *keys_s = nullptr, *vals = nullptr, *vals_s = nullptr;
    int *cstart = nullptr, *cend = nullptr;
    int* rowbase = nullptr;  // ngroup + 1 (exclusive scan of the group row counts)
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    double* wbuf = nullptr;  // internal weight arrays when the caller passes none
    int64_t wbuf_n = 0;
    // per-entry columns: marker index, vy-grid column, and the FP64 slots
    uint32_t* emark = nullptr;
    int* exnv = nullptr;
    double* eval = nullptr;
    int64_t ecap = 0, evcap = 0;
    int* h_rows = nullptr;  // pinned: total row count read back per call
    // cell-ordered mode: list offsets (ncell + 1), scan block totals, partial planes
    int order = EBG_SCATTER_SERIAL_ORDER;
    int* start = nullptr;     // ncell + 1: overflow offsets
    int* bsum = nullptr;
    uint32_t* sliced = nullptr;   // 64 slots per cell, in groups of 64 cells
    int start_n = 0;
    double* inv_mu = nullptr;     // 1 / mat_mu
    int64_t inv_mu_n = 0;
    uint32_t* pend = nullptr;     // queued overflow slots (marker, cell, slot)
    int64_t pend_n = 0;
    double* part = nullptr;
    int64_t part_n = 0;
    ~ScatterState()
    {
        void* ps[] = {keys, keys_s, vals, vals_s, cstart, cend, rowbase, tmp, wbuf, emark, exnv, eval,
                      start, bsum, sliced, part, inv_mu, pend};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        if (h_rows) (void)hipHostFree(h_rows);
    }
};

struct AxisD {
    const double* g;
    int n;
    double x0, inv_h;
    int uniform;
};

// FP64 slots of an entry: 0 dx, 1 dy (normalized distances from the
// upper-left basic node of the marker's cell, MarkerGridMapping.jl:185-226),
// 2 dx on the vy grid (MarkerGridMapping.jl:228-265), then the marker values
// that some requested output consumes, at the slots in Layout::off (-1: not
// stored).  Values are stored as the reference's numerator loops use them:
// the friction angle in degrees and the reciprocal shear modulus
// (BilinearNumerator.jl:296-388).
enum { F_RHO, F_COH, F_FRIC, F_EFLOW, F_ETA, F_INVMU, F_SXY, F_PL, F_DIL, F_MELT, F_SXX, NSTOKES };
// heat-transport slots (BilinearNumerator.jl:31-141): temperature,
// conductivity, rho*Cp, heat production (material value plus the optional
// serpentinization term) and the adiabatic term
enum { F_TK = NSTOKES, F_KT, F_RHOCP, F_HR, F_HA, NFIELD };

struct Layout {
    int nslot;
    int off[NFIELD];
};

struct ScatterArgs {
    int64_t n;
    int xnum, ynum, ncell;
    const double *x, *y;
    const int8_t* flags;
    double xmin, xmax, ymin, ymax;
    AxisD ax, ay, avx;  // basic x, basic y, vy-grid x
    const double *xstp_b, *ystp_b, *gridx_vy, *xstp_vy;
    const int8_t* inside_flags;
    int32_t *out_xn, *out_yn;
    ebg_marker_stokes_props p;
    ebg_stokes_grid_fields f;
    ebg_marker_heat_props h;
    ebg_heat_grid_fields hf;
    const uint32_t *keys_s, *vals_s;
    const int *cstart, *cend;
    const int* rowbase;
    uint32_t *keys, *vals;
    uint32_t* emark;
    int* exnv;
    double* eval;
    int64_t slot_stride;  // entries per FP64 slot (total rows * 64)
    Layout L;
    double *wtnodes, *wtetas, *wtetan, *wtnodes_vy;
};

__device__ __forceinline__ bool inside_marker(const ScatterArgs& a, int64_t m, double xm, double ym)
{
    if (a.inside_flags) return a.inside_flags[m] == 1;
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
        if (a.out_xn) a.out_xn[m] = xn;
        if (a.out_yn) a.out_yn[m] = yn;
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

// Row count of each 64-cell group (its largest cell list), written to
// rows[b + 1] so that the inclusive scan below leaves rowbase[b] in place.
__global__ void __launch_bounds__(256) group_rows_kernel(const int* __restrict__ cstart,
                                                         const int* __restrict__ cend, int ncell,
                                                         int ngroup, int* __restrict__ rows)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int cnt = c < ncell ? cend[c] - cstart[c] : 0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) cnt = max(cnt, __shfl_xor(cnt, d, 64));
    const int b = c / CGROUP;
    if ((c & (CGROUP - 1)) == 0 && b < ngroup) rows[b + 1] = cnt;
    if (c == 0) rows[0] = 0;
}

// In-place inclusive scan of rows[0..n] by one workgroup of 1024 threads.
__global__ void __launch_bounds__(1024) scan_rows_kernel(int* __restrict__ rows, int n)
{
    __shared__ int part[1024];
    const int tid = threadIdx.x;
    const int per = (n + 1 + 1023) / 1024;
    const int lo = tid * per, hi = min(lo + per, n + 1);
    int sum = 0;
    for (int k = lo; k < hi; ++k) sum += rows[k];
    part[tid] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int v = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = tid ? part[tid - 1] : 0;
    for (int k = lo; k < hi; ++k) {
        run += rows[k];
        rows[k] = run;
    }
}

// Fill the sliced cell lists: lane = cell, loop over its entries in order,
// four at a time so that the gathers of four markers are in flight together.
// Each marker is gathered and divided once; the writes of a wave are rows of
// 64 adjacent words.
constexpr int EBATCH = 4;

template <bool HEAT>
__global__ void __launch_bounds__(256) scatter_entries_kernel(ScatterArgs a)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = c / CGROUP, lane = c & (CGROUP - 1);
    if (c >= a.ncell) return;
    const int ks = a.cstart[c];
    const int cnt = a.cend[c] - ks;
    if (cnt == 0) return;
    const int r0 = a.rowbase[b];
    const int cj = c / (a.ynum - 1), ci = c - cj * (a.ynum - 1);
    const double gx = a.ax.g[cj], hx = a.xstp_b[cj], gy = a.ay.g[ci], hy = a.ystp_b[ci];
    const int* o = a.L.off;
    const int64_t S = a.slot_stride;
    const uint32_t* __restrict__ vs = a.vals_s + ks;
    const double* __restrict__ X = a.x;
    const double* __restrict__ Y = a.y;
    for (int t0 = 0; t0 < cnt; t0 += EBATCH) {
        uint32_t m[EBATCH];
        double xm[EBATCH], ym[EBATCH];
#pragma unroll
        for (int u = 0; u < EBATCH; ++u) m[u] = vs[min(t0 + u, cnt - 1)];
#pragma unroll
        for (int u = 0; u < EBATCH; ++u) {
            xm[u] = X[m[u]];
            ym[u] = Y[m[u]];
        }
        double pv[NFIELD][EBATCH];
        if (HEAT) {
#pragma unroll
            for (int u = 0; u < EBATCH; ++u) {
                const uint32_t mu_ = m[u];
                pv[F_TK][u] = o[F_TK] >= 0 ? a.h.TK[mu_] : 0.0;
                pv[F_KT][u] = o[F_KT] >= 0 ? a.h.kt[mu_] : 0.0;
                pv[F_RHOCP][u] = o[F_RHOCP] >= 0 ? a.h.rhocp[mu_] : 0.0;
                pv[F_HA][u] = o[F_HA] >= 0 ? a.h.ha[mu_] : 0.0;
                double hr = 0.0;
                if (o[F_HR] >= 0) {
                    const int id = a.h.matid[mu_];
                    hr = (id >= 1 && id <= a.h.nmat) ? a.h.mat_hr[id - 1] : __builtin_nan("");
                    if (a.h.serpentinization_heat_production)
                        hr = hr + a.h.serpentinization_heat_production[mu_];
                }
                pv[F_HR][u] = hr;
            }
        }
#pragma unroll
        for (int u = 0; u < EBATCH && !HEAT; ++u) {
            const uint32_t mu_ = m[u];
            pv[F_RHO][u] = o[F_RHO] >= 0 ? a.p.rho[mu_] : 0.0;
            pv[F_COH][u] = o[F_COH] >= 0 ? a.p.cohesion[mu_] : 0.0;
            pv[F_FRIC][u] = o[F_FRIC] >= 0 ? a.p.fric[mu_] : 0.0;
            pv[F_EFLOW][u] = o[F_EFLOW] >= 0 ? a.p.eta_flow[mu_] : 0.0;
            pv[F_ETA][u] = o[F_ETA] >= 0 ? a.p.eta[mu_] : 0.0;
            pv[F_INVMU][u] = (o[F_INVMU] >= 0 && a.p.mat_mu) ? a.p.mat_mu[a.p.matid[mu_] - 1] : 1.0;
            pv[F_SXY][u] = o[F_SXY] >= 0 ? a.p.sxy[mu_] : 0.0;
            pv[F_PL][u] = o[F_PL] >= 0 ? (double)a.p.pfailure[mu_] : 0.0;
            pv[F_DIL][u] = o[F_DIL] >= 0 ? (double)a.p.dilatation[mu_] : 0.0;
            pv[F_MELT][u] = o[F_MELT] >= 0 ? a.p.meltfrac[mu_] : 0.0;
            pv[F_SXX][u] = o[F_SXX] >= 0 ? a.p.sxx[mu_] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < EBATCH; ++u) {
            if (t0 + u >= cnt) break;
            const int64_t e = (int64_t)(r0 + t0 + u) * CGROUP + lane;
            a.emark[e] = m[u];
            double* v = a.eval + e;
            v[0] = (xm[u] - gx) / hx;
            v[S] = (ym[u] - gy) / hy;
            if (HEAT) {  // basic-grid nodes only: no vy-grid column
#pragma unroll
                for (int q = F_TK; q < NFIELD; ++q)
                    if (o[q] >= 0) v[o[q] * S] = pv[q][u];
                continue;
            }
            const int xnv = left_node_index(xm[u], a.avx.g, a.avx.n, a.avx.x0, a.avx.inv_h, a.avx.uniform);
            a.exnv[e] = xnv;
            v[2 * S] = (xm[u] - a.gridx_vy[xnv - 1]) / a.xstp_vy[xnv - 1];
            if (o[F_RHO] >= 0) v[o[F_RHO] * S] = pv[F_RHO][u];
            if (o[F_COH] >= 0) v[o[F_COH] * S] = pv[F_COH][u];
            if (o[F_FRIC] >= 0) v[o[F_FRIC] * S] = asin(pv[F_FRIC][u]) * 180.0 / M_PI;
            if (o[F_EFLOW] >= 0) v[o[F_EFLOW] * S] = pv[F_EFLOW][u];
            if (o[F_ETA] >= 0) v[o[F_ETA] * S] = pv[F_ETA][u];
            if (o[F_INVMU] >= 0) v[o[F_INVMU] * S] = 1 / pv[F_INVMU][u];
            if (o[F_SXY] >= 0) v[o[F_SXY] * S] = pv[F_SXY][u];
            if (o[F_PL] >= 0) v[o[F_PL] * S] = pv[F_PL][u];
            if (o[F_DIL] >= 0) v[o[F_DIL] * S] = pv[F_DIL][u];
            if (o[F_MELT] >= 0) v[o[F_MELT] * S] = pv[F_MELT][u];
            if (o[F_SXX] >= 0) v[o[F_SXX] * S] = pv[F_SXX][u];
        }
    }
}

// First entry and one-past-last entry of cell c's sliced list (stride 64).
__device__ __forceinline__ void cell_list(const ScatterArgs& a, int c, int64_t* first, int64_t* last)
{
    if (c < 0) {
        *first = *last = 0;
        return;
    }
    *first = (int64_t)a.rowbase[c / CGROUP] * CGROUP + (c & (CGROUP - 1));
    *last = *first + (int64_t)(a.cend[c] - a.cstart[c]) * CGROUP;
}

// Merge-walk over K cell lists (each in increasing marker index) in
// increasing marker index: the serial loops' order restricted to the node's
// markers.  Unrolled over K with constant indices, so the state stays in
// registers.
template <int K>
struct Merge {
    int64_t cur[K], end[K];
    uint32_t head[K], nxt[K];  // marker index at cur and at cur + 64 (prefetched)
    __device__ __forceinline__ void init(const ScatterArgs& a, const int (&cells)[K])
    {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            cell_list(a, cells[q], &cur[q], &end[q]);
            head[q] = cur[q] < end[q] ? a.emark[cur[q]] : 0xffffffffu;
            nxt[q] = cur[q] + CGROUP < end[q] ? a.emark[cur[q] + CGROUP] : 0xffffffffu;
        }
    }
    // entry of the next marker and its list, or false when all are done
    __device__ __forceinline__ bool next(const ScatterArgs& a, int64_t* pos, int* list)
    {
        uint32_t mn = head[0];
        int q = 0;
#pragma unroll
        for (int t = 1; t < K; ++t)
            if (head[t] < mn) {
                mn = head[t];
                q = t;
            }
        if (mn == 0xffffffffu) return false;
        int64_t k = cur[0], e = end[0];
#pragma unroll
        for (int t = 1; t < K; ++t) {
            k = q == t ? cur[t] : k;
            e = q == t ? end[t] : e;
        }
        // the list's new head was prefetched; fetch the one after it
        const int64_t kf = k + 2 * CGROUP;
        const uint32_t hf = kf < e ? a.emark[kf] : 0xffffffffu;
#pragma unroll
        for (int t = 0; t < K; ++t)
            if (q == t) {
                cur[t] = k + CGROUP;
                head[t] = nxt[t];
                nxt[t] = hf;
            }
        *pos = k;
        *list = q;
        return true;
    }
};

// Basic nodes: inclusive (rho, wtnodes) and exclusive (wtetas and 7 fields).
__global__ void __launch_bounds__(256) scatter_basic_kernel(ScatterArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    if (t >= ny * nx) return;
    const int i = t % ny + 1, j = t / ny + 1;
    // cells whose markers feed node (i,j), by the node's role in the cell:
    // 0 upper-left of (i,j), 1 lower-left of (i-1,j), 2 upper-right of (i,j-1),
    // 3 lower-right of (i-1,j-1)
    const int ci[4] = {i, i - 1, i, i - 1}, cj[4] = {j, j, j - 1, j - 1};
    int cells[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        cells[q] = (ci[q] >= 1 && ci[q] <= ny - 1 && cj[q] >= 1 && cj[q] <= nx - 1)
                       ? (cj[q] - 1) * (ny - 1) + (ci[q] - 1)
                       : -1;
    Merge<4> mg;
    mg.init(a, cells);
    const int* o = a.L.off;
    const int64_t S = a.slot_stride;
    double w_in = 0.0, w_ex = 0.0, s_rho = 0.0, s_coh = 0.0, s_fric = 0.0, s_ef = 0.0, s_eta = 0.0,
           s_mu = 0.0, s_sxy = 0.0, s_pl = 0.0;
    int64_t e;
    int q;
    while (mg.next(a, &e, &q)) {
        const double* v = a.eval + e;
        const double dx = v[0], dy = v[S];
        // WeightFuncs.jl:67-140: (1-dx)(1-dy), (1-dx)dy, dx(1-dy), dx dy by role
        const double w = ((q & 2) ? dx : (1.0 - dx)) * ((q & 1) ? dy : (1.0 - dy));
        const bool ex = ((q & 2) ? (dx >= 0.5) : (dx <= 0.5)) && ((q & 1) ? (dy >= 0.5) : (dy <= 0.5));
        w_in += w * 1.0 * 1.0;
        if (o[F_RHO] >= 0) s_rho += w * v[o[F_RHO] * S] * 1.0;
        if (ex) {
            w_ex += w * 1.0 * 1.0;
            if (o[F_COH] >= 0) s_coh += w * v[o[F_COH] * S] * 1.0;
            if (o[F_FRIC] >= 0) s_fric += w * v[o[F_FRIC] * S] * 1.0;
            if (o[F_EFLOW] >= 0) s_ef += w * v[o[F_EFLOW] * S] * 1.0;
            if (o[F_ETA] >= 0) s_eta += w * v[o[F_ETA] * S] * 1.0;
            if (o[F_INVMU] >= 0) s_mu += w * v[o[F_INVMU] * S] * 1.0;
            if (o[F_SXY] >= 0) s_sxy += w * v[o[F_SXY] * S] * 1.0;
            if (o[F_PL] >= 0) s_pl += w * v[o[F_PL] * S] * 1.0;
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
    const int* o = a.L.off;
    const int64_t S = a.slot_stride;
    double w_c = 0.0, s_eta = 0.0, s_mu = 0.0, s_pl = 0.0, s_dil = 0.0, s_melt = 0.0, s_sxx = 0.0;
    int64_t e, ee;
    cell_list(a, c, &e, &ee);
    for (; e < ee; e += CGROUP) {
        const double* v = a.eval + e;
        const double dx = v[0], dy = v[S];
        const double w = (1.0 - fabs(0.5 - dx)) * (1.0 - fabs(0.5 - dy));
        w_c += w * 1.0 * 1.0;
        if (o[F_ETA] >= 0) s_eta += w * v[o[F_ETA] * S] * 1.0;
        if (o[F_INVMU] >= 0) s_mu += w * v[o[F_INVMU] * S] * 1.0;
        if (o[F_PL] >= 0) s_pl += w * v[o[F_PL] * S] * 1.0;
        if (o[F_DIL] >= 0) s_dil += w * v[o[F_DIL] * S] * 1.0;
        if (o[F_MELT] >= 0) s_melt += w * v[o[F_MELT] * S] * 1.0;
        if (o[F_SXX] >= 0) s_sxx += w * v[o[F_SXX] * S] * 1.0;
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
// Node (i,jv) takes the markers of basic cells rows i-1..i, columns jv-2..jv
// whose vy-grid column is jv-1 or jv.
__global__ void __launch_bounds__(256) scatter_vy_kernel(ScatterArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    if (t >= ny * (nx + 1)) return;
    const int i = t % ny + 1, jv = t / ny + 1;
    int cells[6];
#pragma unroll
    for (int l = 0; l < 6; ++l) {
        const int ci = i - 1 + (l & 1), cj = jv - 2 + (l >> 1);
        cells[l] = (ci >= 1 && ci <= ny - 1 && cj >= 1 && cj <= nx - 1) ? (cj - 1) * (ny - 1) + (ci - 1) : -1;
    }
    Merge<6> mg;
    mg.init(a, cells);
    const int o_rho = a.L.off[F_RHO];
    const int64_t S = a.slot_stride;
    double w_s = 0.0, s_rho = 0.0;
    int64_t e;
    int l;
    while (mg.next(a, &e, &l)) {
        const int xnv = a.exnv[e];
        if (xnv != jv && xnv != jv - 1) continue;
        const double* v = a.eval + e;
        const double dx = v[2 * S], dy = v[S];
        // odd l: the marker's cell row is i, even l: i-1
        const double w = ((xnv == jv) ? (1.0 - dx) : dx) * ((l & 1) ? (1.0 - dy) : dy);
        w_s += w * 1.0 * 1.0;
        if (o_rho >= 0) s_rho += w * v[o_rho * S] * 1.0;
    }
    a.wtnodes_vy[t] = w_s;
    if (a.f.rho1_vy) a.f.rho1_vy[t] = (w_s != 0) ? s_rho / w_s : a.f.rho1_vy[t];
}

// Basic nodes, heat transport properties: the inclusive rule only
// (BilinearNumerator.jl:31-141, BilinearAverage.jl:10-50).
__global__ void __launch_bounds__(256) scatter_heat_kernel(ScatterArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    if (t >= ny * nx) return;
    const int i = t % ny + 1, j = t / ny + 1;
    const int ci[4] = {i, i - 1, i, i - 1}, cj[4] = {j, j, j - 1, j - 1};
    int cells[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        cells[q] = (ci[q] >= 1 && ci[q] <= ny - 1 && cj[q] >= 1 && cj[q] <= nx - 1)
                       ? (cj[q] - 1) * (ny - 1) + (ci[q] - 1)
                       : -1;
    Merge<4> mg;
    mg.init(a, cells);
    const int* o = a.L.off;
    const int64_t S = a.slot_stride;
    double w_in = 0.0, s_tk = 0.0, s_kt = 0.0, s_rc = 0.0, s_hr = 0.0, s_ha = 0.0;
    int64_t e;
    int q;
    while (mg.next(a, &e, &q)) {
        const double* v = a.eval + e;
        const double dx = v[0], dy = v[S];
        const double w = ((q & 2) ? dx : (1.0 - dx)) * ((q & 1) ? dy : (1.0 - dy));
        w_in += w * 1.0 * 1.0;
        if (o[F_TK] >= 0) s_tk += w * v[o[F_TK] * S] * 1.0;
        if (o[F_KT] >= 0) s_kt += w * v[o[F_KT] * S] * 1.0;
        if (o[F_RHOCP] >= 0) s_rc += w * v[o[F_RHOCP] * S] * 1.0;
        if (o[F_HR] >= 0) s_hr += w * v[o[F_HR] * S] * 1.0;
        if (o[F_HA] >= 0) s_ha += w * v[o[F_HA] * S] * 1.0;
    }
    if (a.wtnodes) a.wtnodes[t] = w_in;
    auto ar = [&](double* out, double s) { if (out) out[t] = (w_in != 0) ? s / w_in : out[t]; };
    ar(a.hf.tk1, s_tk);
    ar(a.hf.kt1, s_kt);
    ar(a.hf.rhocp1, s_rc);
    ar(a.hf.hr1, s_hr);
    ar(a.hf.ha1, s_ha);
}

// ---------------------------------------------------------------------------
// Cell-ordered scatter (ebg_scatter_set_order(ctx, EBG_SCATTER_CELL_ORDER)).
//
// Each cell's markers are summed privately, in increasing marker index, into
// the cell's per-corner partial sums (inclusive and exclusive basic-node
// weights and numerators, the vy-node terms of the cell's 3 x 2 vy nodes),
// held in registers; a node's sum is then the fixed-order sum of its
// neighbouring cells' partials (4 for basic nodes, 6 for vy nodes).  No
// floating-point atomics anywhere, so results are deterministic; they differ
// from the serial marker order only in the grouping of the additions (a few
// ulps of the sum of the absolute terms).  Pressure nodes take markers of one
// cell only, so their sums keep the serial order and stay bit-identical.
//
// Binning is a counting sort by cell: one pass computes each inside marker's
// cell (bit-exact) and takes its slot in the cell's list from an integer
// atomic counter (one atomic per run of equal cells among consecutive lanes),
// and a second pass writes the marker indices into sliced lists: slot t of
// the 64 cells of group g at (64 g + t) * 64 + lane, for t < 64, so the lanes
// of a wavefront (one cell each) read and write adjacent words.  The slots
// of cells with more than 64 markers continue in an overflow area, contiguous
// per cell at offsets from a scan of those cells' counts.  Slots follow the
// atomics' arrival order, so every list is then sorted by marker index (a
// bitonic network in registers for up to 64 markers, register runs merged in
// place beyond that).

// slot t < 64 of cell c in the sliced lists
__device__ __forceinline__ int64_t sliced_slot(int c, int t)
{
    return ((int64_t)(c >> 6) << 12) + ((int64_t)t << 6) + (c & 63);
}

// Marker cells and list slots: each inside marker takes its slot in its
// cell's list from an integer atomic counter (consecutive lanes with the same
// cell share one atomic add of the run length) and writes its index into the
// sliced list; a slot beyond 64 is queued (marker, cell, slot) for the
// overflow pass, which needs the final counts.
__global__ void __launch_bounds__(256) bin_slots_kernel(ScatterArgs a, int* __restrict__ count,
                                                        uint32_t* __restrict__ sliced, uint32_t* __restrict__ pending,
                                                        int* __restrict__ npending)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    uint32_t key = 0xffffffffu;
    if (m < a.n) {
        const double xm = a.x[m], ym = a.y[m];
        if (inside_marker(a, m, xm, ym)) {
            const int xn = left_node_index(xm, a.ax.g, a.ax.n, a.ax.x0, a.ax.inv_h, a.ax.uniform);
            const int yn = left_node_index(ym, a.ay.g, a.ay.n, a.ay.x0, a.ay.inv_h, a.ay.uniform);
            key = (uint32_t)((xn - 1) * (a.ynum - 1) + (yn - 1));
            if (a.out_xn) a.out_xn[m] = xn;
            if (a.out_yn) a.out_yn[m] = yn;
        }
    }
    const uint32_t prev = __shfl_up(key, 1);
    const bool head = lane == 0 || prev != key;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);   // lanes 0..lane
    const int leader = 63 - __builtin_clzll(heads & upto);
    const unsigned long long after = heads & ~upto;
    const int next = after ? __builtin_ctzll(after) : 64;
    int base = 0;
    if (head && key != 0xffffffffu) base = atomicAdd(&count[key], next - lane);
    base = __shfl(base, leader);
    if (key == 0xffffffffu) return;
    const int r = base + lane - leader;
    if (r < 64) {
        sliced[sliced_slot((int)key, r)] = (uint32_t)m;
    } else {
        const int q = atomicAdd(npending, 1);
        pending[3 * (int64_t)q] = (uint32_t)m;
        pending[3 * (int64_t)q + 1] = key;
        pending[3 * (int64_t)q + 2] = (uint32_t)r;
    }
}

// exclusive scan of n ints in place-out: out[0..n] (out[n] = total).
// Pass 1: each 256-thread block scans 4096 elements and records its total.
constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_BLOCK = 256 * SCAN_PER_THREAD;

__device__ __forceinline__ int block_exclusive_scan256(int v, int* sh, int* total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    int off = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) off += (k < w) ? sh[k] : 0;
    *total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return off + x - v;
}

// over_only: scan count where it exceeds 64 (the overflow cells), else 0
__global__ void __launch_bounds__(256) scan_blocks_kernel(const int* __restrict__ in, int n, int* __restrict__ out,
                                                          int* __restrict__ bsum, int over_only)
{
    __shared__ int sh[4];
    const int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER_THREAD;
    int v[SCAN_PER_THREAD];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        if (over_only && v[k] <= 64) v[k] = 0;
        s += v[k];
    }
    int total;
    int run = block_exclusive_scan256(s, sh, &total);
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Pass 2 (one block): exclusive scan of the block totals, total at bsum[nb].
__global__ void __launch_bounds__(1024) scan_bsums_kernel(int* __restrict__ bsum, int nb)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int lo = t * per, hi = min(lo + per, nb);
    int s = 0;
    for (int k = lo; k < hi; ++k) s += bsum[k];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int y = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    int run = t ? part[t - 1] : 0;
    for (int k = lo; k < hi; ++k) {
        const int v = bsum[k];
        bsum[k] = run;
        run += v;
    }
    if (t == 1023) bsum[nb] = part[1023];
}

// Pass 3: block offsets added; out[n] = total.
__global__ void __launch_bounds__(256) scan_add_kernel(int* __restrict__ out, int n, const int* __restrict__ bsum,
                                                       int nb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += bsum[i / SCAN_BLOCK];
    if (i == n) out[n] = bsum[nb];
}

// overflow slots: marker q of the queue goes to its cell's overflow area
__global__ void __launch_bounds__(256) bin_overflow_kernel(const uint32_t* __restrict__ pending,
                                                           const int* __restrict__ npending,
                                                           const int* __restrict__ ovstart, uint32_t* __restrict__ overflow)
{
    const int np = *npending;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < np; q += gridDim.x * blockDim.x)
        overflow[ovstart[pending[3 * (int64_t)q + 1]] + pending[3 * (int64_t)q + 2]] = pending[3 * (int64_t)q];
}

__global__ void inv_table_kernel(const double* __restrict__ v, int n, double* __restrict__ inv)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) inv[k] = 1 / v[k];
}

// ascending bitonic sort of N (a power of two) values in registers
template <int N>
__device__ __forceinline__ void bitonic_sort(uint32_t (&v)[N])
{
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t p = v[i], q = v[l];
                    const bool up = (i & k) == 0;
                    v[i] = up ? min(p, q) : max(p, q);
                    v[l] = up ? max(p, q) : min(p, q);
                }
            }
        }
    }
}

template <int N>
__device__ __forceinline__ void sort_list_regs(uint32_t* L, int cnt)
{
    uint32_t v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = i < cnt ? L[i] : 0xffffffffu;
    bitonic_sort<N>(v);
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (i < cnt) L[i] = v[i];
}

// the same on a sliced list: slot i at L[64 i]
template <int N>
__device__ __forceinline__ void sort_sliced_regs(uint32_t* L, int cnt)
{
    uint32_t v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = i < cnt ? L[(int64_t)i << 6] : 0xffffffffu;
    bitonic_sort<N>(v);
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (i < cnt) L[(int64_t)i << 6] = v[i];
}

// lists longer than 64: runs of 64 sorted in registers, then bottom-up
// merges between the list and a scratch area of the same length
__device__ void sort_long_list(uint32_t* L, uint32_t* T, int cnt)
{
    for (int lo = 0; lo < cnt; lo += 64) sort_list_regs<64>(L + lo, min(64, cnt - lo));
    uint32_t *src = L, *dst = T;
    for (int width = 64; width < cnt; width *= 2) {
        for (int lo = 0; lo < cnt; lo += 2 * width) {
            const int mid = min(lo + width, cnt), hi = min(lo + 2 * width, cnt);
            int i = lo, j = mid, k = lo;
            while (i < mid && j < hi) dst[k++] = src[i] <= src[j] ? src[i++] : src[j++];
            while (i < mid) dst[k++] = src[i++];
            while (j < hi) dst[k++] = src[j++];
        }
        uint32_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != L)
        for (int k = 0; k < cnt; ++k) L[k] = src[k];
}

// Sort every cell's list by marker index (lane = cell, wavefront = group of
// 64 cells).  The register network's size is chosen per wavefront (its
// cells' longest list) so that lanes do not diverge.  A cell with more than
// 64 markers first copies its 64 sliced slots to the head of its overflow
// area, whose tail already holds the rest, and sorts the whole list there.
__global__ void __launch_bounds__(256) sort_lists_kernel(int ncell, const int* __restrict__ count,
                                                         const int* __restrict__ ovstart, uint32_t* __restrict__ sliced,
                                                         uint32_t* __restrict__ overflow, uint32_t* __restrict__ scratch)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = c < ncell;
    const int cnt = live ? count[c] : 0;
    const int wmax = __reduce_max_sync(~0ull, cnt);
    if (wmax <= 1) return;
    uint32_t* L = sliced + sliced_slot(c, 0);
    if (wmax <= 16) {
        sort_sliced_regs<16>(L, cnt);
    } else if (wmax <= 32) {
        sort_sliced_regs<32>(L, cnt);
    } else if (cnt <= 64) {
        sort_sliced_regs<64>(L, cnt);
    } else {
        uint32_t* O = overflow + ovstart[c];
        for (int i = 0; i < 64; ++i) O[i] = L[(int64_t)i << 6];
        sort_long_list(O, scratch + ovstart[c], cnt);
    }
}

// marker of slot t of cell c's sorted list
__device__ __forceinline__ uint32_t list_at(const uint32_t* sliced, const uint32_t* overflow, int c, int cnt,
                                            int ovs, int t)
{
    return cnt <= 64 ? sliced[sliced_slot(c, t)] : overflow[ovs + t];
}

// Partial-sum planes of the cell pass: plane p of cell c at part[p * pstride + c].
// Basic inclusive: in_w + q, in_rho + q (q = corner: 0 upper left, 1 lower
// left, 2 upper right, 3 lower right).  Basic exclusive: ex[f] + q for
// f = 0 weight, then cohesion, friction, flow viscosity, viscosity, 1/mu,
// sxy, plastic flag.  vy nodes: vy_w + k, vy_rho + k, k = 3 * row + col
// (row 0: the cell's own row, 1: the row below; col 0..2: vy-grid column
// minus the cell's basic column).  -1: plane not kept.
struct CellPlanes {
    int in_w, in_rho;
    int ex[8];
    int vy_w, vy_rho;
    int heat[6];   // heat scatter, inclusive: weight, TK, kt, rhocp, hr, ha (+ q)
    int nplane;
    int64_t pstride;
    double* part;
    const double* inv_mu;   // 1 / mat_mu per material (IEEE division)
};

enum { EXW, EXCOH, EXFRIC, EXEF, EXETA, EXMU, EXSXY, EXPL };
constexpr int EXF[8] = {-1, F_COH, F_FRIC, F_EFLOW, F_ETA, F_INVMU, F_SXY, F_PL};

// compile-time field masks of the specialised cell passes; CELL_GENERIC reads
// the requested fields at run time
constexpr unsigned CELL_GENERIC = 0xffffffffu;
constexpr unsigned CELL_RHO_ETA_MU = (1u << F_RHO) | (1u << F_ETA) | (1u << F_INVMU);
constexpr unsigned CELL_ETA = 1u << F_ETA;

// Values a compiled field set stages per marker in LDS: x, y and each
// requested property (the weight slot of the exclusive sums is the constant 1).
constexpr int cell_staged(unsigned wm)
{
    int n = 2;
    const int f[11] = {F_RHO, F_COH, F_FRIC, F_EFLOW, F_ETA, F_INVMU, F_SXY, F_PL, F_DIL, F_MELT, F_SXX};
    for (int k = 0; k < 11; ++k) n += (wm >> f[k]) & 1u;
    return n;
}
// markers staged per lane and pass
constexpr int CELL_STAGE = 6;

template <unsigned WM>
__device__ __forceinline__ bool cell_want(const ScatterArgs& a, int f)
{
    return WM == CELL_GENERIC ? a.L.off[f] >= 0 : ((WM >> f) & 1u) != 0;
}

// (x - g) / h, correctly rounded, from rh = RN(1/h)
__device__ __forceinline__ double cell_dist(double x, double g, double h, double rh)
{
    const double d = x - g;
    return div_in_range(d) ? rn_div(d, h, rh) : d / h;
}

struct CellLists {
    const int *count, *ovstart;
    const uint32_t *sliced, *overflow;
};

template <unsigned WM>
__global__ void __launch_bounds__(256) scatter_cells_kernel(ScatterArgs a, CellPlanes P, CellLists Ls)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.ncell) return;
    const int cnt = Ls.count[c];
    const int ovs = cnt > 64 ? Ls.ovstart[c] : 0;
    const int ny1 = a.ynum - 1;
    const int cj = c / ny1, ci = c - cj * ny1;   // 0-based column and row of the cell
    const double gx = a.ax.g[cj], hx = a.xstp_b[cj], gy = a.ay.g[ci], hy = a.ystp_b[ci];
    const double rhx = 1.0 / hx, rhy = 1.0 / hy;
    // the two vy-grid cells a marker of this cell can fall in (1-based left
    // nodes J and J + 1)
    const double gv0 = a.gridx_vy[cj], hv0 = a.xstp_vy[cj], gv1 = a.gridx_vy[cj + 1], hv1 = a.xstp_vy[cj + 1];
    const double rhv0 = 1.0 / hv0, rhv1 = 1.0 / hv1;
    const bool w_rho = cell_want<WM>(a, F_RHO);
    bool w_ex[8];
#pragma unroll
    for (int f = 1; f < 8; ++f) w_ex[f] = cell_want<WM>(a, EXF[f]);
    const bool w_dil = cell_want<WM>(a, F_DIL), w_melt = cell_want<WM>(a, F_MELT),
               w_sxx = cell_want<WM>(a, F_SXX);
    double in_w[4] = {0, 0, 0, 0}, in_rho[4] = {0, 0, 0, 0};
    double ex[8][4];
#pragma unroll
    for (int f = 0; f < 8; ++f)
#pragma unroll
        for (int q = 0; q < 4; ++q) ex[f][q] = 0.0;
    double vw[6] = {0, 0, 0, 0, 0, 0}, vr[6] = {0, 0, 0, 0, 0, 0};
    double w_c = 0.0, p_eta = 0.0, p_mu = 0.0, p_pl = 0.0, p_dil = 0.0, p_melt = 0.0, p_sxx = 0.0;
    const int J = cj + 1;   // 1-based basic column of the cell
    // the values of marker t of the list (loaded one marker ahead of the sums)
    struct MV {
        double xm, ym, rho, pv[8], dil, melt, sxx;
    };
    auto load = [&](int t) {
        MV v;
        const uint32_t m = list_at(Ls.sliced, Ls.overflow, c, cnt, ovs, t);
        v.xm = a.x[m];
        v.ym = a.y[m];
        v.rho = w_rho ? a.p.rho[m] : 0.0;
        v.pv[EXW] = 1.0;
        v.pv[EXCOH] = w_ex[EXCOH] ? a.p.cohesion[m] : 0.0;
        v.pv[EXFRIC] = w_ex[EXFRIC] ? a.p.fric[m] : 0.0;
        v.pv[EXEF] = w_ex[EXEF] ? a.p.eta_flow[m] : 0.0;
        v.pv[EXETA] = w_ex[EXETA] ? a.p.eta[m] : 0.0;
        v.pv[EXMU] = w_ex[EXMU] ? (a.p.mat_mu ? P.inv_mu[a.p.matid[m] - 1] : 1 / 1.0) : 0.0;
        v.pv[EXSXY] = w_ex[EXSXY] ? a.p.sxy[m] : 0.0;
        v.pv[EXPL] = w_ex[EXPL] ? (double)a.p.pfailure[m] : 0.0;
        v.dil = w_dil ? (double)a.p.dilatation[m] : 0.0;
        v.melt = w_melt ? a.p.meltfrac[m] : 0.0;
        v.sxx = w_sxx ? a.p.sxx[m] : 0.0;
        return v;
    };
    auto accumulate = [&](const MV& cur) {
        const double xm = cur.xm, ym = cur.ym, rho = cur.rho, dil = cur.dil, melt = cur.melt, sxx = cur.sxx;
        double pv[8];
#pragma unroll
        for (int f = 0; f < 8; ++f) pv[f] = cur.pv[f];
        if (w_ex[EXFRIC]) pv[EXFRIC] = asin(pv[EXFRIC]) * 180.0 / M_PI;
        const double dx = cell_dist(xm, gx, hx, rhx), dy = cell_dist(ym, gy, hy, rhy);
        // basic nodes (WeightFuncs.jl:67-140): inclusive to all four corners,
        // exclusive to the corners on the marker's side of the centre
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double w = ((q & 2) ? dx : (1.0 - dx)) * ((q & 1) ? dy : (1.0 - dy));
            const bool exq = ((q & 2) ? (dx >= 0.5) : (dx <= 0.5)) && ((q & 1) ? (dy >= 0.5) : (dy <= 0.5));
            in_w[q] += w * 1.0 * 1.0;
            if (w_rho) in_rho[q] += w * rho * 1.0;
            if (exq) {
                ex[EXW][q] += w * 1.0 * 1.0;
#pragma unroll
                for (int f = 1; f < 8; ++f)
                    if (w_ex[f]) ex[f][q] += w * pv[f] * 1.0;
            }
        }
        // pressure node of the cell: central rule (WeightFuncs.jl:178-190)
        {
            const double w = (1.0 - fabs(0.5 - dx)) * (1.0 - fabs(0.5 - dy));
            w_c += w * 1.0 * 1.0;
            if (w_ex[EXETA]) p_eta += w * pv[EXETA] * 1.0;
            if (w_ex[EXMU]) p_mu += w * pv[EXMU] * 1.0;
            if (w_ex[EXPL]) p_pl += w * pv[EXPL] * 1.0;
            if (w_dil) p_dil += w * dil * 1.0;
            if (w_melt) p_melt += w * melt * 1.0;
            if (w_sxx) p_sxx += w * sxx * 1.0;
        }
        // vy nodes: x on the vy grid (MarkerGridMapping.jl:228-265), basic y.
        // The marker's vy cell has left node J or J + 1 (J the cell's basic
        // column); it weights columns (xnv, xnv + 1) by (1 - dxv, dxv), i.e.
        // cell-relative columns (0, 1) or (1, 2).  The third column gets an
        // exact zero, which leaves its non-negative sums unchanged.
        {
            const int xnv = left_node_index(xm, a.avx.g, a.avx.n, a.avx.x0, a.avx.inv_h, a.avx.uniform);
            double wc[3];
            if (xnv == J) {
                const double dxv = cell_dist(xm, gv0, hv0, rhv0);
                wc[0] = 1.0 - dxv;
                wc[1] = dxv;
                wc[2] = 0.0;
            } else if (xnv == J + 1) {
                const double dxv = cell_dist(xm, gv1, hv1, rhv1);
                wc[0] = 0.0;
                wc[1] = 1.0 - dxv;
                wc[2] = dxv;
            } else {   // not reached for markers inside the cell; no vy node takes it
                wc[0] = wc[1] = wc[2] = 0.0;
            }
            const double wy[2] = {1.0 - dy, dy};
#pragma unroll
            for (int row = 0; row < 2; ++row)
#pragma unroll
                for (int col = 0; col < 3; ++col) {
                    const double w = wc[col] * wy[row];
                    vw[3 * row + col] += w * 1.0 * 1.0;
                    if (w_rho) vr[3 * row + col] += w * rho * 1.0;
                }
        }
    };
    if constexpr (WM == CELL_GENERIC) {
        // each marker's loads are issued one marker ahead of its sums
        MV nxt;
        if (cnt > 0) nxt = load(0);
        for (int t = 0; t < cnt; ++t) {
            const MV cur = nxt;
            if (t + 1 < cnt) nxt = load(t + 1);
            accumulate(cur);
        }
    } else {
        // CELL_STAGE markers at a time: their values are loaded together
        // into this lane's column of LDS (consecutive markers of a cell
        // usually share cache lines, which are then read once) and summed
        // from there in list order.  No lane reads another lane's column.
        constexpr int NS = cell_staged(WM);
        extern __shared__ double cell_stage[];
        double* col = cell_stage + threadIdx.x;
        auto put = [&](int u, int k, double v) { col[(u * NS + k) * 256] = v; };
        auto get = [&](int u, int k) { return col[(u * NS + k) * 256]; };
        for (int t0 = 0; t0 < cnt; t0 += CELL_STAGE) {
            const int nb = min(CELL_STAGE, cnt - t0);
#pragma unroll
            for (int u = 0; u < CELL_STAGE; ++u) {
                if (u >= nb) break;
                const MV v = load(t0 + u);
                int k = 0;
                put(u, k++, v.xm);
                put(u, k++, v.ym);
                if (w_rho) put(u, k++, v.rho);
#pragma unroll
                for (int f = 1; f < 8; ++f)
                    if (w_ex[f]) put(u, k++, v.pv[f]);
                if (w_dil) put(u, k++, v.dil);
                if (w_melt) put(u, k++, v.melt);
                if (w_sxx) put(u, k++, v.sxx);
            }
#pragma unroll
            for (int u = 0; u < CELL_STAGE; ++u) {
                if (u >= nb) break;
                MV v;
                int k = 0;
                v.xm = get(u, k++);
                v.ym = get(u, k++);
                v.rho = w_rho ? get(u, k++) : 0.0;
                v.pv[EXW] = 1.0;
#pragma unroll
                for (int f = 1; f < 8; ++f) v.pv[f] = w_ex[f] ? get(u, k++) : 0.0;
                v.dil = w_dil ? get(u, k++) : 0.0;
                v.melt = w_melt ? get(u, k++) : 0.0;
                v.sxx = w_sxx ? get(u, k++) : 0.0;
                accumulate(v);
            }
        }
    }
    double* part = P.part;
    const int64_t S = P.pstride;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        part[(int64_t)(P.in_w + q) * S + c] = in_w[q];
        if (P.in_rho >= 0) part[(int64_t)(P.in_rho + q) * S + c] = in_rho[q];
#pragma unroll
        for (int f = 0; f < 8; ++f)
            if (P.ex[f] >= 0) part[(int64_t)(P.ex[f] + q) * S + c] = ex[f][q];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        part[(int64_t)(P.vy_w + k) * S + c] = vw[k];
        if (P.vy_rho >= 0) part[(int64_t)(P.vy_rho + k) * S + c] = vr[k];
    }
    // pressure node outputs (BilinearAverage.jl:276-322, old value where no weight)
    a.wtetan[c] = w_c;
    auto ar = [&](double* out, double s) { if (out) out[c] = (w_c != 0) ? s / w_c : out[c]; };
    if (a.f.mun1) a.f.mun1[c] = (w_c != 0) ? 1.0 / (p_mu / w_c) : a.f.mun1[c];
    ar(a.f.etan1, p_eta);
    ar(a.f.sxx1, p_sxx);
    ar(a.f.plasticn, p_pl);
    ar(a.f.dilatation_grid, p_dil);
    ar(a.f.extractable_meltfrac_grid, p_melt);
}

// basic nodes from the four neighbouring cells' corner partials, in the
// fixed order upper-left, lower-left, upper-right, lower-right cell
__global__ void __launch_bounds__(256) scatter_basic_cells_kernel(ScatterArgs a, CellPlanes P)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    if (t >= ny * nx) return;
    const int i = t % ny + 1, j = t / ny + 1;
    const int ci[4] = {i, i - 1, i, i - 1}, cj[4] = {j, j, j - 1, j - 1};
    const double* part = P.part;
    const int64_t S = P.pstride;
    double w_in = 0.0, s_rho = 0.0, e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (!(ci[q] >= 1 && ci[q] <= ny - 1 && cj[q] >= 1 && cj[q] <= nx - 1)) continue;
        const int64_t c = (int64_t)(cj[q] - 1) * (ny - 1) + (ci[q] - 1);
        w_in += part[(int64_t)(P.in_w + q) * S + c];
        if (P.in_rho >= 0) s_rho += part[(int64_t)(P.in_rho + q) * S + c];
#pragma unroll
        for (int f = 0; f < 8; ++f)
            if (P.ex[f] >= 0) e[f] += part[(int64_t)(P.ex[f] + q) * S + c];
    }
    const double w_ex = e[EXW];
    a.wtnodes[t] = w_in;
    a.wtetas[t] = w_ex;
    auto ar = [&](double* out, double s, double w) { if (out) out[t] = (w != 0) ? s / w : out[t]; };
    ar(a.f.rho1, s_rho, w_in);
    if (a.f.mus1) a.f.mus1[t] = (w_ex != 0) ? 1.0 / (e[EXMU] / w_ex) : a.f.mus1[t];
    ar(a.f.etas1, e[EXETA], w_ex);
    ar(a.f.sxy1, e[EXSXY], w_ex);
    ar(a.f.plastics, e[EXPL], w_ex);
    ar(a.f.cohesion_grid, e[EXCOH], w_ex);
    ar(a.f.fric_degrees_grid, e[EXFRIC], w_ex);
    ar(a.f.eta_flow, e[EXEF], w_ex);
}

// vy nodes from the six neighbouring cells (rows i-1..i, columns jv-2..jv)
__global__ void __launch_bounds__(256) scatter_vy_cells_kernel(ScatterArgs a, CellPlanes P)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    if (t >= ny * (nx + 1)) return;
    const int i = t % ny + 1, jv = t / ny + 1;
    const double* part = P.part;
    const int64_t S = P.pstride;
    double w_s = 0.0, s_rho = 0.0;
#pragma unroll
    for (int l = 0; l < 6; ++l) {
        const int ci = i - 1 + (l & 1), cj = jv - 2 + (l >> 1);
        if (!(ci >= 1 && ci <= ny - 1 && cj >= 1 && cj <= nx - 1)) continue;
        const int64_t c = (int64_t)(cj - 1) * (ny - 1) + (ci - 1);
        const int k = 3 * ((l & 1) ? 0 : 1) + (jv - cj);   // row 0: the node is on the cell's own row
        w_s += part[(int64_t)(P.vy_w + k) * S + c];
        if (P.vy_rho >= 0) s_rho += part[(int64_t)(P.vy_rho + k) * S + c];
    }
    a.wtnodes_vy[t] = w_s;
    if (a.f.rho1_vy) a.f.rho1_vy[t] = (w_s != 0) ? s_rho / w_s : a.f.rho1_vy[t];
}

// heat scatter, cell pass: inclusive rule of the heat fields
__global__ void __launch_bounds__(256) scatter_heat_cells_kernel(ScatterArgs a, CellPlanes P, CellLists Ls)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.ncell) return;
    const int cnt = Ls.count[c];
    const int ovs = cnt > 64 ? Ls.ovstart[c] : 0;
    const int ny1 = a.ynum - 1;
    const int cj = c / ny1, ci = c - cj * ny1;
    const double gx = a.ax.g[cj], hx = a.xstp_b[cj], gy = a.ay.g[ci], hy = a.ystp_b[ci];
    const double rhx = 1.0 / hx, rhy = 1.0 / hy;
    const int* o = a.L.off;
    double acc[6][4];
#pragma unroll
    for (int f = 0; f < 6; ++f)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[f][q] = 0.0;
    for (int t = 0; t < cnt; ++t) {
        const uint32_t m = list_at(Ls.sliced, Ls.overflow, c, cnt, ovs, t);
        const double xm = a.x[m], ym = a.y[m];
        double v[6];
        v[0] = 1.0;
        v[1] = o[F_TK] >= 0 ? a.h.TK[m] : 0.0;
        v[2] = o[F_KT] >= 0 ? a.h.kt[m] : 0.0;
        v[3] = o[F_RHOCP] >= 0 ? a.h.rhocp[m] : 0.0;
        double hr = 0.0;
        if (o[F_HR] >= 0) {
            const int id = a.h.matid[m];
            hr = (id >= 1 && id <= a.h.nmat) ? a.h.mat_hr[id - 1] : __builtin_nan("");
            if (a.h.serpentinization_heat_production) hr = hr + a.h.serpentinization_heat_production[m];
        }
        v[4] = hr;
        v[5] = o[F_HA] >= 0 ? a.h.ha[m] : 0.0;
        const double dx = cell_dist(xm, gx, hx, rhx), dy = cell_dist(ym, gy, hy, rhy);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double w = ((q & 2) ? dx : (1.0 - dx)) * ((q & 1) ? dy : (1.0 - dy));
            acc[0][q] += w * 1.0 * 1.0;
#pragma unroll
            for (int f = 1; f < 6; ++f)
                if (P.heat[f] >= 0) acc[f][q] += w * v[f] * 1.0;
        }
    }
#pragma unroll
    for (int f = 0; f < 6; ++f)
        if (P.heat[f] >= 0)
#pragma unroll
            for (int q = 0; q < 4; ++q) P.part[(int64_t)(P.heat[f] + q) * P.pstride + c] = acc[f][q];
}

__global__ void __launch_bounds__(256) scatter_heat_nodes_cells_kernel(ScatterArgs a, CellPlanes P)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ynum, nx = a.xnum;
    if (t >= ny * nx) return;
    const int i = t % ny + 1, j = t / ny + 1;
    const int ci[4] = {i, i - 1, i, i - 1}, cj[4] = {j, j, j - 1, j - 1};
    double s[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (!(ci[q] >= 1 && ci[q] <= ny - 1 && cj[q] >= 1 && cj[q] <= nx - 1)) continue;
        const int64_t c = (int64_t)(cj[q] - 1) * (ny - 1) + (ci[q] - 1);
#pragma unroll
        for (int f = 0; f < 6; ++f)
            if (P.heat[f] >= 0) s[f] += P.part[(int64_t)(P.heat[f] + q) * P.pstride + c];
    }
    const double w_in = s[0];
    if (a.wtnodes) a.wtnodes[t] = w_in;
    auto ar = [&](double* out, double v) { if (out) out[t] = (w_in != 0) ? v / w_in : out[t]; };
    ar(a.hf.tk1, s[1]);
    ar(a.hf.kt1, s[2]);
    ar(a.hf.rhocp1, s[3]);
    ar(a.hf.hr1, s[4]);
    ar(a.hf.ha1, s[5]);
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

}  // extern "C"

namespace ebg {
namespace {

// Grid geometry, marker coordinates and the entry slots of the requested
// fields, shared by both scatter orders.
void scatter_geometry(ebg_ctx* ctx, int64_t n, const double* x, const double* y, const bool (&want)[NFIELD],
                      ScatterArgs& a)
{
    const Grid2D& g = ctx->grid;
    const int nx = (int)g.xnum, ny = (int)g.ynum;
    a.n = n;
    a.xnum = nx;
    a.ynum = ny;
    a.ncell = (nx - 1) * (ny - 1);
    a.x = x;
    a.y = y;
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
    int slot = 3;  // slots 0..2: dx, dy, vy-grid dx
    for (int q = 0; q < NFIELD; ++q) a.L.off[q] = want[q] ? slot++ : -1;
    a.L.nslot = slot;
}

// Cell-ordered binning: keys and ranks, the overflow offsets, the sliced
// lists sorted by marker index, and the partial-sum planes sized for P.
int scatter_cells_bin(ebg_ctx* ctx, int64_t n, const double* x, const double* y, const bool (&want)[NFIELD],
                      ScatterArgs& a, CellPlanes& P, CellLists& Ls)
{
    const Grid2D& g = ctx->grid;
    const int nx = (int)g.xnum, ny = (int)g.ynum;
    const int ncell = (nx - 1) * (ny - 1);
    const int ngroup = (ncell + 63) / 64;
    ScatterState& S = *(ScatterState*)ctx->scatter;
    hipStream_t s = ctx->stream;
    if (S.cap < n || S.ncell != ncell) {
        // the same buffers as the serial order: keys, ranks (vals), the
        // overflow area (vals_s) and its merge scratch (keys_s), counts (cstart)
        ScatterState fresh;
        std::swap(S.keys, fresh.keys);
        std::swap(S.keys_s, fresh.keys_s);
        std::swap(S.vals, fresh.vals);
        std::swap(S.vals_s, fresh.vals_s);
        std::swap(S.cstart, fresh.cstart);
        std::swap(S.cend, fresh.cend);
        std::swap(S.rowbase, fresh.rowbase);
        std::swap(S.tmp, fresh.tmp);
        S.tmp_bytes = 0;
        const size_t cap = (size_t)std::max<int64_t>(n, 1);
        EBG_HIP(hipMalloc(&S.keys, cap * 4));
        EBG_HIP(hipMalloc(&S.keys_s, cap * 4));
        EBG_HIP(hipMalloc(&S.vals, cap * 4));
        EBG_HIP(hipMalloc(&S.vals_s, cap * 4));
        EBG_HIP(hipMalloc(&S.cstart, (size_t)(ncell + 1) * 4));
        EBG_HIP(hipMalloc(&S.cend, (size_t)ncell * 4));
        EBG_HIP(hipMalloc(&S.rowbase, (size_t)((ncell + CGROUP - 1) / CGROUP + 1) * 4));
        S.cap = (int64_t)cap;
        S.ncell = ncell;
    }
    const int nb = (ncell + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (S.start_n < ncell + 1) {
        if (S.start) (void)hipFree(S.start);
        if (S.bsum) (void)hipFree(S.bsum);
        if (S.sliced) (void)hipFree(S.sliced);
        S.start = nullptr;
        S.bsum = nullptr;
        S.sliced = nullptr;
        S.start_n = 0;
        EBG_HIP(hipMalloc(&S.start, (size_t)(ncell + 1) * 4));
        EBG_HIP(hipMalloc(&S.bsum, (size_t)(nb + 1) * 4));
        EBG_HIP(hipMalloc(&S.sliced, (size_t)ngroup * 4096 * 4));
        S.start_n = ncell + 1;
    }
    if (S.pend_n < 3 * std::max<int64_t>(n, 1)) {
        if (S.pend) (void)hipFree(S.pend);
        S.pend = nullptr;
        S.pend_n = 0;
        EBG_HIP(hipMalloc(&S.pend, (size_t)std::max<int64_t>(n, 1) * 12));
        S.pend_n = 3 * std::max<int64_t>(n, 1);
    }
    const int64_t need = (int64_t)P.nplane * ncell;
    if (S.part_n < need) {
        if (S.part) (void)hipFree(S.part);
        S.part = nullptr;
        S.part_n = 0;
        EBG_HIP(hipMalloc(&S.part, (size_t)std::max<int64_t>(need, 1) * 8));
        S.part_n = need;
    }
    P.part = S.part;
    P.pstride = ncell;
    Ls.count = S.cstart;
    Ls.ovstart = S.start;
    Ls.sliced = S.sliced;
    Ls.overflow = S.vals_s;
    scatter_geometry(ctx, n, x, y, want, a);
    a.keys = S.keys;
    // counts and the pending-queue length (one more int after the counts)
    hipLaunchKernelGGL(fill_int_kernel, dim3(nbk((int64_t)ncell + 1)), dim3(256), 0, s, S.cstart, ncell + 1, 0);
    int* npending = S.cstart + ncell;
    if (n > 0)
        hipLaunchKernelGGL(bin_slots_kernel, dim3(nbk(n)), dim3(256), 0, s, a, S.cstart, S.sliced, S.pend,
                           npending);
    // offsets of the overflow areas (cells of more than 64 markers)
    hipLaunchKernelGGL(scan_blocks_kernel, dim3(nb), dim3(256), 0, s, S.cstart, ncell, S.start, S.bsum, 1);
    hipLaunchKernelGGL(scan_bsums_kernel, dim3(1), dim3(1024), 0, s, S.bsum, nb);
    hipLaunchKernelGGL(scan_add_kernel, dim3(nbk((int64_t)ncell + 1)), dim3(256), 0, s, S.start, ncell, S.bsum, nb);
    if (n > 0) {
        hipLaunchKernelGGL(bin_overflow_kernel, dim3(1024), dim3(256), 0, s, S.pend, npending, S.start, S.vals_s);
        hipLaunchKernelGGL(sort_lists_kernel, dim3(nbk((int64_t)ngroup * 64)), dim3(256), 0, s, ncell, S.cstart,
                           S.start, S.sliced, S.vals_s, S.vals);
    }
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

// Binning shared by the Stokes and heat scatters: cell keys of the inside
// markers, the stable radix sort, per-cell ranges, the sliced row layout and
// the entry pass that gathers each marker's distances and the values of the
// requested slots.  On return a is ready for the node kernels.
int scatter_bin(ebg_ctx* ctx, int64_t n, const double* x, const double* y, const bool (&want)[NFIELD],
                bool heat, ScatterArgs& a)
{
    const Grid2D& g = ctx->grid;
    const int nx = (int)g.xnum, ny = (int)g.ynum;
    const int ncell = (nx - 1) * (ny - 1);
    if (!ctx->scatter) ctx->scatter = new ScatterState();
    ScatterState& S = *(ScatterState*)ctx->scatter;
    hipStream_t s = ctx->stream;
    if (S.cap < n || S.ncell != ncell) {
        ScatterState fresh;
        std::swap(S.keys, fresh.keys);
        std::swap(S.keys_s, fresh.keys_s);
        std::swap(S.vals, fresh.vals);
        std::swap(S.vals_s, fresh.vals_s);
        std::swap(S.cstart, fresh.cstart);
        std::swap(S.cend, fresh.cend);
        std::swap(S.rowbase, fresh.rowbase);
        std::swap(S.tmp, fresh.tmp);
        S.tmp_bytes = 0;
        const size_t cap = (size_t)std::max<int64_t>(n, 1);
        EBG_HIP(hipMalloc(&S.keys, cap * 4));
        EBG_HIP(hipMalloc(&S.keys_s, cap * 4));
        EBG_HIP(hipMalloc(&S.vals, cap * 4));
        EBG_HIP(hipMalloc(&S.vals_s, cap * 4));
        EBG_HIP(hipMalloc(&S.cstart, (size_t)(ncell + 1) * 4));
        EBG_HIP(hipMalloc(&S.cend, (size_t)ncell * 4));
        EBG_HIP(hipMalloc(&S.rowbase, (size_t)((ncell + CGROUP - 1) / CGROUP + 1) * 4));
        S.cap = (int64_t)cap;
        S.ncell = ncell;
    }
    scatter_geometry(ctx, n, x, y, want, a);
    a.keys = S.keys;
    a.vals = S.vals;
    a.keys_s = S.keys_s;
    a.vals_s = S.vals_s;
    a.cstart = S.cstart;
    a.cend = S.cend;
    const int ngroup = (ncell + CGROUP - 1) / CGROUP;
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
    // rows of each 64-cell group, their offsets, and the total for sizing
    hipLaunchKernelGGL(group_rows_kernel, dim3(nbk((int64_t)ngroup * CGROUP)), dim3(256), 0, s, S.cstart,
                       S.cend, ncell, ngroup, S.rowbase);
    hipLaunchKernelGGL(scan_rows_kernel, dim3(1), dim3(1024), 0, s, S.rowbase, ngroup);
    if (!S.h_rows) EBG_HIP(hipHostMalloc((void**)&S.h_rows, sizeof(int)));
    EBG_HIP(hipMemcpyAsync(S.h_rows, S.rowbase + ngroup, sizeof(int), hipMemcpyDeviceToHost, s));
    EBG_HIP(hipStreamSynchronize(s));
    const int64_t nent = (int64_t)*S.h_rows * CGROUP;
    if (S.ecap < nent) {
        if (S.emark) (void)hipFree(S.emark);
        if (S.exnv) (void)hipFree(S.exnv);
        S.emark = nullptr;
        S.exnv = nullptr;
        S.ecap = 0;
        EBG_HIP(hipMalloc(&S.emark, (size_t)nent * 4));
        EBG_HIP(hipMalloc(&S.exnv, (size_t)nent * 4));
        S.ecap = nent;
    }
    if (S.evcap < nent * a.L.nslot) {
        if (S.eval) (void)hipFree(S.eval);
        S.eval = nullptr;
        S.evcap = 0;
        EBG_HIP(hipMalloc(&S.eval, (size_t)nent * a.L.nslot * 8));
        S.evcap = nent * a.L.nslot;
    }
    a.rowbase = S.rowbase;
    a.emark = S.emark;
    a.exnv = S.exnv;
    a.eval = S.eval;
    a.slot_stride = nent;
    if (nent > 0) {
        if (heat)
            hipLaunchKernelGGL(scatter_entries_kernel<true>, dim3(nbk((int64_t)ngroup * CGROUP)), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL(scatter_entries_kernel<false>, dim3(nbk((int64_t)ngroup * CGROUP)), dim3(256), 0, s,
                               a);
    }
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

}  // namespace
}  // namespace ebg

extern "C" {

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
    if (!ctx->scatter) ctx->scatter = new ScatterState();
    ScatterState& S = *(ScatterState*)ctx->scatter;
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
    a.inside_flags = props->inside_flags;
    a.out_xn = fields->marker_xn;
    a.out_yn = fields->marker_yn;
    a.p = *props;
    a.f = *fields;
    a.wtnodes = wtn;
    a.wtetas = wts;
    a.wtetan = wte;
    a.wtnodes_vy = wtv;
    // entry slots: the marker values some requested output consumes
    const ebg_stokes_grid_fields& f = *fields;
    bool want[NFIELD] = {
        props->rho && (f.rho1 || f.rho1_vy),
        props->cohesion && f.cohesion_grid,
        props->fric && f.fric_degrees_grid,
        props->eta_flow && f.eta_flow,
        props->eta && (f.etas1 || f.etan1),
        f.mus1 || f.mun1,  // 1/mu is 1 without a material table
        props->sxy && f.sxy1,
        props->pfailure && (f.plastics || f.plasticn),
        props->dilatation && f.dilatation_grid,
        props->meltfrac && f.extractable_meltfrac_grid,
        props->sxx && f.sxx1};
    if (S.order == EBG_SCATTER_CELL_ORDER) {
        CellPlanes P{};
        int np_ = 0;
        P.in_w = np_;
        np_ += 4;
        P.in_rho = want[F_RHO] ? (np_ += 4) - 4 : -1;
        P.ex[EXW] = (np_ += 4) - 4;
        for (int k = 1; k < 8; ++k) P.ex[k] = want[EXF[k]] ? (np_ += 4) - 4 : -1;
        P.vy_w = (np_ += 6) - 6;
        P.vy_rho = want[F_RHO] ? (np_ += 6) - 6 : -1;
        for (int k = 0; k < 6; ++k) P.heat[k] = -1;
        P.nplane = np_;
        CellLists Ls{};
        int rc = scatter_cells_bin(ctx, n, props->x, props->y, want, a, P, Ls);
        if (rc) return rc;
        if (props->mat_mu) {
            if (S.inv_mu_n < props->nmat) {
                if (S.inv_mu) (void)hipFree(S.inv_mu);
                S.inv_mu = nullptr;
                S.inv_mu_n = 0;
                EBG_HIP(hipMalloc(&S.inv_mu, (size_t)props->nmat * 8));
                S.inv_mu_n = props->nmat;
            }
            hipLaunchKernelGGL(inv_table_kernel, dim3(nbk(props->nmat)), dim3(256), 0, s, props->mat_mu,
                               (int)props->nmat, S.inv_mu);
            P.inv_mu = S.inv_mu;
        }
        unsigned mask = 0;
        for (int q = 0; q < NSTOKES; ++q)
            if (want[q]) mask |= 1u << q;
        const void* k = (const void*)scatter_cells_kernel<CELL_GENERIC>;
        size_t lds = 0;
        if (mask == CELL_RHO_ETA_MU) {
            k = (const void*)scatter_cells_kernel<CELL_RHO_ETA_MU>;
            lds = (size_t)cell_staged(CELL_RHO_ETA_MU) * CELL_STAGE * 256 * 8;
        } else if (mask == CELL_ETA) {
            k = (const void*)scatter_cells_kernel<CELL_ETA>;
            lds = (size_t)cell_staged(CELL_ETA) * CELL_STAGE * 256 * 8;
        }
        void* args[] = {&a, &P, &Ls};
        EBG_HIP(hipLaunchKernel(k, dim3(nbk(np)), dim3(256), args, lds, s));
        hipLaunchKernelGGL(scatter_basic_cells_kernel, dim3(nbk(nb)), dim3(256), 0, s, a, P);
        hipLaunchKernelGGL(scatter_vy_cells_kernel, dim3(nbk(nvy)), dim3(256), 0, s, a, P);
        EBG_HIP(hipGetLastError());
        return EBG_OK;
    }
    int rc = scatter_bin(ctx, n, props->x, props->y, want, false, a);
    if (rc) return rc;
    hipLaunchKernelGGL(scatter_basic_kernel, dim3(nbk(nb)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(scatter_pressure_kernel, dim3(nbk(np)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(scatter_vy_kernel, dim3(nbk(nvy)), dim3(256), 0, s, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_scatter_set_order(ebg_ctx* ctx, int order)
{
    EBG_CHECK(ctx != nullptr, EBG_ERR_ARG, "ebg_scatter_set_order: ctx is NULL");
    EBG_CHECK(order == EBG_SCATTER_SERIAL_ORDER || order == EBG_SCATTER_CELL_ORDER, EBG_ERR_ARG,
              "ebg_scatter_set_order: order must be EBG_SCATTER_SERIAL_ORDER or EBG_SCATTER_CELL_ORDER");
    if (!ctx->scatter) ctx->scatter = new ScatterState();
    ((ScatterState*)ctx->scatter)->order = order;
    return EBG_OK;
}

int ebg_interpolate_marker_heat_parameters_to_grid(ebg_ctx* ctx, const ebg_marker_heat_props* props,
                                                   ebg_heat_grid_fields* fields)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_interpolate_marker_heat_parameters_to_grid: no grid");
    EBG_CHECK(props && fields, EBG_ERR_ARG, "ebg_interpolate_marker_heat_parameters_to_grid: NULL arguments");
    const int64_t n = props->marknum;
    EBG_CHECK(n >= 0 && n < (int64_t)0xffffffffLL, EBG_ERR_ARG, "heat scatter: marknum out of range");
    EBG_CHECK(n == 0 || (props->x && props->y), EBG_ERR_ARG, "heat scatter: NULL marker coordinates");
    const ebg_heat_grid_fields& f = *fields;
    EBG_CHECK(!f.tk1 || props->TK, EBG_ERR_ARG, "heat scatter: tk1 needs TK");
    EBG_CHECK(!f.kt1 || props->kt, EBG_ERR_ARG, "heat scatter: kt1 needs kt");
    EBG_CHECK(!f.rhocp1 || props->rhocp, EBG_ERR_ARG, "heat scatter: rhocp1 needs rhocp");
    EBG_CHECK(!f.ha1 || props->ha, EBG_ERR_ARG, "heat scatter: ha1 needs ha");
    EBG_CHECK(!f.hr1 || (props->matid && props->mat_hr && props->nmat >= 1), EBG_ERR_ARG,
              "heat scatter: hr1 needs matid, mat_hr and nmat >= 1");
    const Grid2D& g = ctx->grid;
    ScatterArgs a{};
    a.inside_flags = props->inside_flags;
    a.out_xn = fields->marker_xn;
    a.out_yn = fields->marker_yn;
    a.h = *props;
    a.hf = *fields;
    a.wtnodes = fields->wtnodes;
    bool want[NFIELD] = {};
    want[F_TK] = f.tk1 != nullptr;
    want[F_KT] = f.kt1 != nullptr;
    want[F_RHOCP] = f.rhocp1 != nullptr;
    want[F_HR] = f.hr1 != nullptr;
    want[F_HA] = f.ha1 != nullptr;
    if (!ctx->scatter) ctx->scatter = new ScatterState();
    if (((ScatterState*)ctx->scatter)->order == EBG_SCATTER_CELL_ORDER) {
        ScatterState& S = *(ScatterState*)ctx->scatter;
        CellPlanes P{};
        P.in_w = P.in_rho = P.vy_w = P.vy_rho = -1;
        for (int k = 0; k < 8; ++k) P.ex[k] = -1;
        const int hf[6] = {-1, F_TK, F_KT, F_RHOCP, F_HR, F_HA};
        int np_ = 0;
        for (int k = 0; k < 6; ++k) P.heat[k] = (k == 0 || want[hf[k]]) ? (np_ += 4) - 4 : -1;
        P.nplane = np_;
        CellLists Ls{};
        int rc = scatter_cells_bin(ctx, n, props->x, props->y, want, a, P, Ls);
        if (rc) return rc;
        const int ncell = (int)((g.xnum - 1) * (g.ynum - 1));
        hipLaunchKernelGGL(scatter_heat_cells_kernel, dim3(nbk(ncell)), dim3(256), 0, ctx->stream, a, P, Ls);
        hipLaunchKernelGGL(scatter_heat_nodes_cells_kernel, dim3(nbk(g.xnum * g.ynum)), dim3(256), 0, ctx->stream,
                           a, P);
        EBG_HIP(hipGetLastError());
        return EBG_OK;
    }
    int rc = scatter_bin(ctx, n, props->x, props->y, want, true, a);
    if (rc) return rc;
    hipLaunchKernelGGL(scatter_heat_kernel, dim3(nbk(g.xnum * g.ynum)), dim3(256), 0, ctx->stream, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

}  // extern "C"
