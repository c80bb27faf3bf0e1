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
    ~ScatterState()
    {
        void* ps[] = {keys, keys_s, vals, vals_s, cstart, cend, rowbase, tmp, wbuf, emark, exnv, eval};
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
        EBG_HIP(hipMalloc(&S.cstart, (size_t)ncell * 4));
        EBG_HIP(hipMalloc(&S.cend, (size_t)ncell * 4));
        EBG_HIP(hipMalloc(&S.rowbase, (size_t)((ncell + CGROUP - 1) / CGROUP + 1) * 4));
        S.cap = (int64_t)cap;
        S.ncell = ncell;
    }
    a.n = n;
    a.xnum = nx;
    a.ynum = ny;
    a.ncell = ncell;
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
    a.keys = S.keys;
    a.vals = S.vals;
    a.keys_s = S.keys_s;
    a.vals_s = S.vals_s;
    a.cstart = S.cstart;
    a.cend = S.cend;
    int slot = 3;  // slots 0..2: dx, dy, vy-grid dx
    for (int q = 0; q < NFIELD; ++q) a.L.off[q] = want[q] ? slot++ : -1;
    a.L.nslot = slot;
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
    int rc = scatter_bin(ctx, n, props->x, props->y, want, false, a);
    if (rc) return rc;
    hipLaunchKernelGGL(scatter_basic_kernel, dim3(nbk(nb)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(scatter_pressure_kernel, dim3(nbk(np)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(scatter_vy_kernel, dim3(nbk(nvy)), dim3(256), 0, s, a);
    EBG_HIP(hipGetLastError());
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
    int rc = scatter_bin(ctx, n, props->x, props->y, want, true, a);
    if (rc) return rc;
    hipLaunchKernelGGL(scatter_heat_kernel, dim3(nbk(g.xnum * g.ynum)), dim3(256), 0, ctx->stream, a);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

}  // extern "C"
