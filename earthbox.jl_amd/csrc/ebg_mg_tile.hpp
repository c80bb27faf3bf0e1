// LDS-tiled smoother for the multigrid preconditioner.
//
// A level's (i, j) index space [1, ynum+1] x [1, xnum+1] (the union of the vx,
// vy and pressure index ranges) is cut into TI x TJ tiles.  One workgroup owns
// one tile: it copies the window [I0-1, I0+TI] x [J0-1, J0+TJ] of all eight
// level fields (vx, vy, pr, etas, etan, RX, RY, RC; entries outside a field's
// bounds read 0) into LDS, runs nsweeps smoothing sweeps there, and writes the
// owned velocities and pressures back.  One sweep is exactly the global sweep
// restricted to the tile: the ghost-node boundary pass (BoundaryConditions.jl
// 22-53) on the window, then the red/black phases vx, vy, p of
// SolveStokes2dOpt.jl:9-76 (v += dR/diag*theta_s, p += etan*dRc*theta_c).
//
// Halo entries stay at their values from the start of the launch, so a
// multi-tile launch is a Schwarz (block Gauss-Seidel) smoother.  Tiles are
// launched in four colours ((bi & 1) + 2 (bj & 1)), so no tile reads an entry
// that a concurrently running tile writes: results are deterministic.  The
// additive variant launches every tile at once and writes a second buffer,
// so all tiles read the previous iterate (block Jacobi between tiles, four
// times fewer launches; also deterministic).  When
// one tile covers the whole level the sweeps equal the global phased sweeps,
// and the kernel also applies the pressure gauge shift and writes the
// residuals, replacing four launches per smoothing call.
#pragma once
#include "ebg_mg.hpp"

namespace ebg {

enum { TF_VX = 0, TF_VY, TF_PR, TF_ES, TF_EN, TF_RX, TF_RY, TF_RC, TF_N };

template <typename T, bool FAST>
struct TileView {
    static constexpr bool fast_div = FAST;
    T* s;             // LDS, field f at s + f * ww
    int w, ww;        // window rows (leading dimension) and entries
    int i0, j0;       // global (i, j) of window entry (0, 0)
    // LDS copies of the step arrays over the window: entry k of the reference
    // array (0-based) at [k - (j0 - 1)] for x and [k - (i0 - 1)] for y
    const T *xstp_b, *ystp_b, *xstp_vy, *ystp_vx;
    __device__ int o(int i, int j) const { return (j - j0) * w + (i - i0); }
    __device__ T get(int f, int i, int j) const { return s[f * ww + o(i, j)]; }
    __device__ T& at(int f, int i, int j) const { return s[f * ww + o(i, j)]; }
    __device__ T vx(int i, int j) const { return get(TF_VX, i, j); }
    __device__ T vy(int i, int j) const { return get(TF_VY, i, j); }
    __device__ T pr(int i, int j) const { return get(TF_PR, i, j); }
    __device__ T es(int i, int j) const { return get(TF_ES, i, j); }
    __device__ T en(int i, int j) const { return get(TF_EN, i, j); }
    __device__ T RX(int i, int j) const { return get(TF_RX, i, j); }
    __device__ T RY(int i, int j) const { return get(TF_RY, i, j); }
    __device__ T RC(int i, int j) const { return get(TF_RC, i, j); }
    __device__ T xb(int k) const { return xstp_b[k - (j0 - 1)]; }
    __device__ T yb(int k) const { return ystp_b[k - (i0 - 1)]; }
    __device__ T xv(int k) const { return xstp_vy[k - (j0 - 1)]; }
    __device__ T yv(int k) const { return ystp_vx[k - (i0 - 1)]; }
};

// LDS bytes of one tile window: eight fields plus four step-array windows
template <typename T>
inline size_t tile_lds_bytes(int ti, int tj)
{
    return ((size_t)TF_N * (ti + 2) * (tj + 2) + 2 * (size_t)(ti + 2) + 2 * (size_t)(tj + 2)) * sizeof(T);
}

struct TileSpec {
    int ti, tj;      // tile extent in i and j
    int nti, ntj;    // tiles along i and j
    int colours;     // 1: every tile in one launch, 4: one colour per launch
};

// Bounds of field f at (i, j) on a level with xnum x ynum basic nodes.
__device__ __forceinline__ bool tile_in_bounds(int f, int i, int j, int xnum, int ynum)
{
    switch (f) {
        case TF_VX: case TF_RX: return i >= 1 && i <= ynum + 1 && j >= 1 && j <= xnum;
        case TF_VY: case TF_RY: return i >= 1 && i <= ynum && j >= 1 && j <= xnum + 1;
        case TF_ES: return i >= 1 && i <= ynum && j >= 1 && j <= xnum;
        default: return i >= 1 && i <= ynum - 1 && j >= 1 && j <= xnum - 1;  // pr, etan, RC
    }
}

template <typename T>
__device__ __forceinline__ T* tile_field_ptr(const MGLevelDev<T>& l, int f)
{
    switch (f) {
        case TF_VX: return l.vx;
        case TF_VY: return l.vy;
        case TF_PR: return l.pr;
        case TF_ES: return l.etas;
        case TF_EN: return l.etan;
        case TF_RX: return l.RX;
        case TF_RY: return l.RY;
        default: return l.RC;
    }
}

// leading dimension (rows) of field f
__device__ __forceinline__ int tile_field_ld(int f, int ynum)
{
    switch (f) {
        case TF_VX: case TF_RX: return ynum + 1;
        case TF_VY: case TF_RY: case TF_ES: return ynum;
        default: return ynum - 1;
    }
}

struct TileSmoothArgs {
    int colour, nsweeps;
    int finalize;        // single tile: gauge shift (if normalize) and residuals
    int normalize;
};

// FAST: FP32 residual divisions by the hardware reciprocal (preconditioner).
template <typename T, bool FAST>
__global__ void __launch_bounds__(1024) mg_tile_smooth_kernel(MGLevelDev<T> l, TileSpec ts,
                                                              TileSmoothArgs a, T theta_s,
                                                              T theta_c, T d_top, T d_bottom,
                                                              T d_left, T d_right, T prnorm,
                                                              const T* prnorm_dev, T* ovx, T* ovy,
                                                              T* opr, T* shift)
{
    extern __shared__ __align__(16) unsigned char tile_smem[];
    T* sm = reinterpret_cast<T*>(tile_smem);
    int bi, bj;
    if (ts.colours == 1) {
        bi = blockIdx.x % ts.nti;
        bj = blockIdx.x / ts.nti;
    } else {
        const int ci = a.colour & 1, cj = a.colour >> 1;
        const int ni = (ts.nti - ci + 1) / 2;
        bi = 2 * (blockIdx.x % ni) + ci;
        bj = 2 * (blockIdx.x / ni) + cj;
    }
    const int xnum = l.xnum, ynum = l.ynum;
    const int I0 = 1 + bi * ts.ti, J0 = 1 + bj * ts.tj;
    TileView<T, FAST> v;
    v.s = sm;
    v.w = ts.ti + 2;
    v.ww = v.w * (ts.tj + 2);
    v.i0 = I0 - 1;
    v.j0 = J0 - 1;
    const int wj = ts.tj + 2;
    T* steps = sm + (size_t)TF_N * v.ww;
    T* sxb = steps;
    T* sxv = sxb + wj;
    T* syb = sxv + wj;
    T* syv = syb + v.w;
    v.xstp_b = sxb;
    v.xstp_vy = sxv;
    v.ystp_b = syb;
    v.ystp_vx = syv;
    const int nt = blockDim.x;
    // (i, j) of entry e of the owned block and of the window: every thread
    // starts at e = threadIdx.x, whose indices are computed once here (the
    // divisions by the runtime extents then run only for later entries)
    const int tid = threadIdx.x;
    const int own_i = I0 + tid % ts.ti, own_j = J0 + tid / ts.ti;
    const int win_i = v.i0 + tid % v.w, win_j = v.j0 + tid / v.w;
    auto owned = [&](int e, int& i, int& j) {
        if (e == tid) { i = own_i; j = own_j; }
        else { i = I0 + e % ts.ti; j = J0 + e / ts.ti; }
    };
    auto windowed = [&](int e, int& i, int& j) {
        if (e == tid) { i = win_i; j = win_j; }
        else { i = v.i0 + e % v.w; j = v.j0 + e / v.w; }
    };
    // ---- step windows: x entries k in [j0-1, j0+tj], y entries in [i0-1, i0+ti]
    for (int e = threadIdx.x; e < wj; e += nt) {
        const int k = v.j0 - 1 + e;
        sxb[e] = (k >= 0 && k <= xnum - 2) ? l.xstp_b[k] : T(1);
        sxv[e] = (k >= 0 && k <= xnum - 1) ? l.xstp_vy[k] : T(1);
    }
    for (int e = threadIdx.x; e < v.w; e += nt) {
        const int k = v.i0 - 1 + e;
        syb[e] = (k >= 0 && k <= ynum - 2) ? l.ystp_b[k] : T(1);
        syv[e] = (k >= 0 && k <= ynum - 1) ? l.ystp_vx[k] : T(1);
    }
    // ---- load the window: two entries per thread per pass, every field's
    // load issued before the first LDS store (one wait per pass)
    for (int e0 = threadIdx.x; e0 < v.ww; e0 += 2 * nt) {
        T val[2][TF_N];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = e0 + u * nt;
            int i, j;
            windowed(e, i, j);
#pragma unroll
            for (int f = 0; f < TF_N; ++f) {
                val[u][f] = T(0);
                if (e < v.ww && tile_in_bounds(f, i, j, xnum, ynum))
                    val[u][f] = tile_field_ptr(l, f)[(size_t)(j - 1) * tile_field_ld(f, ynum) + (i - 1)];
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = e0 + u * nt;
            if (e < v.ww) {
#pragma unroll
                for (int f = 0; f < TF_N; ++f) sm[f * v.ww + e] = val[u][f];
            }
        }
    }
    __syncthreads();
    const int nown = ts.ti * ts.tj;
    for (int sw = 0; sw < a.nsweeps; ++sw) {
        // ---- ghost-node pass: ghosts from their interior neighbours and zero
        // boundary-normal velocities.  The thread of ghost (1, j) (or (i, 1))
        // also zeroes the wall node it read at the corners, so no entry is
        // read and written by different threads in this pass.
        const bool has_row1 = v.i0 <= 1, has_col1 = v.j0 <= 1;
        for (int e = threadIdx.x; e < v.ww; e += nt) {
            int i, j;
            windowed(e, i, j);
            if (j >= 1 && j <= xnum) {
                const bool wall = (j == 1 || j == xnum);
                if (i == 1) {
                    const T old = v.vx(2, j);
                    v.at(TF_VX, 1, j) = d_top * old;
                    if (wall) v.at(TF_VX, 2, j) = T(0);
                } else if (i == ynum + 1) {
                    v.at(TF_VX, i, j) = wall ? d_bottom * T(0) : d_bottom * v.vx(ynum, j);
                } else if (wall && i >= 3 && i <= ynum) {
                    v.at(TF_VX, i, j) = T(0);
                } else if (wall && i == 2 && !has_row1) {
                    v.at(TF_VX, i, j) = T(0);
                }
            }
            if (i >= 1 && i <= ynum) {
                const bool wall = (i == 1 || i == ynum);
                if (j == 1) {
                    const T old = v.vy(i, 2);
                    v.at(TF_VY, i, 1) = d_left * old;
                    if (wall) v.at(TF_VY, i, 2) = T(0);
                } else if (j == xnum + 1) {
                    v.at(TF_VY, i, j) = wall ? d_right * T(0) : d_right * v.vy(i, xnum);
                } else if (wall && j >= 3 && j <= xnum) {
                    v.at(TF_VY, i, j) = T(0);
                } else if (wall && j == 2 && !has_col1) {
                    v.at(TF_VY, i, j) = T(0);
                }
            }
        }
        __syncthreads();
        for (int color = 0; color <= 1; ++color) {
            for (int e = threadIdx.x; e < nown; e += nt) {
                int i, j;
                owned(e, i, j);
                if (i >= 2 && i <= ynum && j >= 2 && j <= xnum - 1 && ((i + j) & 1) == color &&
                    !(l.fix_vx && l.fix_vx[(size_t)(j - 1) * (ynum + 1) + (i - 1)])) {
                    T coef;
                    const T dR = vx_residual_v<T>(v, i, j, &coef);
                    v.at(TF_VX, i, j) += vdiv<TileView<T, FAST>, T>(dR, coef) * theta_s;
                }
            }
            __syncthreads();
            for (int e = threadIdx.x; e < nown; e += nt) {
                int i, j;
                owned(e, i, j);
                if (i >= 2 && i <= ynum - 1 && j >= 2 && j <= xnum && ((i + j) & 1) == color &&
                    !(l.fix_vy && l.fix_vy[(size_t)(j - 1) * ynum + (i - 1)])) {
                    T coef;
                    const T dR = vy_residual_v<T>(v, i, j, &coef);
                    v.at(TF_VY, i, j) += vdiv<TileView<T, FAST>, T>(dR, coef) * theta_s;
                }
            }
            __syncthreads();
            for (int e = threadIdx.x; e < nown; e += nt) {
                int i, j;
                owned(e, i, j);
                if (i <= ynum - 1 && j <= xnum - 1 && ((i + j) & 1) == color &&
                    !(l.fix_p && l.fix_p[(size_t)(j - 1) * (ynum - 1) + (i - 1)])) {
                    const T dR = pr_residual_v<T>(v, i, j);
                    v.at(TF_PR, i, j) += v.en(i, j) * dR * theta_c;
                }
            }
            __syncthreads();
        }
    }
    if (a.finalize) {
        // single tile: gauge shift p += prnorm - p(1,1) (PressureUpdate.jl:24-37)
        if (a.normalize) {
            const T shift = (prnorm_dev ? prnorm_dev[0] : prnorm) - v.pr(1, 1);
            __syncthreads();
            for (int e = threadIdx.x; e < nown; e += nt) {
                int i, j;
                owned(e, i, j);
                if (i <= ynum - 1 && j <= xnum - 1) v.at(TF_PR, i, j) += shift;
            }
            __syncthreads();
        }
        // residuals as mg_residual_kernel
        for (int e = threadIdx.x; e < nown; e += nt) {
            int i, j;
                owned(e, i, j);
            if (j <= xnum && i <= ynum + 1) {
                T r = T(0);
                if (i >= 2 && i <= ynum && j >= 2 && j <= xnum - 1) r = vx_residual_v<T>(v, i, j, (T*)nullptr);
                if (l.fix_vx && l.fix_vx[(size_t)(j - 1) * (ynum + 1) + (i - 1)]) r = T(0);
                l.rx[(size_t)(j - 1) * (ynum + 1) + (i - 1)] = r;
            }
            if (i <= ynum && j <= xnum + 1) {
                T r = T(0);
                if (i >= 2 && i <= ynum - 1 && j >= 2 && j <= xnum) r = vy_residual_v<T>(v, i, j, (T*)nullptr);
                if (l.fix_vy && l.fix_vy[(size_t)(j - 1) * ynum + (i - 1)]) r = T(0);
                l.ry[(size_t)(j - 1) * ynum + (i - 1)] = r;
            }
            if (i <= ynum - 1 && j <= xnum - 1) {
                const size_t k = (size_t)(j - 1) * (ynum - 1) + (i - 1);
                l.rc[k] = (l.fix_p && l.fix_p[k]) ? T(0) : pr_residual_v<T>(v, i, j);
            }
        }
    }
    // gauge shift of the level's pressures, prnorm - p(1,1), for the
    // pressure update launched after the smoothing (block of node (1,1))
    if (shift && I0 == 1 && J0 == 1 && threadIdx.x == 0)
        shift[0] = (prnorm_dev ? prnorm_dev[0] : prnorm) - v.pr(1, 1);
    // ---- write back the owned velocities and pressures (to the level's own
    // arrays, or to its second buffers in an all-tiles launch)
    for (int e = threadIdx.x; e < nown; e += nt) {
        int i, j;
                owned(e, i, j);
        if (tile_in_bounds(TF_VX, i, j, xnum, ynum)) ovx[(size_t)(j - 1) * (ynum + 1) + (i - 1)] = v.vx(i, j);
        if (tile_in_bounds(TF_VY, i, j, xnum, ynum)) ovy[(size_t)(j - 1) * ynum + (i - 1)] = v.vy(i, j);
        if (tile_in_bounds(TF_PR, i, j, xnum, ynum)) opr[(size_t)(j - 1) * (ynum - 1) + (i - 1)] = v.pr(i, j);
    }
}

}  // namespace ebg
