// Stokes-continuity solve on the assembled-operator conventions.
//
// Replaces SystemSolver.solve_system_optimized / ParallelSolver.parallel_direct_solver
// (solver/system_solver/SystemSolver.jl:67-91, parallel_solver/ParallelSolver.jl:52-126):
// FGMRES(m) in FP64 on the matrix-free assembled operator (ebg_stokes_op.hpp),
// right-preconditioned by one V-cycle of the reference's multigrid
// (ebg_mg.hpp) run in FP32 (mixed precision) or FP64 (plain).  Also the
// companions K10 (unpack + velocity BCs, SystemSolver.jl:141-271) and K11
// (|L*S - R|, StokesResiduals.jl:63-82).
#include "ebg_common.hpp"
#include "ebg_krylov.hpp"
#include "ebg_mg.hpp"
#include "ebg_stokes_op.hpp"
#include "ebg_learned.hpp"

#include <rocblas/rocblas.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

namespace ebg {

static inline unsigned nb(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------------ kernels
template <bool WANT_R, class Lay>
__global__ void __launch_bounds__(256) stokes_rows_kernel(StokesOpDev s, const double* __restrict__ x,
                                                          double* __restrict__ y, Lay L)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny1 = s.ynum - 1;
    if (c >= L.nc) return;
    const int i = c % ny1 + 1, j = c / ny1 + 1;
    y[L.vx(c)] = x_row<WANT_R>(s, i, j, x, L);
    y[L.vy(c)] = y_row<WANT_R>(s, i, j, x, L);
    y[L.p(c)] = p_row<WANT_R>(s, i, j, x, L);
}

// y = b - L*x (planar)
__global__ void __launch_bounds__(256) residual_planar_kernel(StokesOpDev s, const double* __restrict__ x,
                                                              const double* __restrict__ b,
                                                              double* __restrict__ r, LayoutPlanar L)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny1 = s.ynum - 1;
    if (c >= L.nc) return;
    const int i = c % ny1 + 1, j = c / ny1 + 1;
    r[L.vx(c)] = b[L.vx(c)] - x_row<false>(s, i, j, x, L);
    r[L.vy(c)] = b[L.vy(c)] - y_row<false>(s, i, j, x, L);
    r[L.p(c)] = b[L.p(c)] - p_row<false>(s, i, j, x, L);
}

__global__ void interleaved_to_planar_kernel(const double* __restrict__ a, double* __restrict__ p, int nc)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    p[c] = a[3 * (int64_t)c];
    p[nc + c] = a[3 * (int64_t)c + 1];
    p[2 * (int64_t)nc + c] = a[3 * (int64_t)c + 2];
}
__global__ void planar_to_interleaved_kernel(const double* __restrict__ p, double* __restrict__ a, int nc)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    a[3 * (int64_t)c] = p[c];
    a[3 * (int64_t)c + 1] = p[nc + c];
    a[3 * (int64_t)c + 2] = p[2 * (int64_t)nc + c];
}

// K11: |L*S - R| in the reference's interleaved order
__global__ void abs_diff_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                double* __restrict__ out, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = fabs(a[t] - b[t]);
}

// K10: reload_solu_vel! (SystemSolver.jl:188-221), interleaved S
__global__ void unpack_kernel(StokesOpDev s, const double* __restrict__ S, double* __restrict__ vx1,
                              double* __restrict__ vy1, double* __restrict__ pr1)
{
    // one thread per vx-grid node (covers vy and pr nodes too)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int xnum = s.xnum, ynum = s.ynum;
    const int ni = ynum + 1;
    if (t >= ni * (xnum + 1)) return;
    const int I = t % ni + 1, J = t / ni + 1;  // node indices on the vx/vy grids
    if (J <= xnum) {  // vx1[I, J]: cell (I-1, J-1) for J-1 < xnum-1
        double v = 0.0;
        const int i = I - 1, j = J - 1;
        if (i >= 1 && i <= ynum - 1 && j >= 1 && j < xnum - 1) v = S[3 * (int64_t)((j - 1) * (ynum - 1) + (i - 1))];
        vx1[(size_t)(J - 1) * (ynum + 1) + (I - 1)] = v;
    }
    if (I <= ynum) {  // vy1[I, J]
        double v = 0.0;
        const int i = I - 1, j = J - 1;
        if (i >= 1 && i < ynum - 1 && j >= 1 && j <= xnum - 1) v = S[3 * (int64_t)((j - 1) * (ynum - 1) + (i - 1)) + 1];
        vy1[(size_t)(J - 1) * ynum + (I - 1)] = v;
    }
    if (I <= ynum - 1 && J <= xnum - 1)
        pr1[(size_t)(J - 1) * (ynum - 1) + (I - 1)] = S[3 * (int64_t)((J - 1) * (ynum - 1) + (I - 1)) + 2] * s.pscale;
}

// apply_vel_bc! (SystemSolver.jl:237-271): top/bottom rows first, then
// left/right columns (the column pass overwrites the corners, as the reference).
__global__ void vel_bc_rows_kernel(StokesOpDev s, double* vx1, double* vy1)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int xnum = s.xnum, ynum = s.ynum;
    if (j <= xnum) {
        double* col = vx1 + (size_t)(j - 1) * (ynum + 1);
        col[0] = EBG_BX(btopx, j, 1) + EBG_BX(btopx, j, 2) * col[1];
        col[ynum] = EBG_BX(bbottomx, j, 1) + EBG_BX(bbottomx, j, 2) * col[ynum - 1];
    }
    if (j <= xnum + 1) {
        double* col = vy1 + (size_t)(j - 1) * ynum;
        col[0] = EBG_BX(btopy, j, 1) + EBG_BX(btopy, j, 2) * col[1];
        col[ynum - 1] = EBG_BX(bbottomy, j, 1) + EBG_BX(bbottomy, j, 2) * col[ynum - 2];
    }
}
__global__ void vel_bc_cols_kernel(StokesOpDev s, double* vx1, double* vy1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int xnum = s.xnum, ynum = s.ynum;
    if (i <= ynum + 1) {
        double* row = vx1 + (i - 1);
        const size_t ld = ynum + 1;
        row[0] = EBG_BY(bleftx, i, 1) + EBG_BY(bleftx, i, 2) * row[ld];
        row[(size_t)(xnum - 1) * ld] = EBG_BY(brightx, i, 1) + EBG_BY(brightx, i, 2) * row[(size_t)(xnum - 2) * ld];
    }
    if (i <= ynum) {
        double* row = vy1 + (i - 1);
        const size_t ld = ynum;
        row[0] = EBG_BY(blefty, i, 1) + EBG_BY(blefty, i, 2) * row[ld];
        row[(size_t)xnum * ld] = EBG_BY(brighty, i, 1) + EBG_BY(brighty, i, 2) * row[(size_t)(xnum - 1) * ld];
    }
}

// ---- deterministic reductions (declared in ebg_krylov.hpp) ----
constexpr int COMPAT_BLOCKS = 256;  // partial sums of the mode-0 compatibility RHS

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// partial[q*gridDim.x + b] = block b's grid-stride share of <w, V[q]> for
// every q < k.  The basis is taken DOT_CHUNK vectors per pass, with the
// accumulators in registers (a fixed-size, fully unrolled array), so w is
// re-read once per chunk and the basis exactly once.  A non-null gate that
// holds 0 makes the whole grid return at once.
__global__ void __launch_bounds__(256) multidot_chunks_kernel(const double* __restrict__ w,
                                                             const double* __restrict__ V, int64_t ldv,
                                                             int k, int64_t n, double* __restrict__ partial,
                                                             const int* __restrict__ gate)
{
    if (gate && *gate == 0) return;
    __shared__ double sh[4][DOT_CHUNK];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int q0 = 0; q0 < k; q0 += DOT_CHUNK) {
        const int kc = min(DOT_CHUNK, k - q0);
        const double* Vc = V + (int64_t)q0 * ldv;
        double acc[DOT_CHUNK];
#pragma unroll
        for (int q = 0; q < DOT_CHUNK; ++q) acc[q] = 0.0;
        // two slice elements per step (their loads in flight together), added
        // in slice order
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        for (; t + stride < n; t += 2 * stride) {
            const double w0 = w[t], w1 = w[t + stride];
            double v0[DOT_CHUNK], v1[DOT_CHUNK];
#pragma unroll
            for (int q = 0; q < DOT_CHUNK; ++q)
                if (q < kc) {
                    v0[q] = Vc[(int64_t)q * ldv + t];
                    v1[q] = Vc[(int64_t)q * ldv + t + stride];
                }
#pragma unroll
            for (int q = 0; q < DOT_CHUNK; ++q)
                if (q < kc) {
                    acc[q] += w0 * v0[q];
                    acc[q] += w1 * v1[q];
                }
        }
        if (t < n) {
            const double wv = w[t];
#pragma unroll
            for (int q = 0; q < DOT_CHUNK; ++q)
                if (q < kc) acc[q] += wv * Vc[(int64_t)q * ldv + t];
        }
#pragma unroll
        for (int q = 0; q < DOT_CHUNK; ++q) {
            const double v = wave_sum(acc[q]);
            if (lane == 0) sh[wid][q] = v;
        }
        __syncthreads();
        if ((int)threadIdx.x < kc) {
            const int q = threadIdx.x;
            partial[(int64_t)(q0 + q) * gridDim.x + blockIdx.x] = ((sh[0][q] + sh[1][q]) + sh[2][q]) + sh[3][q];
        }
        __syncthreads();
    }
}

// lane's fixed-order share of sum_b col[b], then the shuffle tree (lane 0
// holds the total)
__device__ __forceinline__ double column_sum(const double* __restrict__ col, int nblocks, int lane)
{
    double v = 0.0;
    for (int b = lane; b < nblocks; b += 64) v += col[b];
    return wave_sum(v);
}

// out[q] = sum_b partial[q*nblocks + b]: one wavefront per column q (four per
// block), lanes take blocks b = lane, lane + 64, ..., then a fixed shuffle
// tree.  Gate as in the multi-dot.
__global__ void __launch_bounds__(256) reduce_columns_kernel(const double* __restrict__ partial, int nblocks, int k,
                                                            double* __restrict__ out, int accumulate,
                                                            const int* __restrict__ gate)
{
    if (gate && *gate == 0) return;
    const int q = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= k) return;   // whole wavefronts leave together
    const double v = column_sum(partial + (int64_t)q * nblocks, nblocks, lane);
    if (lane == 0) out[q] = accumulate ? out[q] + v : v;
}

// w -= sum_q h[q] V[q]
__global__ void __launch_bounds__(256) update_kernel(double* __restrict__ w, const double* __restrict__ V,
                                                    int64_t ldv, const double* __restrict__ h, int k,
                                                    int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double v = w[t];
    for (int q = 0; q < k; ++q) v -= h[q] * V[(int64_t)q * ldv + t];
    w[t] = v;
}

// x += sum_q y[q] Z[q]
__global__ void __launch_bounds__(256) combine_kernel(double* __restrict__ x, const double* __restrict__ Z,
                                                     int64_t ldz, const double* __restrict__ y, int k,
                                                     int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double v = x[t];
    for (int q = 0; q < k; ++q) v += y[q] * Z[(int64_t)q * ldz + t];
    x[t] = v;
}

__global__ void scale_kernel(double* __restrict__ out, const double* __restrict__ in, const double* nrm,
                             int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = in[t] / nrm[0];
}

// ---- Arnoldi step of the 2-D solve on the device ----
// w -= sum_q h[q] V[q] over the multi-dot's grid-stride slices (one launch of
// RED_BLOCKS blocks), and partial[b] = block b's share of |w|^2 after the
// update, in the slices and summation order of a one-column multi-dot of w
// with itself.  Gate as in the multi-dot.
__global__ void __launch_bounds__(256) update_norm_kernel(double* __restrict__ w, const double* __restrict__ V,
                                                         int64_t ldv, const double* __restrict__ h, int k,
                                                         int64_t n, double* __restrict__ partial,
                                                         const int* __restrict__ gate)
{
    if (gate && *gate == 0) return;
    __shared__ double sh[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        double v = w[t];
        for (int q = 0; q < k; ++q) v -= h[q] * V[(int64_t)q * ldv + t];
        w[t] = v;
        acc += v * v;
    }
    const double v = wave_sum(acc);
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// Device state of one restart cycle: the Hessenberg columns (leading
// dimension ldh), the Givens cosines and sines, the rotated right side g, the
// scalars wn2 (|w|^2 after the first pass) and hn (the new basis norm), the
// second-pass gate, and the per-step status in host-mapped memory:
// status[2j] = |g[j+1]| (the residual norm estimate), status[2j+1] = hn.
struct ArnoldiDev {
    double* H;
    int ldh;
    double *cs, *sn, *g;
    double *wn2, *hn;
    int* gate;
    double* status;
};

// one wavefront: |w|^2 after the first Gram-Schmidt pass and the decision on
// the second ("twice is enough": only when the first pass left no more than
// REORTH_RATIO2 of the squared norm); hd[k+1] holds |w|^2 before the pass
__global__ void __launch_bounds__(64) arnoldi_decide_kernel(const double* __restrict__ npart, int nblocks,
                                                           const double* __restrict__ hd, int k, ArnoldiDev a)
{
    const double v = column_sum(npart, nblocks, threadIdx.x);
    if (threadIdx.x == 0) {
        a.wn2[0] = v;
        a.gate[0] = (v > REORTH_RATIO2 * hd[k + 1]) ? 0 : 1;
    }
}

// one wavefront: Hessenberg column k (both passes' coefficients), the new
// basis norm, the earlier rotations applied to the column, the new rotation,
// the rotated right side and the step's status.  The column and the earlier
// rotations are staged in LDS by all lanes; lane 0 runs the recurrence.
__global__ void __launch_bounds__(64) arnoldi_finish_kernel(const double* __restrict__ npart, int nblocks,
                                                           const double* __restrict__ hd,
                                                           const double* __restrict__ hd2, int k, ArnoldiDev a)
{
    __shared__ double col[MAXK + 1], cs[MAXK], sn[MAXK];
    const bool second = a.gate[0] != 0;
    double wn2 = second ? column_sum(npart, nblocks, threadIdx.x) : 0.0;
    for (int q = threadIdx.x; q <= k; q += 64) {
        col[q] = second ? hd[q] + hd2[q] : hd[q];
        if (q < k) {
            cs[q] = a.cs[q];
            sn[q] = a.sn[q];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (!second) wn2 = a.wn2[0];
        const double hn = sqrt(wn2);
        a.hn[0] = hn;
        col[k + 1] = hn;
        for (int q = 0; q < k; ++q) {
            const double t = cs[q] * col[q] + sn[q] * col[q + 1];
            col[q + 1] = -sn[q] * col[q] + cs[q] * col[q + 1];
            col[q] = t;
        }
        const double d = hypot(col[k], col[k + 1]);
        const double c = col[k] / d, sv = col[k + 1] / d;
        a.cs[k] = c;
        a.sn[k] = sv;
        col[k] = d;
        col[k + 1] = 0.0;
        const double gk = a.g[k];
        a.g[k + 1] = -sv * gk;
        a.g[k] = c * gk;
        a.status[2 * k] = fabs(-sv * gk);
        a.status[2 * k + 1] = hn;
    }
    __syncthreads();
    double* out = a.H + (int64_t)k * a.ldh;
    for (int q = threadIdx.x; q <= k + 1; q += 64) out[q] = col[q];
}

// v /= hn[0] in place, skipped for a zero (or NaN) norm
__global__ void scale_inplace_kernel(double* v, const double* hn, int64_t n)
{
    const double h = hn[0];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && h > 0) v[t] = v[t] / h;
}

// ---- preconditioner packing ----
template <typename T>
__global__ void pack_rhs_kernel(StokesOpDev s, const double* __restrict__ r, MGLevelDev<T> l,
                                T* prnorm_dev, int nc)
{
    // covers the vx grid; RX/RY/RC of level 0 (MG scaling: RC = r_p / pscale),
    // and the level's solution cleared for the V-cycle
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int xnum = s.xnum, ynum = s.ynum, ni = ynum + 1;
    if (t >= ni * (xnum + 1)) return;
    const int I = t % ni + 1, J = t / ni + 1;
    const int i = I - 1, j = J - 1;
    const bool cellv = (i >= 1 && i <= ynum - 1 && j >= 1 && j <= xnum - 1);
    const int c = (j - 1) * (ynum - 1) + (i - 1);
    if (J <= xnum) {
        const size_t k = (size_t)(J - 1) * (ynum + 1) + (I - 1);
        l.RX[k] = cellv ? (T)r[c] : T(0);
        l.vx[k] = T(0);
    }
    if (I <= ynum) {
        const size_t k = (size_t)(J - 1) * ynum + (I - 1);
        l.RY[k] = cellv ? (T)r[nc + c] : T(0);
        l.vy[k] = T(0);
    }
    if (I <= ynum - 1 && J <= xnum - 1) {
        const int cp = (J - 1) * (ynum - 1) + (I - 1);
        l.pr[cp] = T(0);
        l.RC[cp] = (T)(r[2 * (int64_t)nc + cp] / s.pscale);
        if (cp == 0)
            prnorm_dev[0] = (s.pressure_bc_mode == 0) ? (T)(s.pscale * (r[2 * (int64_t)nc] / kbond(s))) : T(0);
    }
}

template <typename T>
__global__ void unpack_corr_kernel(StokesOpDev s, const double* __restrict__ r, MGLevelDev<T> l,
                                   double* __restrict__ z, int nc)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int ny1 = s.ynum - 1;
    const int i = c % ny1 + 1, j = c / ny1 + 1;
    const double kb = kbond(s);
    // vx unknown of cell (i,j) is vx1[i+1, j+1]
    if (j < s.xnum - 1 && out_zone_x(s, i, j))
        z[c] = (double)l.vx[(size_t)j * (s.ynum + 1) + i];
    else
        z[c] = r[c] / kb;
    if (i < s.ynum - 1 && out_zone_y(s, i, j))
        z[nc + c] = (double)l.vy[(size_t)j * s.ynum + i];
    else
        z[nc + c] = r[nc + c] / kb;
    if (internal_continuity(s, i, j))
        z[2 * (int64_t)nc + c] = (double)l.pr[c] / s.pscale;
    else
        z[2 * (int64_t)nc + c] = r[2 * (int64_t)nc + c] / kb;
}

// finest-level nodes whose correction is known exactly (identity rows):
// internal velocity zones and the pressure rows of BC modes 1 and 2
template <typename T>
__global__ void set_fixed_kernel(StokesOpDev s, const double* __restrict__ r, MGLevelDev<T> l, int nc)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int ny1 = s.ynum - 1;
    const int i = c % ny1 + 1, j = c / ny1 + 1;
    const double kb = kbond(s);
    if (l.fix_vx && j < s.xnum - 1 && !out_zone_x(s, i, j))
        l.vx[(size_t)j * (s.ynum + 1) + i] = (T)(r[c] / kb);
    if (l.fix_vy && i < s.ynum - 1 && !out_zone_y(s, i, j))
        l.vy[(size_t)j * s.ynum + i] = (T)(r[nc + c] / kb);
    if (l.fix_p && !internal_continuity(s, i, j))
        l.pr[c] = (T)(s.pscale * (r[2 * (int64_t)nc + c] / kb));
}

// Pressure mode 0 replaces the continuity equation of cell (1,1) by the gauge
// row.  With homogeneous boundary velocities the correction's divergence sums
// to zero over all cells, so the multigrid's continuity right-hand side at
// (1,1) is set to the value that makes the coarse problems compatible:
// RC(1,1) = -sum_{c != (1,1)} RC_c dx dy / (dx_1 dy_1).
template <typename T>
__global__ void __launch_bounds__(256) compat_rc_partial_kernel(MGLevelDev<T> l, double* partial)
{
    __shared__ double sh[256];
    const int ny1 = l.ynum - 1, n = ny1 * (l.xnum - 1);
    double acc = 0.0;
    for (int c = 1 + blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const int i = c % ny1, j = c / ny1;
        acc += (double)l.RC[c] * (double)l.xstp_b[j] * (double)l.ystp_b[i];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

template <typename T>
__global__ void __launch_bounds__(256) compat_rc_finish_kernel(MGLevelDev<T> l, const double* partial,
                                                               int nparts)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) acc += partial[k];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) l.RC[0] = (T)(-sh[0] / ((double)l.xstp_b[0] * (double)l.ystp_b[0]));
}

template <typename T>
__global__ void to_precision_kernel(const double* __restrict__ a, T* __restrict__ b, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (T)a[t];
}


// compute_stokes_residuals! (solver/residuals/StokesResiduals.jl:94-137):
// XStokes.jl:66-117, YStokes.jl:59-112, Continuity.jl:53-84 per node, same
// expression order.  One thread per (i, j) of the (ynum+1) x (xnum+1) index
// space; every node of each output array is written (0 where the reference's
// zeroed array is not updated).
#define GR_M(a, ld, i, j) (a)[((size_t)(j) - 1) * (size_t)(ld) + (size_t)((i) - 1)]
__global__ void __launch_bounds__(256) stokes_grid_residuals_kernel(StokesOpDev s, const double* __restrict__ vx1,
                                                                    const double* __restrict__ vy1,
                                                                    const double* __restrict__ pr1,
                                                                    double* __restrict__ resx,
                                                                    double* __restrict__ resy,
                                                                    double* __restrict__ resc)
{
    const int xnum = s.xnum, ynum = s.ynum;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)(ynum + 1) * (xnum + 1)) return;
    const int i = 1 + (int)(t % (ynum + 1)), j = 1 + (int)(t / (ynum + 1));
    const int lx = ynum + 1, ly = ynum, lp = ynum - 1, le = ynum, ln = ynum - 1;
    if (j < xnum + 1) {
        double r = 0.0;
        if ((j != s.zx_j || i < s.zx_i0 || i > s.zx_i1) && !(i == 1 || i == ynum + 1 || j == 1 || j == xnum)) {
            r = GR_M(s.RX, lx, i, j) -
                (2.0 * (GR_M(s.etan, ln, i - 1, j) * (GR_M(vx1, lx, i, j + 1) - GR_M(vx1, lx, i, j)) / s.xstp[j - 1] -
                        GR_M(s.etan, ln, i - 1, j - 1) * (GR_M(vx1, lx, i, j) - GR_M(vx1, lx, i, j - 1)) / s.xstp[j - 2]) -
                 (GR_M(pr1, lp, i - 1, j) - GR_M(pr1, lp, i - 1, j - 1))) /
                    s.xstpc[j - 1];
            r = r - (GR_M(s.etas, le, i, j) * ((GR_M(vx1, lx, i + 1, j) - GR_M(vx1, lx, i, j)) / s.ystpc[i - 1] +
                                               (GR_M(vy1, ly, i, j + 1) - GR_M(vy1, ly, i, j)) / s.xstpc[j - 1]) -
                     GR_M(s.etas, le, i - 1, j) * ((GR_M(vx1, lx, i, j) - GR_M(vx1, lx, i - 1, j)) / s.ystpc[i - 2] +
                                                   (GR_M(vy1, ly, i - 1, j + 1) - GR_M(vy1, ly, i - 1, j)) / s.xstpc[j - 1])) /
                        s.ystp[i - 2];
        }
        GR_M(resx, lx, i, j) = r;
    }
    if (i < ynum + 1) {
        double r = 0.0;
        if ((j != s.zy_j || i < s.zy_i0 || i > s.zy_i1) && !(i == 1 || i == ynum || j == 1 || j == xnum + 1)) {
            r = GR_M(s.RY, ly, i, j) -
                (2.0 * (GR_M(s.etan, ln, i, j - 1) * (GR_M(vy1, ly, i + 1, j) - GR_M(vy1, ly, i, j)) / s.ystp[i - 1] -
                        GR_M(s.etan, ln, i - 1, j - 1) * (GR_M(vy1, ly, i, j) - GR_M(vy1, ly, i - 1, j)) / s.ystp[i - 2]) -
                 (GR_M(pr1, lp, i, j - 1) - GR_M(pr1, lp, i - 1, j - 1))) /
                    s.ystpc[i - 1];
            r = r - (GR_M(s.etas, le, i, j) * ((GR_M(vy1, ly, i, j + 1) - GR_M(vy1, ly, i, j)) / s.xstpc[j - 1] +
                                               (GR_M(vx1, lx, i + 1, j) - GR_M(vx1, lx, i, j)) / s.ystpc[i - 1]) -
                     GR_M(s.etas, le, i, j - 1) * ((GR_M(vy1, ly, i, j) - GR_M(vy1, ly, i, j - 1)) / s.xstpc[j - 2] +
                                                   (GR_M(vx1, lx, i + 1, j - 1) - GR_M(vx1, lx, i, j - 1)) / s.ystpc[i - 1])) /
                        s.xstp[j - 2];
        }
        GR_M(resy, ly, i, j) = r;
    }
    if (i < ynum && j < xnum)
        GR_M(resc, lp, i, j) = GR_M(s.RC, lp, i, j) -
                               ((GR_M(vx1, lx, i + 1, j + 1) - GR_M(vx1, lx, i + 1, j)) / s.xstp[j - 1] +
                                (GR_M(vy1, ly, i + 1, j + 1) - GR_M(vy1, ly, i, j + 1)) / s.ystp[i - 1]);
}

// Change norms of (v, v_old) over up to 4 segments (MathTools.jl:64-84):
// part[block][4] = sum (v_old - v)^2, sum v^2, max |v_old - v|, max |v|
struct NormSegs {
    int nseg;
    int64_t n[4];
    const double* v[4];
    const double* vo[4];
};
__global__ void __launch_bounds__(256) change_norms_kernel(NormSegs g, double* __restrict__ part)
{
    __shared__ double sh[4][256];
    double sd = 0.0, sv = 0.0, md = 0.0, mv = 0.0;
    for (int q = 0; q < g.nseg; ++q)
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < g.n[q];
             k += (int64_t)gridDim.x * blockDim.x) {
            const double v = g.v[q][k];
            const double d = fabs((g.vo[q] ? g.vo[q][k] : 0.0) - v);
            sd += d * d;
            sv += v * v;
            md = fmax(md, d);
            mv = fmax(mv, fabs(v));
        }
    sh[0][threadIdx.x] = sd;
    sh[1][threadIdx.x] = sv;
    sh[2][threadIdx.x] = md;
    sh[3][threadIdx.x] = mv;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
            sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + o]);
            sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[(size_t)blockIdx.x * 4 + threadIdx.x] = sh[threadIdx.x][0];
}

// max and min of two arrays: part[block][4] = max a, min a, max b, min b
__global__ void __launch_bounds__(256) maxmin2_kernel(const double* __restrict__ a, int64_t na,
                                                      const double* __restrict__ b, int64_t nb,
                                                      double* __restrict__ part)
{
    __shared__ double sh[4][256];
    double r[4] = {-INFINITY, INFINITY, -INFINITY, INFINITY};
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < na; k += (int64_t)gridDim.x * blockDim.x) {
        r[0] = fmax(r[0], a[k]);
        r[1] = fmin(r[1], a[k]);
    }
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += (int64_t)gridDim.x * blockDim.x) {
        r[2] = fmax(r[2], b[k]);
        r[3] = fmin(r[3], b[k]);
    }
    for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = r[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + o]);
            sh[1][threadIdx.x] = fmin(sh[1][threadIdx.x], sh[1][threadIdx.x + o]);
            sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + o]);
            sh[3][threadIdx.x] = fmin(sh[3][threadIdx.x], sh[3][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[(size_t)blockIdx.x * 4 + threadIdx.x] = sh[threadIdx.x][0];
}

// ------------------------------------------------------------------ state
// A captured V-cycle.  The launch sequence of the tiled V-cycle depends only
// on the hierarchy (its device buffers, tiling and coarse solve) and the
// smoother settings, not on the vector being preconditioned (the right-hand
// side is packed into, and the correction unpacked from, the finest level's
// own arrays outside the graph).  It is captured once into a HIP graph on a
// private stream and replayed on the context's stream, which removes the host
// launch gaps between its ~70 short kernels.  The key records everything the
// captured launches read from the host; any change re-captures.  The additive
// smoother swaps each level's two solution buffers after every launch, an
// even number of times per V-cycle, so the host's buffer assignment after a
// V-cycle equals the one before it and a replay leaves the correction where
// the unpacking reads it.
struct VcycleGraph {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
    ~VcycleGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

template <typename T>
static std::vector<uint64_t> vcycle_key(const MGHierarchy<T>& H, const std::vector<int>& sm, const ebg_solve_opts& o,
                                        const T* prn, const T (&bcd)[4])
{
    std::vector<uint64_t> k;
    auto put = [&](const void* p, size_t bytes) {
        const size_t w = (bytes + 7) / 8;
        const size_t at = k.size();
        k.resize(at + w, 0);
        std::memcpy(k.data() + at, p, bytes);
    };
    auto u = [&](uint64_t v) { k.push_back(v); };
    u((uint64_t)(uintptr_t)&H);
    for (const auto& l : H.lev) put(&l, sizeof(l));     // ints and pointers, no padding
    for (const auto& x : H.xfer) put(&x, sizeof(x));    // pointers only
    for (const auto& t : H.tiles) {
        u((uint64_t)t.ti); u((uint64_t)t.tj); u((uint64_t)t.nti); u((uint64_t)t.ntj);
        u((uint64_t)t.lds); u((uint64_t)t.whole_ok); u((uint64_t)t.whole_lds);
    }
    u((uint64_t)H.fast_div); u((uint64_t)H.additive_tiles); u((uint64_t)H.normalize_finest);
    u((uint64_t)(uintptr_t)H.coarse_inv); u((uint64_t)(uintptr_t)H.coarse_b);
    u((uint64_t)(uintptr_t)H.coarse_x); u((uint64_t)H.coarse_n); u((uint64_t)(uintptr_t)H.scratch);
    for (int v : sm) u((uint64_t)(int64_t)v);
    u((uint64_t)o.local_sweeps);
    const T th[2] = {(T)o.theta_s, (T)o.theta_c};
    put(th, sizeof(th));
    put(bcd, sizeof(bcd));
    u((uint64_t)(uintptr_t)prn);
    return k;
}

struct StokesState {
    int xnum = 0, ynum = 0, nc = 0;
    StokesOpDev op{};
    DevBuf etas, etan, rho, RX, RY, RC, bcs;
    // solver workspace
    DevBuf b, x, r, w, V, Z, partial, hdev, ydev, nrm, stage;
    int m_alloc = 0;
    std::unique_ptr<MGHierarchy<float>> Hf;
    std::unique_ptr<MGHierarchy<double>> Hd;
    DevBuf prnorm_f, prnorm_d;
    DevBuf fix_vx, fix_vy, fix_p;
    bool any_fix_v = false, any_fix_p = false;
    // per-precision hierarchy keys: levels, tile edge and grid generation it
    // was built for, and whether the viscosity must be restricted again
    uint64_t grid_gen = 0;  // grid generation the operator was built on
    struct HKey {
        int nlev = -1, tile = -1, coarse = -1;
        uint64_t grid = 0;
        bool visc_dirty = true;
        bool dense_dirty = true;   // the coarsest level's inverse must be rebuilt
        double bcd[4] = {0, 0, 0, 0};
    } key_f, key_d;
    // direct coarsest-level solve: FP64 assembly/LU workspace, pivots, info,
    // the per-precision inverses and coarse vectors, and the solver handle
    DevBuf cA, cW, cipiv, cinfo, cinv_f, cinv_d, cvec_f, cvec_d;
    void* solver = nullptr;
    ~StokesState()
    {
        if (solver) (void)rocblas_destroy_handle((rocblas_handle)solver);
        release_status();
    }
    // learned correction: candidate vector, judge residuals and acceptance count
    DevBuf zc, jr0, jr1, judge, pair_e, pair_r, pair_z;
    // captured V-cycles per precision (ebg_stokes_set_graphs); a failed
    // capture switches them off for this state
    bool use_graphs = true, graph_broken = false;
    VcycleGraph vg_f, vg_d;
    // device Arnoldi state (plain and mixed solves): Hessenberg columns,
    // rotations and right side, the second-pass gate, the per-step status in
    // host-mapped memory, and the events the host waits on one step behind
    DevBuf arn, gate;
    double* status_h = nullptr;
    double* status_d = nullptr;
    hipEvent_t step_done[2] = {nullptr, nullptr};
    int ensure_status()
    {
        if (!status_h) {
            EBG_HIP(hipHostMalloc((void**)&status_h, sizeof(double) * 2 * MAXK,
                                  hipHostMallocMapped | hipHostMallocCoherent));
            EBG_HIP(hipHostGetDevicePointer((void**)&status_d, status_h, 0));
        }
        for (hipEvent_t& e : step_done)
            if (!e) EBG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return EBG_OK;
    }
    void release_status()
    {
        for (hipEvent_t& e : step_done)
            if (e) (void)hipEventDestroy(e);
        if (status_h) (void)hipHostFree(status_h);
        status_h = status_d = nullptr;
    }
};

template <typename T> static VcycleGraph& vgraph(StokesState& st);
template <> VcycleGraph& vgraph<float>(StokesState& st) { return st.vg_f; }
template <> VcycleGraph& vgraph<double>(StokesState& st) { return st.vg_d; }

// the tiled V-cycle, replayed from its captured graph when graphs are on;
// a capture that fails turns graphs off for this state and runs eagerly
template <typename T>
static void vcycle_tiled_graph(StokesState& st, MGHierarchy<T>& H, const ebg_solve_opts& o, const std::vector<int>& sm,
                               const T* prn, const T (&bcd)[4], hipStream_t s)
{
    auto run = [&](hipStream_t q) {
        mg_vcycle_tiled(H, sm, o.local_sweeps, (T)o.theta_s, (T)o.theta_c, T(1), T(1), T(0), prn, bcd, q);
    };
    if (!st.use_graphs) {
        run(s);
        return;
    }
    VcycleGraph& g = vgraph<T>(st);
    std::vector<uint64_t> key = vcycle_key(H, sm, o, prn, bcd);
    if (!g.exec || key != g.key) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        g.key.clear();
        bool ok = g.cap || hipStreamCreateWithFlags(&g.cap, hipStreamNonBlocking) == hipSuccess;
        hipGraph_t graph = nullptr;
        ok = ok && hipStreamBeginCapture(g.cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            run(g.cap);
            ok = hipStreamEndCapture(g.cap, &graph) == hipSuccess && graph;
        }
        ok = ok && hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        if (!ok || vcycle_key(H, sm, o, prn, bcd) != key) {   // capture failed or moved the buffers
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            st.use_graphs = false;
            st.graph_broken = true;
            run(s);
            return;
        }
        g.key = std::move(key);
    }
    if (hipGraphLaunch(g.exec, s) != hipSuccess) {
        (void)hipGetLastError();
        st.use_graphs = false;
        st.graph_broken = true;
        run(s);
    }
}

// The system's operator points into the context's grid arrays, so a system
// is usable only while the grid it was set with is current.
static StokesState* state(ebg_ctx* ctx)
{
    if (!ctx || !ctx->stokes) return nullptr;
    StokesState* st = (StokesState*)ctx->stokes;
    return (ctx->has_grid && st->grid_gen == ctx->grid.generation) ? st : nullptr;
}

static int copy_in(DevBuf& d, const double* src, size_t n, hipStream_t s)
{
    int rc = d.ensure(n * sizeof(double));
    if (rc) return rc;
    EBG_HIP(hipMemcpyAsync(d.p, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    return EBG_OK;
}

// run one V-cycle on r (planar FP64) into z (planar FP64)
template <typename T>
static void precondition(StokesState& st, MGHierarchy<T>& H, T* prn, const ebg_solve_opts& o,
                         const std::vector<int>& sm, const double* r, double* z, hipStream_t s)
{
    MGLevelDev<T>& l0 = H.lev[0];
    const int nvx = (st.ynum + 1) * (st.xnum + 1);
    hipLaunchKernelGGL(pack_rhs_kernel<T>, dim3(nb(nvx)), dim3(256), 0, s, st.op, r, l0, prn, st.nc);
    if (st.op.pressure_bc_mode == 0) {  // deterministic two-stage sum
        hipLaunchKernelGGL(compat_rc_partial_kernel<T>, dim3(COMPAT_BLOCKS), dim3(256), 0, s, l0,
                           st.partial.as<double>());
        hipLaunchKernelGGL(compat_rc_finish_kernel<T>, dim3(1), dim3(256), 0, s, l0,
                           (const double*)st.partial.as<double>(), COMPAT_BLOCKS);
    }
    if (l0.fix_vx || l0.fix_p)
        hipLaunchKernelGGL(set_fixed_kernel<T>, dim3(nb(st.nc)), dim3(256), 0, s, st.op, r, l0, st.nc);
    const T bcd[4] = {(T)o.bc_d[0], (T)o.bc_d[1], (T)o.bc_d[2], (T)o.bc_d[3]};
    if (o.smoother == 1)
        vcycle_tiled_graph(st, H, o, sm, (const T*)prn, bcd, s);
    else
        mg_vcycle(H, sm, (T)o.theta_s, (T)o.theta_c, T(1), T(1), T(0), (const T*)prn, bcd, s);
    hipLaunchKernelGGL(unpack_corr_kernel<T>, dim3(nb(st.nc)), dim3(256), 0, s, st.op, r, l0, z, st.nc);
}

template <typename T>
static int ensure_hierarchy(ebg_ctx* ctx, StokesState& st, std::unique_ptr<MGHierarchy<T>>& H,
                            DevBuf& prn, StokesState::HKey& key, const ebg_solve_opts& o)
{
    if (!H || key.nlev != o.nlevels || key.coarse != o.coarse_unknowns || key.grid != ctx->grid.generation) {
        H.reset(new MGHierarchy<T>());
        std::vector<HostLevelGrid> fine{HostLevelGrid{}};
        const Grid2D& g = ctx->grid;
        HostLevelGrid& h = fine[0];
        h.xnum = (int)g.xnum;
        h.ynum = (int)g.ynum;
        h.gridx_b = g.h_gridx_b;
        h.gridy_b = g.h_gridy_b;
        h.xstp_b = g.h_xstp_b;
        h.ystp_b = g.h_ystp_b;
        h.gridx_vy = g.h_gridx_vy;
        h.xstp_vy = g.h_xstp_vy;
        h.gridy_vx = g.h_gridy_vx;
        h.ystp_vx = g.h_ystp_vx;
        h.gridx_pr.resize(g.xnum - 1);
        h.gridy_pr.resize(g.ynum - 1);
        // pressure nodes (PressureGridTools.jl:34-86)
        h.gridx_pr[0] = (g.h_gridx_b[0] + g.h_gridx_b[1]) / 2.0;
        for (int64_t k = 1; k < g.xnum - 1; ++k)
            h.gridx_pr[k] = h.gridx_pr[k - 1] + (g.h_gridx_b[k + 1] - g.h_gridx_b[k - 1]) / 2.0;
        h.gridy_pr[0] = (g.h_gridy_b[0] + g.h_gridy_b[1]) / 2.0;
        for (int64_t k = 1; k < g.ynum - 1; ++k)
            h.gridy_pr[k] = h.gridy_pr[k - 1] + (g.h_gridy_b[k + 1] - g.h_gridy_b[k - 1]) / 2.0;
        const double xsize = g.h_gridx_b.back() - g.h_gridx_b.front();
        const double ysize = g.h_gridy_b.back() - g.h_gridy_b.front();
        int nl = o.nlevels;
        if (o.coarse_unknowns > 0) {   // stop at the first coarse level small enough for a direct solve
            auto sizes = mg_level_sizes((int)g.xnum, (int)g.ynum, 10);
            if (nl > 0 && nl < (int)sizes.size()) sizes.resize(nl);
            nl = (int)sizes.size();
            for (int k = 1; k < (int)sizes.size(); ++k)
                if (mg_level_unknowns(sizes[k].first, sizes[k].second) <= o.coarse_unknowns) {
                    nl = k + 1;
                    break;
                }
        }
        int rc = mg_build(*H, (int)g.xnum, (int)g.ynum, xsize, ysize, nl, &fine, true);
        if (rc) return rc;
        if ((rc = prn.ensure(sizeof(T) * 4))) return rc;
        key.nlev = o.nlevels;
        key.coarse = o.coarse_unknowns;
        key.dense_dirty = true;
        key.grid = ctx->grid.generation;
        key.tile = -1;
        key.visc_dirty = true;
    }
    if (o.smoother == 1 && key.tile != o.tile) {
        int rc = mg_plan_tiles(*H, o.tile, true, sizeof(T) == 4);
        if (rc) return rc;
        key.tile = o.tile;
    }
    H->additive_tiles = o.additive_tiles != 0;
    return EBG_OK;
}

// Fixed-node masks of the finest level (host-built from the zone indices and
// the pressure BC mode) attached to the hierarchy's level 0.
static int build_masks(StokesState& st)
{
    const StokesOpDev& op = st.op;
    const int nx = st.xnum, ny = st.ynum;
    std::vector<int8_t> mx((size_t)(ny + 1) * nx, 0), my((size_t)ny * (nx + 1), 0), mp((size_t)(ny - 1) * (nx - 1), 0);
    bool fv = false, fp = false;
    for (int j = 1; j <= nx - 1; ++j)
        for (int i = 1; i <= ny - 1; ++i) {
            const bool zx = !(j != op.zx_j || i < op.zx_i0 || i > op.zx_i1);
            const bool zy = !(j != op.zy_j || i < op.zy_i0 || i > op.zy_i1);
            if (zx && j < nx - 1) { mx[(size_t)j * (ny + 1) + i] = 1; fv = true; }
            if (zy && i < ny - 1) { my[(size_t)j * ny + i] = 1; fv = true; }
            bool internal;
            if (op.pressure_bc_mode == 1) internal = (1 < i && i < ny - 1);
            else if (op.pressure_bc_mode == 2) internal = (1 < j && j < nx - 1);
            else internal = true;  // mode 0 keeps the gauge through normalisation
            if (!internal) { mp[(size_t)(j - 1) * (ny - 1) + (i - 1)] = 1; fp = true; }
        }
    int rc;
    if ((rc = st.fix_vx.ensure(mx.size())) || (rc = st.fix_vy.ensure(my.size())) || (rc = st.fix_p.ensure(mp.size())))
        return rc;
    EBG_HIP(hipMemcpy(st.fix_vx.p, mx.data(), mx.size(), hipMemcpyHostToDevice));
    EBG_HIP(hipMemcpy(st.fix_vy.p, my.data(), my.size(), hipMemcpyHostToDevice));
    EBG_HIP(hipMemcpy(st.fix_p.p, mp.data(), mp.size(), hipMemcpyHostToDevice));
    st.any_fix_v = fv;
    st.any_fix_p = fp;
    return EBG_OK;
}

template <typename T>
static void attach_masks(StokesState& st, MGHierarchy<T>& H)
{
    MGLevelDev<T>& l0 = H.lev[0];
    l0.fix_vx = st.any_fix_v ? st.fix_vx.as<int8_t>() : nullptr;
    l0.fix_vy = st.any_fix_v ? st.fix_vy.as<int8_t>() : nullptr;
    l0.fix_p = st.any_fix_p ? st.fix_p.as<int8_t>() : nullptr;
    H.normalize_finest = (st.op.pressure_bc_mode == 0);
}

template <typename T>
static void load_viscosity(StokesState& st, MGHierarchy<T>& H, hipStream_t s)
{
    MGLevelDev<T>& l0 = H.lev[0];
    const int64_t ns = (int64_t)st.ynum * st.xnum, nn = (int64_t)(st.ynum - 1) * (st.xnum - 1);
    hipLaunchKernelGGL(to_precision_kernel<T>, dim3(nb(ns)), dim3(256), 0, s, st.etas.as<double>(), l0.etas, ns);
    hipLaunchKernelGGL(to_precision_kernel<T>, dim3(nb(nn)), dim3(256), 0, s, st.etan.as<double>(), l0.etan, nn);
    mg_restrict_viscosity(H, s);
}

// deterministic dot products of w against V[0..k-1] -> hdev[0..k-1]

static void multidot(StokesState& st, const double* w, const double* V, int64_t ldv, int k, int64_t n,
                     double* out, int accumulate, hipStream_t s)
{
    hipLaunchKernelGGL(multidot_chunks_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, w, V, ldv, k, n,
                       st.partial.as<double>(), (const int*)nullptr);
    hipLaunchKernelGGL(reduce_columns_kernel, dim3((k + 3) / 4), dim3(256), 0, s, st.partial.as<double>(),
                       RED_BLOCKS, k, out, accumulate, (const int*)nullptr);
}

// Direct solve of the coarsest level when it is small enough: the inverse is
// rebuilt after a new viscosity or different ghost mirror factors.
template <typename T>
static int prepare_coarse(StokesState& st, MGHierarchy<T>& H, StokesState::HKey& key, DevBuf& inv,
                          DevBuf& vec, const ebg_solve_opts& o, hipStream_t s)
{
    H.coarse_inv = nullptr;
    const int k = H.nlev() - 1;
    if (o.coarse_unknowns <= 0 || k < 1) return EBG_OK;
    const int64_t n = mg_level_unknowns(H.lev[k].xnum, H.lev[k].ynum);
    if (n > o.coarse_unknowns) return EBG_OK;
    const bool bc_changed = std::memcmp(key.bcd, o.bc_d, sizeof(key.bcd)) != 0;
    int rc;
    if (key.dense_dirty || bc_changed) {
        if (!st.solver) {
            rocblas_handle h = nullptr;
            EBG_CHECK(rocblas_create_handle(&h) == rocblas_status_success, EBG_ERR_HIP,
                      "coarse solve: rocblas_create_handle failed");
            st.solver = h;
        }
        if ((rc = st.cA.ensure((size_t)n * n * sizeof(double))) || (rc = st.cW.ensure((size_t)n * n * sizeof(double))) ||
            (rc = st.cipiv.ensure((size_t)n * sizeof(int))) ||
            (rc = st.cinfo.ensure(sizeof(int))) || (rc = inv.ensure((size_t)n * n * sizeof(T))) ||
            (rc = vec.ensure((size_t)2 * n * sizeof(T))))
            return rc;
        if ((rc = mg_coarse_setup(H, k, o.bc_d, st.cA.as<double>(), st.cW.as<double>(), st.cipiv.as<int>(),
                                  st.cinfo.as<int>(), inv.as<T>(), st.solver, s)))
            return rc;
        std::memcpy(key.bcd, o.bc_d, sizeof(key.bcd));
        key.dense_dirty = false;
    }
    H.coarse_inv = inv.as<T>();
    H.coarse_b = vec.as<T>();
    H.coarse_x = vec.as<T>() + n;
    H.coarse_n = (int)n;
    return EBG_OK;
}

// hierarchy of the requested precision, smoothing counts, viscosity and masks
static int prepare_preconditioner(ebg_ctx* ctx, StokesState& st, const ebg_solve_opts& o, std::vector<int>& sm)
{
    hipStream_t s = ctx->stream;
    int rc;
    if ((rc = st.partial.ensure((size_t)RED_BLOCKS * (MAXK + 1) * 8))) return rc;
    if (o.precision == 1) {
        if ((rc = ensure_hierarchy(ctx, st, st.Hf, st.prnorm_f, st.key_f, o))) return rc;
        sm = mg_smoothing_counts(st.Hf->nlev(), o.finest_sweeps, o.sweep_multiplier, o.sweep_cap);
        if (st.key_f.visc_dirty) {
            load_viscosity(st, *st.Hf, s);
            st.key_f.dense_dirty = true;
        }
        st.key_f.visc_dirty = false;
        attach_masks(st, *st.Hf);
        if ((rc = prepare_coarse(st, *st.Hf, st.key_f, st.cinv_f, st.cvec_f, o, s))) return rc;
    } else {
        if ((rc = ensure_hierarchy(ctx, st, st.Hd, st.prnorm_d, st.key_d, o))) return rc;
        sm = mg_smoothing_counts(st.Hd->nlev(), o.finest_sweeps, o.sweep_multiplier, o.sweep_cap);
        if (st.key_d.visc_dirty) {
            load_viscosity(st, *st.Hd, s);
            st.key_d.dense_dirty = true;
        }
        st.key_d.visc_dirty = false;
        attach_masks(st, *st.Hd);
        if ((rc = prepare_coarse(st, *st.Hd, st.key_d, st.cinv_d, st.cvec_d, o, s))) return rc;
    }
    return EBG_OK;
}

static void apply_preconditioner(StokesState& st, const ebg_solve_opts& o, const std::vector<int>& sm,
                                 const double* r, double* z, hipStream_t s)
{
    if (o.precision == 1) precondition(st, *st.Hf, st.prnorm_f.as<float>(), o, sm, r, z, s);
    else precondition(st, *st.Hd, st.prnorm_d.as<double>(), o, sm, r, z, s);
}

// judge[1], judge[2] = ||v - L z||^2, ||v - L zc||^2; keep zc in z when it is
// smaller and count it in judge[0]
__global__ void learned_select_kernel(double* __restrict__ z, const double* __restrict__ zc, int64_t n,
                                      double* __restrict__ judge)
{
    const bool take = judge[2] < judge[1];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (take && t < n) z[t] = zc[t];
    if (t == 0 && take) judge[0] += 1.0;
}

// The learned correction as a candidate preconditioner output: zc = z + N(...),
// kept only where it lowers the FP64 residual of the preconditioned direction
// (otherwise the plain V-cycle output stays).
static int learned_step(ebg_ctx* ctx, StokesState& st, const double* v, double* z, bool judge,
                        hipStream_t s)
{
    const int nc = st.nc;
    const int64_t n = 3 * (int64_t)nc;
    LayoutPlanar LP{nc};
    int rc = cnn_correct((CnnNet*)ctx->cnn, st.xnum - 1, st.ynum - 1, st.etan.as<double>(), v, z,
                         st.zc.as<double>(), s);
    if (rc) return rc;
    if (!judge) {
        EBG_HIP(hipMemcpyAsync(z, st.zc.p, n * 8, hipMemcpyDeviceToDevice, s));
        return EBG_OK;
    }
    double* jd = st.judge.as<double>();
    hipLaunchKernelGGL(residual_planar_kernel, dim3(nb(nc)), dim3(256), 0, s, st.op, (const double*)z, v,
                       st.jr0.as<double>(), LP);
    hipLaunchKernelGGL(residual_planar_kernel, dim3(nb(nc)), dim3(256), 0, s, st.op,
                       (const double*)st.zc.as<double>(), v, st.jr1.as<double>(), LP);
    multidot(st, st.jr0.as<double>(), st.jr0.as<double>(), 0, 1, n, jd + 1, 0, s);
    multidot(st, st.jr1.as<double>(), st.jr1.as<double>(), 0, 1, n, jd + 2, 0, s);
    hipLaunchKernelGGL(learned_select_kernel, dim3(nb(n)), dim3(256), 0, s, z, (const double*)st.zc.as<double>(),
                       n, jd);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

// One restart cycle of the Arnoldi process with the Hessenberg columns, the
// Givens rotations and the residual estimate on the device (plain and mixed
// solves).  Step j writes A z_j straight into the basis slot j + 1, takes
// |w|^2 along with the Gram-Schmidt dots, fuses the update with the new norm
// and decides the second pass on the device.  The host enqueues step j + 1
// before it waits for step j's status (written by the step's last kernel into
// host-mapped memory), so the device never idles on the host.  A step
// enqueued past the converged one writes only basis slot j + 2, column j + 1
// of H, rotation j + 1 and g[j + 1..j + 2], none of which the solution update
// over the first j + 1 columns reads.  On return kcols is the number of
// columns to combine; g[0] holds the cycle's initial residual norm.  k0 > 0
// continues a cycle whose first k0 steps ran on the host (their Hessenberg
// columns, rotations and g[0..k0] already copied to the device state).
static int arnoldi_cycle_device(StokesState& st, const ebg_solve_opts& o, const std::vector<int>& sm, int64_t n,
                                double bnorm, int& total, int& kcols, hipStream_t s, int k0 = 0)
{
    const int m = o.restart;
    const int nc = st.nc;
    LayoutPlanar LP{nc};
    double* V = st.V.as<double>();
    double* Z = st.Z.as<double>();
    double* hd = st.hdev.as<double>();
    double* part = st.partial.as<double>();
    double* npart = part + (size_t)RED_BLOCKS * MAXK;
    ArnoldiDev a;
    a.H = st.arn.as<double>();
    a.ldh = m + 1;
    a.cs = a.H + (size_t)(m + 1) * m;
    a.sn = a.cs + m;
    a.g = a.sn + m;
    a.wn2 = st.nrm.as<double>() + 2;
    a.hn = st.nrm.as<double>() + 3;
    a.gate = st.gate.as<int>();
    a.status = st.status_d;
    const volatile double* status = st.status_h;
    // 1: converged at step j, 0: go on, -1: non-finite basis norm
    auto check = [&](int j) {
        const double res = status[2 * j], hn = status[2 * j + 1];
        if (!std::isfinite(hn)) {
            set_error("ebg_stokes_solve: preconditioned Krylov vector is non-finite "
                      "(multigrid preconditioner diverged)");
            return -1;
        }
        return (res / bnorm <= o.rtol || hn == 0) ? 1 : 0;
    };
    int k = k0;
    for (; k < m && total < o.max_iterations; ++k) {
        double* vk = V + (int64_t)k * n;
        double* zk = Z + (int64_t)k * n;
        double* w = V + (int64_t)(k + 1) * n;
        apply_preconditioner(st, o, sm, vk, zk, s);
        hipLaunchKernelGGL((stokes_rows_kernel<false, LayoutPlanar>), dim3(nb(nc)), dim3(256), 0, s, st.op, zk, w,
                           LP);
        // classical Gram-Schmidt: <w, V_q> for q <= k and |w|^2 in one pass
        multidot(st, w, V, n, k + 2, n, hd, 0, s);
        hipLaunchKernelGGL(update_norm_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, w, V, n, (const double*)hd, k + 1,
                           n, npart, (const int*)nullptr);
        hipLaunchKernelGGL(arnoldi_decide_kernel, dim3(1), dim3(64), 0, s, (const double*)npart, RED_BLOCKS,
                           (const double*)hd, k, a);
        // the second pass runs only where the gate says so
        hipLaunchKernelGGL(multidot_chunks_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, (const double*)w,
                           (const double*)V, n, k + 1, n, part, (const int*)a.gate);
        hipLaunchKernelGGL(reduce_columns_kernel, dim3((k + 4) / 4), dim3(256), 0, s, (const double*)part,
                           RED_BLOCKS, k + 1, hd + MAXK, 0, (const int*)a.gate);
        hipLaunchKernelGGL(update_norm_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, w, V, n,
                           (const double*)(hd + MAXK), k + 1, n, npart, (const int*)a.gate);
        hipLaunchKernelGGL(arnoldi_finish_kernel, dim3(1), dim3(64), 0, s, (const double*)npart, RED_BLOCKS,
                           (const double*)hd, (const double*)(hd + MAXK), k, a);
        hipLaunchKernelGGL(scale_inplace_kernel, dim3(nb(n)), dim3(256), 0, s, w, (const double*)a.hn, n);
        EBG_HIP(hipGetLastError());
        EBG_HIP(hipEventRecord(st.step_done[k & 1], s));
        ++total;
        if (k > k0) {
            EBG_HIP(hipEventSynchronize(st.step_done[(k - 1) & 1]));
            const int c = check(k - 1);
            if (c < 0) return EBG_ERR_NUMERIC;
            if (c) {            // step k was enqueued past convergence
                --total;
                kcols = k;
                return EBG_OK;
            }
        }
    }
    if (k > k0) {
        EBG_HIP(hipEventSynchronize(st.step_done[(k - 1) & 1]));
        if (check(k - 1) < 0) return EBG_ERR_NUMERIC;
    }
    kcols = k;
    return EBG_OK;
}

// rhs: right-hand side in planar layout, or NULL for the system's R
static int fgmres(ebg_ctx* ctx, StokesState& st, const ebg_solve_opts& o, ebg_solve_stats* stats,
                  const double* rhs)
{
    hipStream_t s = ctx->stream;
    const int nc = st.nc;
    const int64_t n = 3 * (int64_t)nc;
    const int m = o.restart;
    int rc;
    if ((rc = st.b.ensure(n * 8)) || (rc = st.x.ensure(n * 8)) || (rc = st.r.ensure(n * 8)) ||
        (rc = st.w.ensure(n * 8)) || (rc = st.V.ensure((size_t)(m + 1) * n * 8)) ||
        (rc = st.Z.ensure((size_t)m * n * 8)) || (rc = st.partial.ensure((size_t)RED_BLOCKS * (MAXK + 1) * 8)) ||
        (rc = st.hdev.ensure(MAXK * 8 * 2)) || (rc = st.ydev.ensure(MAXK * 8)) || (rc = st.nrm.ensure(64)))
        return rc;
    const bool dev_arnoldi = (o.learned == 0);
    if ((rc = st.arn.ensure(((size_t)(m + 1) * m + 3 * (size_t)m + 1) * 8)) || (rc = st.gate.ensure(64)) ||
        (rc = st.ensure_status()))
        return rc;
    double* const arn_cs = st.arn.as<double>() + (size_t)(m + 1) * m;
    double* const arn_g = arn_cs + 2 * (size_t)m;
    std::vector<int> sm;
    const bool f32 = (o.precision == 1);
    st.use_graphs = ctx->stokes_graphs != 0 && !st.graph_broken;
    if ((rc = prepare_preconditioner(ctx, st, o, sm))) return rc;
    if (o.learned && ((rc = st.zc.ensure(n * 8)) || (rc = st.jr0.ensure(n * 8)) || (rc = st.jr1.ensure(n * 8)) ||
                      (rc = st.judge.ensure(64))))
        return rc;
    if (o.learned) hipLaunchKernelGGL(fill_kernel<double>, dim3(1), dim3(64), 0, s, st.judge.as<double>(), 8, 0.0);
    LayoutPlanar LP{nc};
    double* b = st.b.as<double>();
    double* x = st.x.as<double>();
    double* r = st.r.as<double>();
    double* w = st.w.as<double>();
    double* V = st.V.as<double>();
    double* Z = st.Z.as<double>();
    double* hd = st.hdev.as<double>();
    if (rhs)
        EBG_HIP(hipMemcpyAsync(b, rhs, n * 8, hipMemcpyDeviceToDevice, s));
    else
        hipLaunchKernelGGL((stokes_rows_kernel<true, LayoutPlanar>), dim3(nb(nc)), dim3(256), 0, s, st.op,
                           (const double*)nullptr, b, LP);
    auto norm = [&](const double* v, double* out) {
        multidot(st, v, v, 0, 1, n, out, 0, s);
    };
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), hbuf(2 * MAXK);
    double bnorm2 = 0;
    norm(b, st.nrm.as<double>());
    EBG_HIP(hipMemcpyAsync(&bnorm2, st.nrm.p, 8, hipMemcpyDeviceToHost, s));
    EBG_HIP(hipStreamSynchronize(s));
    const double bnorm = std::sqrt(bnorm2);
    if (!(bnorm > 0) || !std::isfinite(bnorm)) {
        if (stats) { stats->iterations = 0; stats->relres = 0; stats->relres_initial = 0; stats->learned_evaluated = 0; }
        hipLaunchKernelGGL(fill_kernel<double>, dim3(nb(n)), dim3(256), 0, s, x, n, 0.0);
        if (bnorm == 0) return EBG_OK;
        set_error("ebg_stokes_solve: right-hand side is not finite");
        return EBG_ERR_NUMERIC;
    }
    int total = 0, restarts = 0, learned_dirs = 0;
    double relres = 1.0, relres0 = -1.0;
    // learned 1: the judge's acceptance count after the last evaluation, the
    // rejections in a row, and whether the network is still evaluated
    double judged = 0.0;
    int rejects = 0, evaluated = 0;
    bool learned_on = (o.learned == 1 || o.learned == 2);
    while (true) {
        hipLaunchKernelGGL(residual_planar_kernel, dim3(nb(nc)), dim3(256), 0, s, st.op, x, b, r, LP);
        norm(r, st.nrm.as<double>());
        double beta2;
        EBG_HIP(hipMemcpyAsync(&beta2, st.nrm.p, 8, hipMemcpyDeviceToHost, s));
        EBG_HIP(hipStreamSynchronize(s));
        const double beta = std::sqrt(beta2);
        relres = beta / bnorm;
        if (relres0 < 0) relres0 = relres;
        if (!std::isfinite(relres)) {
            set_error("ebg_stokes_solve: residual became non-finite");
            return EBG_ERR_NUMERIC;
        }
        if (relres <= o.rtol || total >= o.max_iterations) break;
        // the network only within the first restart cycle: accepted
        // corrections vary the preconditioner from step to step, and a solve
        // that needs a restart finishes on the plain V-cycle (rift solves
        // stalled for thousands of iterations otherwise)
        if (restarts > 0) learned_on = false;
        EBG_HIP(hipMemcpyAsync(st.nrm.p, &beta, 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(scale_kernel, dim3(nb(n)), dim3(256), 0, s, V, r, st.nrm.as<double>(), n);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        bool done = false, pending = false;
        // plain and mixed solves, and learned ones once the network is off
        // (mode 3 keeps adding its directions), run the cycle on the device
        const bool dev_cycle = dev_arnoldi || (!learned_on && o.learned != 3);
        auto fetch_hg = [&](int kc) -> int {
            if (kc > 0) {
                EBG_HIP(hipMemcpyAsync(H.data(), st.arn.p, sizeof(double) * (size_t)(m + 1) * kc,
                                       hipMemcpyDeviceToHost, s));
                EBG_HIP(hipMemcpyAsync(g.data(), arn_g, sizeof(double) * kc, hipMemcpyDeviceToHost, s));
                EBG_HIP(hipStreamSynchronize(s));
            }
            return EBG_OK;
        };
        if (dev_cycle) {
            EBG_HIP(hipMemcpyAsync(arn_g, &beta, 8, hipMemcpyHostToDevice, s));
            if ((rc = arnoldi_cycle_device(st, o, sm, n, bnorm, total, k, s)) || (rc = fetch_hg(k))) return rc;
        }
        for (; !dev_cycle && k < m && total < o.max_iterations; ++k) {
            double* vk = V + (int64_t)k * n;
            double* zk = Z + (int64_t)k * n;
            const bool learned_dir = pending;
            bool judging = false;
            if (pending) {
                // learned 3: the correction of the previous V-cycle output is
                // an extra flexible direction (any z_k keeps A Z = V H valid)
                rc = cnn_correct((CnnNet*)ctx->cnn, st.xnum - 1, st.ynum - 1, st.etan.as<double>(),
                                 vk - n, zk - n, zk, s);
                if (rc) return rc;
                pending = false;
            } else {
                apply_preconditioner(st, o, sm, vk, zk, s);
                if (learned_on && (rc = learned_step(ctx, st, vk, zk, o.learned == 1, s))) return rc;
                judging = learned_on && o.learned == 1;
                pending = (o.learned == 3);
            }
            hipLaunchKernelGGL((stokes_rows_kernel<false, LayoutPlanar>), dim3(nb(nc)), dim3(256), 0, s,
                               st.op, zk, w, LP);
            // classical Gram-Schmidt; a second pass only when the first one
            // cancelled more than half of the norm ("twice is enough")
            norm(w, st.nrm.as<double>() + 1);
            multidot(st, w, V, n, k + 1, n, hd, 0, s);
            hipLaunchKernelGGL(update_kernel, dim3(nb(n)), dim3(256), 0, s, w, V, n, hd, k + 1, n);
            norm(w, st.nrm.as<double>());
            double wn[2], acc = judged;
            EBG_HIP(hipMemcpyAsync(hbuf.data(), hd, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, s));
            EBG_HIP(hipMemcpyAsync(wn, st.nrm.p, sizeof(wn), hipMemcpyDeviceToHost, s));
            if (judging) EBG_HIP(hipMemcpyAsync(&acc, st.judge.p, 8, hipMemcpyDeviceToHost, s));
            EBG_HIP(hipStreamSynchronize(s));
            double wn2 = wn[0];
            for (int q = 0; q <= k; ++q) H[(size_t)k * (m + 1) + q] = hbuf[q];
            if (!(wn2 > REORTH_RATIO2 * wn[1])) {
                multidot(st, w, V, n, k + 1, n, hd + MAXK, 0, s);
                hipLaunchKernelGGL(update_kernel, dim3(nb(n)), dim3(256), 0, s, w, V, n, hd + MAXK, k + 1, n);
                norm(w, st.nrm.as<double>());
                EBG_HIP(hipMemcpyAsync(hbuf.data() + MAXK, hd + MAXK, sizeof(double) * (k + 1),
                                       hipMemcpyDeviceToHost, s));
                EBG_HIP(hipMemcpyAsync(&wn2, st.nrm.p, 8, hipMemcpyDeviceToHost, s));
                EBG_HIP(hipStreamSynchronize(s));
                for (int q = 0; q <= k; ++q) H[(size_t)k * (m + 1) + q] += hbuf[MAXK + q];
            }
            if (judging) {
                // the plain V-cycle is the fallback: after learned_patience
                // rejections in a row the network is no longer evaluated
                ++evaluated;
                rejects = acc > judged ? 0 : rejects + 1;
                judged = acc;
                if (o.learned_patience > 0 && rejects >= o.learned_patience) learned_on = false;
            }
            // a learned direction that lies (almost) in the current space would
            // turn rounding noise into the next basis vector: drop it
            if (learned_dir) {
                if (!(wn2 > 1e-4 * wn[1])) {
                    --k;
                    continue;
                }
                ++learned_dirs;
            }
            const double hn = std::sqrt(wn2);
            H[(size_t)k * (m + 1) + k + 1] = hn;
            if (!std::isfinite(hn)) {
                set_error("ebg_stokes_solve: preconditioned Krylov vector is non-finite "
                          "(multigrid preconditioner diverged)");
                return EBG_ERR_NUMERIC;
            }
            if (hn > 0) {
                EBG_HIP(hipMemcpyAsync(st.nrm.p, &hn, 8, hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(scale_kernel, dim3(nb(n)), dim3(256), 0, s, V + (int64_t)(k + 1) * n, w,
                                   st.nrm.as<double>(), n);
            }
            // Givens rotations on column k
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
                done = true;
                break;
            }
            // the network is off for the rest of the solve: the cycle's
            // remaining steps run on the device, from the host's state
            if (!learned_on && o.learned != 3 && k + 1 < m && total < o.max_iterations) {
                double* ad = st.arn.as<double>();
                EBG_HIP(hipMemcpyAsync(ad, H.data(), sizeof(double) * (size_t)(m + 1) * (k + 1),
                                       hipMemcpyHostToDevice, s));
                EBG_HIP(hipMemcpyAsync(arn_cs, cs.data(), sizeof(double) * (k + 1), hipMemcpyHostToDevice, s));
                EBG_HIP(hipMemcpyAsync(arn_cs + m, sn.data(), sizeof(double) * (k + 1), hipMemcpyHostToDevice, s));
                EBG_HIP(hipMemcpyAsync(arn_g, g.data(), sizeof(double) * (k + 2), hipMemcpyHostToDevice, s));
                int kc = 0;
                if ((rc = arnoldi_cycle_device(st, o, sm, n, bnorm, total, kc, s, k + 1)) || (rc = fetch_hg(kc)))
                    return rc;
                k = kc;
                break;
            }
        }
        // back substitution and x += Z y
        std::vector<double> y(k);
        for (int q = k - 1; q >= 0; --q) {
            double v = g[q];
            for (int p = q + 1; p < k; ++p) v -= H[(size_t)p * (m + 1) + q] * y[p];
            y[q] = v / H[(size_t)q * (m + 1) + q];
        }
        EBG_HIP(hipMemcpyAsync(st.ydev.p, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(combine_kernel, dim3(nb(n)), dim3(256), 0, s, x, Z, n, st.ydev.as<double>(), k, n);
        ++restarts;
        (void)done;
    }
    if (stats) {
        stats->iterations = total;
        stats->restarts = restarts;
        stats->relres = relres;
        stats->relres_initial = relres0;
        stats->learned_accepted = learned_dirs;
        stats->learned_evaluated = evaluated;
        if (o.learned == 1) {
            double acc = 0;
            EBG_HIP(hipMemcpyAsync(&acc, st.judge.p, 8, hipMemcpyDeviceToHost, s));
            EBG_HIP(hipStreamSynchronize(s));
            stats->learned_accepted = (int)acc;
        }
    }
    (void)f32;
    return relres <= o.rtol ? EBG_OK : EBG_NOT_CONVERGED;
}

}  // namespace ebg

using namespace ebg;

extern "C" {

int ebg_stokes_set_graphs(ebg_ctx* ctx, int on)
{
    EBG_CHECK(ctx != nullptr, EBG_ERR_ARG, "ebg_stokes_set_graphs: ctx is NULL");
    EBG_CHECK(on == 0 || on == 1, EBG_ERR_ARG, "ebg_stokes_set_graphs: on must be 0 or 1");
    ctx->stokes_graphs = on;
    return EBG_OK;
}

int ebg_stokes_release(ebg_ctx* ctx)
{
    if (ctx && ctx->stokes) {
        StokesState* st = (StokesState*)ctx->stokes;
        delete st;
        ctx->stokes = nullptr;
    }
    return EBG_OK;
}

int ebg_stokes_set_system(ebg_ctx* ctx, const double* etas0, const double* etan0, const double* rho1_vy,
                          const double* RX1, const double* RY1, const double* RC1,
                          const double* btopx, const double* btopy, const double* bbottomx,
                          const double* bbottomy, const double* bleftx, const double* blefty,
                          const double* brightx, const double* brighty, const int64_t* bintern_zone,
                          const double* bintern_velocity, int pressure_bc_mode, double pressure_bc,
                          double pscale, double timestep, double gravity_y,
                          int iuse_interface_stabilization)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_stokes_set_system: no grid set");
    EBG_CHECK(etas0 && etan0 && RX1 && RY1 && RC1, EBG_ERR_ARG, "ebg_stokes_set_system: NULL field");
    EBG_CHECK(rho1_vy || iuse_interface_stabilization != 1, EBG_ERR_ARG,
              "ebg_stokes_set_system: rho1_vy required with interface stabilization");
    EBG_CHECK(btopx && btopy && bbottomx && bbottomy && bleftx && blefty && brightx && brighty,
              EBG_ERR_ARG, "ebg_stokes_set_system: NULL boundary array");
    EBG_CHECK(pressure_bc_mode >= 0 && pressure_bc_mode <= 2, EBG_ERR_ARG,
              "ebg_stokes_set_system: pressure_bc_mode must be 0, 1 or 2");
    EBG_CHECK(std::isfinite(pscale) && pscale > 0, EBG_ERR_ARG, "ebg_stokes_set_system: pscale must be > 0");
    if (!ctx->stokes) ctx->stokes = new StokesState();
    StokesState& st = *(StokesState*)ctx->stokes;
    const Grid2D& g = ctx->grid;
    hipStream_t s = ctx->stream;
    const int xnum = (int)g.xnum, ynum = (int)g.ynum;
    if (st.xnum != xnum || st.ynum != ynum) {
        st.Hf.reset();
        st.Hd.reset();
    }
    st.xnum = xnum;
    st.ynum = ynum;
    st.grid_gen = g.generation;
    st.nc = (xnum - 1) * (ynum - 1);
    int rc;
    const size_t ny = ynum, nx = xnum;
    if ((rc = copy_in(st.etas, etas0, ny * nx, s))) return rc;
    if ((rc = copy_in(st.etan, etan0, (ny - 1) * (nx - 1), s))) return rc;
    if (rho1_vy) { if ((rc = copy_in(st.rho, rho1_vy, ny * (nx + 1), s))) return rc; }
    else if ((rc = st.rho.ensure(8))) return rc;
    if ((rc = copy_in(st.RX, RX1, (ny + 1) * nx, s))) return rc;
    if ((rc = copy_in(st.RY, RY1, ny * (nx + 1), s))) return rc;
    if ((rc = copy_in(st.RC, RC1, (ny - 1) * (nx - 1), s))) return rc;
    // boundary arrays: 4 x (nx+1)x2 then 4 x (ny+1)x2, host -> device
    const size_t bx = (nx + 1) * 2, by = (ny + 1) * 2;
    std::vector<double> hb(4 * bx + 4 * by);
    std::memcpy(&hb[0], btopx, bx * 8);
    std::memcpy(&hb[bx], btopy, bx * 8);
    std::memcpy(&hb[2 * bx], bbottomx, bx * 8);
    std::memcpy(&hb[3 * bx], bbottomy, bx * 8);
    std::memcpy(&hb[4 * bx], bleftx, by * 8);
    std::memcpy(&hb[4 * bx + by], blefty, by * 8);
    std::memcpy(&hb[4 * bx + 2 * by], brightx, by * 8);
    std::memcpy(&hb[4 * bx + 3 * by], brighty, by * 8);
    if ((rc = st.bcs.ensure(hb.size() * 8))) return rc;
    EBG_HIP(hipMemcpyAsync(st.bcs.p, hb.data(), hb.size() * 8, hipMemcpyHostToDevice, s));
    EBG_HIP(hipStreamSynchronize(s));
    double* B = st.bcs.as<double>();
    StokesOpDev& op = st.op;
    op.xnum = xnum;
    op.ynum = ynum;
    op.xstp = g.xstp_b;
    op.ystp = g.ystp_b;
    op.xstpc = g.xstp_vy;
    op.ystpc = g.ystp_vx;
    op.xstpavg = g.xstpavg;
    op.ystpavg = g.ystpavg;
    op.etan = st.etan.as<double>();
    op.etas = st.etas.as<double>();
    op.rho1_vy = st.rho.as<double>();
    op.RX = st.RX.as<double>();
    op.RY = st.RY.as<double>();
    op.RC = st.RC.as<double>();
    op.btopx = B;
    op.btopy = B + bx;
    op.bbottomx = B + 2 * bx;
    op.bbottomy = B + 3 * bx;
    op.bleftx = B + 4 * bx;
    op.blefty = B + 4 * bx + by;
    op.brightx = B + 4 * bx + 2 * by;
    op.brighty = B + 4 * bx + 3 * by;
    const int64_t nz[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    const int64_t* z = bintern_zone ? bintern_zone : nz;
    op.zx_j = (int)z[0];
    op.zx_i0 = (int)z[1];
    op.zx_i1 = (int)z[2];
    op.zy_j = (int)z[4];
    op.zy_i0 = (int)z[5];
    op.zy_i1 = (int)z[6];
    op.zvx = bintern_velocity ? bintern_velocity[3] : 0.0;
    op.zvy = bintern_velocity ? bintern_velocity[7] : 0.0;
    op.pressure_bc_mode = pressure_bc_mode;
    op.pressure_bc = pressure_bc;
    op.pscale = pscale;
    op.timestep = timestep;
    op.gravity_y = gravity_y;
    op.stab = iuse_interface_stabilization;
    st.key_f.visc_dirty = true;
    st.key_d.visc_dirty = true;
    return build_masks(st);
}

int ebg_stokes_apply(ebg_ctx* ctx, const double* S, double* LS)
{
    StokesState* st = state(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_stokes_apply: no system set (or the grid changed since ebg_stokes_set_system)");
    EBG_CHECK(S && LS && S != LS, EBG_ERR_ARG, "ebg_stokes_apply: bad vectors");
    hipLaunchKernelGGL((stokes_rows_kernel<false, LayoutInterleaved>), dim3(nb(st->nc)), dim3(256), 0,
                       ctx->stream, st->op, S, LS, LayoutInterleaved{st->nc});
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_stokes_rhs(ebg_ctx* ctx, double* R)
{
    StokesState* st = state(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_stokes_rhs: no system set (or the grid changed since ebg_stokes_set_system)");
    EBG_CHECK(R, EBG_ERR_ARG, "ebg_stokes_rhs: NULL R");
    hipLaunchKernelGGL((stokes_rows_kernel<true, LayoutInterleaved>), dim3(nb(st->nc)), dim3(256), 0,
                       ctx->stream, st->op, (const double*)nullptr, R, LayoutInterleaved{st->nc});
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_stokes_residual(ebg_ctx* ctx, const double* S, double* resnl)
{
    StokesState* st = state(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_stokes_residual: no system set (or the grid changed since ebg_stokes_set_system)");
    EBG_CHECK(S && resnl && S != resnl, EBG_ERR_ARG, "ebg_stokes_residual: bad vectors");
    const int64_t n = 3 * (int64_t)st->nc;
    int rc;
    if ((rc = st->stage.ensure(2 * n * 8))) return rc;
    double* LS = st->stage.as<double>();
    double* R = LS + n;
    LayoutInterleaved L{st->nc};
    hipLaunchKernelGGL((stokes_rows_kernel<false, LayoutInterleaved>), dim3(nb(st->nc)), dim3(256), 0,
                       ctx->stream, st->op, S, LS, L);
    hipLaunchKernelGGL((stokes_rows_kernel<true, LayoutInterleaved>), dim3(nb(st->nc)), dim3(256), 0,
                       ctx->stream, st->op, (const double*)nullptr, R, L);
    hipLaunchKernelGGL(abs_diff_kernel, dim3(nb(n)), dim3(256), 0, ctx->stream, LS, R, resnl, n);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_stokes_unpack(ebg_ctx* ctx, const double* S, double* vx1, double* vy1, double* pr1)
{
    StokesState* st = state(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_stokes_unpack: no system set (or the grid changed since ebg_stokes_set_system)");
    EBG_CHECK(S && vx1 && vy1 && pr1, EBG_ERR_ARG, "ebg_stokes_unpack: NULL array");
    const int xnum = st->xnum, ynum = st->ynum;
    hipLaunchKernelGGL(unpack_kernel, dim3(nb((int64_t)(ynum + 1) * (xnum + 1))), dim3(256), 0,
                       ctx->stream, st->op, S, vx1, vy1, pr1);
    hipLaunchKernelGGL(vel_bc_rows_kernel, dim3(nb(xnum + 1)), dim3(256), 0, ctx->stream, st->op, vx1, vy1);
    hipLaunchKernelGGL(vel_bc_cols_kernel, dim3(nb(ynum + 1)), dim3(256), 0, ctx->stream, st->op, vx1, vy1);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_stokes_solve_opts_default(ebg_solve_opts* o)
{
    EBG_CHECK(o, EBG_ERR_ARG, "ebg_stokes_solve_opts_default: NULL");
    o->precision = 1;
    o->restart = 30;
    o->max_iterations = 2000;
    o->rtol = 1e-8;
    o->nlevels = 0;
    // with the coarsest level solved exactly, short smoothing is enough: two
    // finest-level sweeps, doubling per level, at most eight
    o->finest_sweeps = 2;
    o->sweep_multiplier = 2;
    o->sweep_cap = 8;
    o->theta_s = 0.9;
    o->theta_c = 1.0;
    o->use_initial_guess = 0;
    o->learned = 0;
    for (int k = 0; k < 4; ++k) o->bc_d[k] = 1.0;
    o->smoother = 1;
    o->local_sweeps = 2;  // more sweeps per halo refresh let the tile pressures drift
    o->tile = 16;
    o->flexible = 1;
    o->coarse_unknowns = 6000;
    // levels below the finest: all tiles per launch (a quarter of the
    // launches of the four colours; a few more iterations, less time)
    o->additive_tiles = 1;
    o->learned_patience = 4;
    return EBG_OK;
}

static int solve_impl(ebg_ctx* ctx, const double* R, double* S, const ebg_solve_opts* opts,
                      ebg_solve_stats* stats);

int ebg_stokes_solve(ebg_ctx* ctx, double* S, const ebg_solve_opts* opts, ebg_solve_stats* stats)
{
    return solve_impl(ctx, nullptr, S, opts, stats);
}

int ebg_stokes_solve_rhs(ebg_ctx* ctx, const double* R, double* S, const ebg_solve_opts* opts,
                         ebg_solve_stats* stats)
{
    EBG_CHECK(R, EBG_ERR_ARG, "ebg_stokes_solve_rhs: NULL R");
    return solve_impl(ctx, R, S, opts, stats);
}

static int solve_impl(ebg_ctx* ctx, const double* R, double* S, const ebg_solve_opts* opts,
                      ebg_solve_stats* stats)
{
    StokesState* st = state(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_stokes_solve: no system set (or the grid changed since ebg_stokes_set_system)");
    EBG_CHECK(S, EBG_ERR_ARG, "ebg_stokes_solve: NULL S");
    ebg_solve_opts o;
    if (opts) o = *opts;
    else ebg_stokes_solve_opts_default(&o);
    EBG_CHECK(o.restart >= 1 && o.restart < MAXK - 1, EBG_ERR_ARG, "ebg_stokes_solve: restart must be in [1, 254]");
    EBG_CHECK(o.precision == 0 || o.precision == 1, EBG_ERR_ARG, "ebg_stokes_solve: precision must be 0 or 1");
    EBG_CHECK(o.rtol > 0 && o.max_iterations >= 0, EBG_ERR_ARG, "ebg_stokes_solve: bad tolerance");
    EBG_CHECK(o.finest_sweeps >= 1 && o.sweep_multiplier >= 1, EBG_ERR_ARG, "ebg_stokes_solve: bad sweeps");
    EBG_CHECK(o.smoother == 0 || o.smoother == 1, EBG_ERR_ARG, "ebg_stokes_solve: smoother must be 0 or 1");
    EBG_CHECK(o.smoother == 0 || (o.local_sweeps >= 1 && o.tile >= 4 && o.tile <= 32), EBG_ERR_ARG,
              "ebg_stokes_solve: tiled smoother needs local_sweeps >= 1 and tile in [4, 32]");
    EBG_CHECK(o.learned >= 0 && o.learned <= 3, EBG_ERR_ARG, "ebg_stokes_solve: learned must be 0..3");
    EBG_CHECK(o.additive_tiles == 0 || o.additive_tiles == 1, EBG_ERR_ARG,
              "ebg_stokes_solve: additive_tiles must be 0 or 1");
    EBG_CHECK(o.learned_patience >= 0, EBG_ERR_ARG, "ebg_stokes_solve: learned_patience must be >= 0");
    EBG_CHECK(o.learned == 0 || ctx->cnn, EBG_ERR_STATE,
              "ebg_stokes_solve: learned correction requested but no network (ebg_cnn_create)");
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t n = 3 * (int64_t)st->nc;
    int rc;
    if ((rc = st->x.ensure(n * 8))) return rc;
    if (o.use_initial_guess)
        hipLaunchKernelGGL(interleaved_to_planar_kernel, dim3(nb(st->nc)), dim3(256), 0, ctx->stream, S,
                           st->x.as<double>(), st->nc);
    else
        hipLaunchKernelGGL(fill_kernel<double>, dim3(nb(n)), dim3(256), 0, ctx->stream, st->x.as<double>(), n, 0.0);
    const double* rhs = nullptr;
    if (R) {
        if ((rc = st->pair_r.ensure(n * 8))) return rc;
        hipLaunchKernelGGL(interleaved_to_planar_kernel, dim3(nb(st->nc)), dim3(256), 0, ctx->stream, R,
                           st->pair_r.as<double>(), st->nc);
        rhs = st->pair_r.as<double>();
    }
    rc = fgmres(ctx, *st, o, stats, rhs);
    if (rc < 0) return rc;
    hipLaunchKernelGGL(planar_to_interleaved_kernel, dim3(nb(st->nc)), dim3(256), 0, ctx->stream,
                       st->x.as<double>(), S, st->nc);
    EBG_HIP(hipStreamSynchronize(ctx->stream));
    if (stats) {
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        stats->status = rc;
    }
    return rc;
}

int ebg_stokes_precondition(ebg_ctx* ctx, const double* r, double* z, const ebg_solve_opts* opts)
{
    StokesState* st = state(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_stokes_precondition: no system set (or the grid changed since ebg_stokes_set_system)");
    EBG_CHECK(r && z, EBG_ERR_ARG, "ebg_stokes_precondition: NULL vector");
    ebg_solve_opts o;
    if (opts) o = *opts;
    else ebg_stokes_solve_opts_default(&o);
    EBG_CHECK(o.precision == 0 || o.precision == 1, EBG_ERR_ARG, "ebg_stokes_precondition: bad precision");
    hipStream_t s = ctx->stream;
    const int nc = st->nc;
    const int64_t n = 3 * (int64_t)nc;
    std::vector<int> sm;
    int rc;
    if ((rc = prepare_preconditioner(ctx, *st, o, sm))) return rc;
    if ((rc = st->pair_r.ensure(n * 8)) || (rc = st->pair_z.ensure(n * 8))) return rc;
    hipLaunchKernelGGL(interleaved_to_planar_kernel, dim3(nb(nc)), dim3(256), 0, s, r, st->pair_r.as<double>(), nc);
    apply_preconditioner(*st, o, sm, st->pair_r.as<double>(), st->pair_z.as<double>(), s);
    hipLaunchKernelGGL(planar_to_interleaved_kernel, dim3(nb(nc)), dim3(256), 0, s,
                       (const double*)st->pair_z.as<double>(), z, nc);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_cnn_make_sample(ebg_ctx* ctx, const double* e, const ebg_solve_opts* opts, float* in_img,
                        float* tgt_img)
{
    StokesState* st = state(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_cnn_make_sample: no system set (or the grid changed since ebg_stokes_set_system)");
    EBG_CHECK(ctx->cnn, EBG_ERR_STATE, "ebg_cnn_make_sample: no network (ebg_cnn_create)");
    EBG_CHECK(e && in_img && tgt_img, EBG_ERR_ARG, "ebg_cnn_make_sample: NULL pointer");
    ebg_solve_opts o;
    if (opts) o = *opts;
    else ebg_stokes_solve_opts_default(&o);
    EBG_CHECK(o.precision == 0 || o.precision == 1, EBG_ERR_ARG, "ebg_cnn_make_sample: bad precision");
    hipStream_t s = ctx->stream;
    const int nc = st->nc;
    const int64_t n = 3 * (int64_t)nc;
    std::vector<int> sm;
    int rc;
    if ((rc = prepare_preconditioner(ctx, *st, o, sm))) return rc;
    if ((rc = st->pair_e.ensure(n * 8)) || (rc = st->pair_r.ensure(n * 8)) || (rc = st->pair_z.ensure(n * 8)))
        return rc;
    LayoutPlanar LP{nc};
    double* pe = st->pair_e.as<double>();
    double* pr = st->pair_r.as<double>();
    double* pz = st->pair_z.as<double>();
    hipLaunchKernelGGL(interleaved_to_planar_kernel, dim3(nb(nc)), dim3(256), 0, s, e, pe, nc);
    hipLaunchKernelGGL((stokes_rows_kernel<false, LayoutPlanar>), dim3(nb(nc)), dim3(256), 0, s, st->op,
                       (const double*)pe, pr, LP);
    apply_preconditioner(*st, o, sm, pr, pz, s);
    rc = cnn_pack_sample((CnnNet*)ctx->cnn, st->xnum - 1, st->ynum - 1, st->etan.as<double>(), pr, pz, pe,
                         in_img, tgt_img, s);
    if (rc) return rc;
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}


int ebg_compute_stokes_residuals(ebg_ctx* ctx, const double* vx1, const double* vy1, const double* pr1,
                                 double* resx1, double* resy1, double* resc1)
{
    StokesState* st = state(ctx);
    EBG_CHECK(st, EBG_ERR_STATE, "ebg_compute_stokes_residuals: no system set (or the grid changed since ebg_stokes_set_system)");
    EBG_CHECK(vx1 && vy1 && pr1 && resx1 && resy1 && resc1, EBG_ERR_ARG,
              "ebg_compute_stokes_residuals: NULL array");
    const int64_t n = (int64_t)(st->ynum + 1) * (st->xnum + 1);
    hipLaunchKernelGGL(stokes_grid_residuals_kernel, dim3(nb(n)), dim3(256), 0, ctx->stream, st->op, vx1, vy1,
                       pr1, resx1, resy1, resc1);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

constexpr int NORM_BLOCKS = 256;

// blocks' partials reduced on the host in block order (deterministic)
static int device_partials(ebg_ctx* ctx, std::vector<double>& h, DevBuf& buf)
{
    EBG_HIP(hipMemcpyAsync(h.data(), buf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    EBG_HIP(hipStreamSynchronize(ctx->stream));
    return EBG_OK;
}

int ebg_change_norms(ebg_ctx* ctx, int nseg, const int64_t* n, const double* const* v,
                     const double* const* v_old, double* out)
{
    EBG_CHECK(ctx && out && n && v, EBG_ERR_ARG, "ebg_change_norms: NULL argument");
    EBG_CHECK(nseg >= 1 && nseg <= 4, EBG_ERR_ARG, "ebg_change_norms: nseg must be 1..4");
    NormSegs g{};
    g.nseg = nseg;
    for (int q = 0; q < nseg; ++q) {
        EBG_CHECK(n[q] >= 0 && (v[q] || n[q] == 0), EBG_ERR_ARG, "ebg_change_norms: bad segment");
        g.n[q] = n[q];
        g.v[q] = v[q];
        g.vo[q] = v_old ? v_old[q] : nullptr;
    }
    static thread_local DevBuf buf;
    int rc = buf.ensure(NORM_BLOCKS * 4 * sizeof(double));
    if (rc) return rc;
    hipLaunchKernelGGL(change_norms_kernel, dim3(NORM_BLOCKS), dim3(256), 0, ctx->stream, g, buf.as<double>());
    EBG_HIP(hipGetLastError());
    std::vector<double> h(NORM_BLOCKS * 4);
    if ((rc = device_partials(ctx, h, buf))) return rc;
    double sd = 0, sv = 0, md = 0, mv = 0;
    for (int b = 0; b < NORM_BLOCKS; ++b) {
        sd += h[4 * b];
        sv += h[4 * b + 1];
        md = std::max(md, h[4 * b + 2]);
        mv = std::max(mv, h[4 * b + 3]);
    }
    const double l2d = std::sqrt(sd), l2v = std::sqrt(sv);
    out[0] = l2d;
    out[1] = l2v != 0.0 ? l2d / l2v : 1e38;
    out[2] = md;
    out[3] = mv != 0.0 ? md / mv : 1e38;
    return EBG_OK;
}

int ebg_calculate_new_timestep(ebg_ctx* ctx, const double* vx1, const double* vy1, double timestep_old,
                               double marker_cell_displ_max, double* timestep_new, int* updated)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_calculate_new_timestep: no grid set");
    EBG_CHECK(vx1 && vy1 && timestep_new, EBG_ERR_ARG, "ebg_calculate_new_timestep: NULL argument");
    const Grid2D& gr = ctx->grid;
    static thread_local DevBuf buf;
    int rc = buf.ensure(NORM_BLOCKS * 4 * sizeof(double));
    if (rc) return rc;
    hipLaunchKernelGGL(maxmin2_kernel, dim3(NORM_BLOCKS), dim3(256), 0, ctx->stream, vx1, (gr.ynum + 1) * gr.xnum,
                       vy1, gr.ynum * (gr.xnum + 1), buf.as<double>());
    EBG_HIP(hipGetLastError());
    std::vector<double> h(NORM_BLOCKS * 4);
    if ((rc = device_partials(ctx, h, buf))) return rc;
    double mxa = -INFINITY, mna = INFINITY, mxb = -INFINITY, mnb = INFINITY;
    for (int b = 0; b < NORM_BLOCKS; ++b) {
        mxa = std::max(mxa, h[4 * b]);
        mna = std::min(mna, h[4 * b + 1]);
        mxb = std::max(mxb, h[4 * b + 2]);
        mnb = std::min(mnb, h[4 * b + 3]);
    }
    // time_step/update/Update.jl:71-98, 198-218
    const double vxmax = std::max(std::fabs(mxa), std::fabs(mna));
    const double vymax = std::max(std::fabs(mxb), std::fabs(mnb));
    double dt = timestep_old;
    if (vxmax > 0) dt = std::min(dt, marker_cell_displ_max * gr.xstpavg / vxmax);
    if (vymax > 0) dt = std::min(dt, marker_cell_displ_max * gr.ystpavg / vymax);
    *timestep_new = dt;
    if (updated) *updated = dt != timestep_old;
    return EBG_OK;
}

}  // extern "C"
