// Direct solve of the coarsest multigrid level of the Stokes preconditioner.
//
// The tiled Gauss-Seidel smoother needs dozens of sweeps on the coarse levels
// and each sweep there is a chain of barrier-bound phases on one or a few
// workgroups.  Replacing the bottom of the hierarchy by an exact solve turns
// that chain into one matrix-vector product: the level's operator is assembled
// column by column from the smoother's own residual arithmetic (so ghost rules,
// mirror factors and steps are exactly the smoother's), inverted once per
// system, and applied as a dense row-major GEMV, one wavefront per row.
//
// The inverse comes from a blocked Gauss-Jordan sweep (128-unknown pivot
// blocks): each step inverts its diagonal block in one workgroup's LDS with
// partial pivoting inside the block, and the rest of the step is three
// rocBLAS GEMMs, so the n sequential pivot steps cost a barrier each instead
// of the three kernel launches per column of an unblocked LU panel.  There is
// no pivoting across blocks; a zero or non-finite block pivot sends the system
// to rocSOLVER (LU with partial pivoting, then A^T X = I).
#include "ebg_common.hpp"
#include "ebg_mg.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace ebg {

namespace {

// The value field x = e_c (one unknown set to one) after the smoother's ghost
// pass (mg_bc_kernel): interior unknowns, top/bottom vx ghosts mirrored with
// d_top/d_bottom, left/right vy ghosts with d_left/d_right, wall nodes zero.
// Right-hand sides read zero, so a residual evaluated on this view is minus
// column c of the operator.  Viscosity and steps come from level l in double.
template <typename S>
struct UnitView {
    static constexpr bool fast_div = false;
    const MGLevelDev<S>& l;
    int kind, ci, cj;   // 0 vx, 1 vy, 2 pr; 1-based node
    double dt, db, dl, dr;
    __device__ double vx(int i, int j) const
    {
        if (kind != 0 || j != cj || j < 2 || j > l.xnum - 1) return 0.0;
        if (i == ci) return 1.0;
        if (i == 1 && ci == 2) return dt;
        if (i == l.ynum + 1 && ci == l.ynum) return db;
        return 0.0;
    }
    __device__ double vy(int i, int j) const
    {
        if (kind != 1 || i != ci || i < 2 || i > l.ynum - 1) return 0.0;
        if (j == cj) return 1.0;
        if (j == 1 && cj == 2) return dl;
        if (j == l.xnum + 1 && cj == l.xnum) return dr;
        return 0.0;
    }
    __device__ double pr(int i, int j) const { return (kind == 2 && i == ci && j == cj) ? 1.0 : 0.0; }
    __device__ double es(int i, int j) const { return (double)MG_ES(l, i, j); }
    __device__ double en(int i, int j) const { return (double)MG_EN(l, i, j); }
    __device__ double RX(int, int) const { return 0.0; }
    __device__ double RY(int, int) const { return 0.0; }
    __device__ double RC(int, int) const { return 0.0; }
    __device__ double xb(int k) const { return (double)l.xstp_b[k]; }
    __device__ double yb(int k) const { return (double)l.ystp_b[k]; }
    __device__ double xv(int k) const { return (double)l.xstp_vy[k]; }
    __device__ double yv(int k) const { return (double)l.ystp_vx[k]; }
};

struct UnknownMap {
    int xnum, ynum, nvx, nvy;
    __device__ __host__ int index(int kind, int i, int j) const
    {
        if (kind == 0) {
            if (i < 2 || i > ynum || j < 2 || j > xnum - 1) return -1;
            return (j - 2) * (ynum - 1) + (i - 2);
        }
        if (kind == 1) {
            if (i < 2 || i > ynum - 1 || j < 2 || j > xnum) return -1;
            return nvx + (j - 2) * (ynum - 2) + (i - 2);
        }
        if (i < 1 || i > ynum - 1 || j < 1 || j > xnum - 1) return -1;
        return nvx + nvy + (j - 1) * (ynum - 1) + (i - 1);
    }
    __device__ __host__ void node(int u, int& kind, int& i, int& j) const
    {
        if (u < nvx) {
            kind = 0;
            i = 2 + u % (ynum - 1);
            j = 2 + u / (ynum - 1);
        } else if (u < nvx + nvy) {
            u -= nvx;
            kind = 1;
            i = 2 + u % (ynum - 2);
            j = 2 + u / (ynum - 2);
        } else {
            u -= nvx + nvy;
            kind = 2;
            i = 1 + u % (ynum - 1);
            j = 1 + u / (ynum - 1);
        }
    }
};

inline UnknownMap unknown_map(int xnum, int ynum)
{
    return UnknownMap{xnum, ynum, (ynum - 1) * (xnum - 2), (ynum - 2) * (xnum - 1)};
}

constexpr int CAND = 75;   // candidate rows per column: 3 grids x 5 x 5 offsets

// A[row, c] = -residual_row(e_c) for every row within two nodes of unknown c
// (the stencils and ghost mirrors reach no farther); other entries stay zero.
template <typename S>
__global__ void __launch_bounds__(256) coarse_assemble_kernel(MGLevelDev<S> l, UnknownMap m, int n, double dt,
                                                              double db, double dl, double dr,
                                                              double* __restrict__ A)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * CAND) return;
    const int c = (int)(t / CAND), q = (int)(t % CAND);
    int kind, ci, cj;
    m.node(c, kind, ci, cj);
    const int rk = q / 25, di = (q % 25) / 5 - 2, dj = q % 5 - 2;
    const int ri = ci + di, rj = cj + dj;
    const int row = m.index(rk, ri, rj);
    if (row < 0) return;
    const UnitView<S> v{l, kind, ci, cj, dt, db, dl, dr};
    double r;
    if (rk == 0) r = vx_residual_v<double>(v, ri, rj, (double*)nullptr);
    else if (rk == 1) r = vy_residual_v<double>(v, ri, rj, (double*)nullptr);
    else r = pr_residual_v<double>(v, ri, rj);
    A[(size_t)c * n + row] = -r;
}

// continuity row of cell (1,1) -> p(1,1) = 0
__global__ void coarse_gauge_kernel(double* __restrict__ A, int n, int g)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) A[(size_t)c * n + g] = (c == g) ? 1.0 : 0.0;
}

__global__ void zero_kernel(double* __restrict__ a, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = 0.0;
}

__global__ void identity_kernel(double* __restrict__ a, int n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (int64_t)n * n) a[t] = (t % (n + 1) == 0) ? 1.0 : 0.0;
}

template <typename T>
__global__ void to_type_kernel(const double* __restrict__ a, T* __restrict__ b, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (T)a[t];
}

// ---- blocked Gauss-Jordan inverse (column-major A, leading dimension n) ----
constexpr int GJ_NB = 128;
constexpr int GJ_MIN_N = 4 * GJ_NB;  // smaller systems go to rocSOLVER directly

// R (b x n, ld b) = rows k0..k0+b-1 of A
__global__ void gj_rows_kernel(const double* __restrict__ A, int n, int k0, int b, double* __restrict__ R)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)b * n) return;
    const int r = (int)(t % b);
    const int64_t j = t / b;
    R[t] = A[(size_t)j * n + k0 + r];
}

// Inverse of the b x b pivot block P = C[0..b-1, 0..b-1] (ld ldc) by
// Gauss-Jordan with partial pivoting inside the block, the swept array M
// held in registers: wavefront w owns columns w, w + 16, ..., lane l rows l
// and l + 64.  Pivot rows are chosen in place (no interchanges): step j
// sweeps on (piv[j], j), so M ends as M[piv[r]][q] = P^-1[r][piv[q]].  The
// owner wavefront of column j picks the pivot and publishes the column in
// LDS; every wavefront takes row p from the owning lane's registers, so a
// step costs one barrier.  Sets *status when a pivot is zero or not finite.
constexpr int GJ_WAVES = 16;
constexpr int GJ_COLS = GJ_NB / GJ_WAVES;  // columns per wavefront

__device__ __forceinline__ double gj_readlane(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__global__ void __launch_bounds__(1024) gj_block_inverse_kernel(const double* __restrict__ C, int ldc, int b,
                                                                double* __restrict__ Pinv, int* status)
{
    __shared__ double colbuf[2][GJ_NB];
    __shared__ int pbuf[2];
    __shared__ int piv[GJ_NB], rowof[GJ_NB];
    const int tid = threadIdx.x;
    const int w = tid >> 6, l = tid & 63;
    double m[GJ_COLS][2];
#pragma unroll
    for (int k = 0; k < GJ_COLS; ++k)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = l + 64 * h, c = w + GJ_WAVES * k;
            m[k][h] = (i < b && c < b) ? C[(size_t)c * ldc + i] : 0.0;
        }
    // rows already used as pivots (padding rows count as used)
    bool used0 = l >= b, used1 = l + 64 >= b;
    for (int j = 0; j < b; ++j) {
        const int par = j & 1;
        const int owner = j % GJ_WAVES, kj = j / GJ_WAVES;
        if (w == owner) {
            // largest |M[r][j]| over the rows not yet pivoted; the key keeps
            // the magnitude's high bits and the row in the low seven
            double v0 = 0.0, v1 = 0.0;
#pragma unroll
            for (int k = 0; k < GJ_COLS; ++k)
                if (k == kj) {
                    v0 = m[k][0];
                    v1 = m[k][1];
                }
            unsigned long long key = 0;
            if (!used0)
                key = ((unsigned long long)__double_as_longlong(fabs(v0)) & ~0x7full) | (unsigned)(GJ_NB - 1 - l);
            if (!used1) {
                const unsigned long long k1 =
                    ((unsigned long long)__double_as_longlong(fabs(v1)) & ~0x7full) | (unsigned)(GJ_NB - 1 - (l + 64));
                key = k1 > key ? k1 : key;
            }
            for (int d = 32; d >= 1; d >>= 1) {
                const unsigned long long o = (unsigned long long)__shfl_xor((long long)key, d, 64);
                key = o > key ? o : key;
            }
            colbuf[par][l] = v0;
            colbuf[par][l + 64] = v1;
            if (l == 0) {
                const int p = GJ_NB - 1 - (int)(key & 0x7full);
                pbuf[par] = p;
                piv[j] = p;
                rowof[p] = j;
            }
        }
        __syncthreads();
        const int p = pbuf[par];
        const double d = colbuf[par][p];
        if (tid == 0 && (d == 0.0 || !isfinite(d))) *status = 1;
        const double rd = 1.0 / d;
        const double pij0 = colbuf[par][l], pij1 = colbuf[par][l + 64];
        const int hp = p >> 6, lp = p & 63;
        // row p of this wavefront's columns, scaled: M[p][c] / d
        double rpr[GJ_COLS];
        if (hp) {
#pragma unroll
            for (int k = 0; k < GJ_COLS; ++k) rpr[k] = gj_readlane(m[k][1], lp) * rd;
        } else {
#pragma unroll
            for (int k = 0; k < GJ_COLS; ++k) rpr[k] = gj_readlane(m[k][0], lp) * rd;
        }
        // every entry: M_ic -= M_ij (M_pc / d); row p and column j are
        // rewritten below
#pragma unroll
        for (int k = 0; k < GJ_COLS; ++k) {
            m[k][0] = fma(-pij0, rpr[k], m[k][0]);
            m[k][1] = fma(-pij1, rpr[k], m[k][1]);
        }
        if (l == lp) {  // row p: M_pc / d
#pragma unroll
            for (int k = 0; k < GJ_COLS; ++k) {
                if (hp) m[k][1] = rpr[k];
                else m[k][0] = rpr[k];
            }
        }
        if (w == owner) {  // column j: -M_ij / d, pivot 1 / d
#pragma unroll
            for (int k = 0; k < GJ_COLS; ++k)
                if (k == kj) {
                    m[k][0] = (l == p) ? rd : -pij0 * rd;
                    m[k][1] = (l + 64 == p) ? rd : -pij1 * rd;
                }
        }
        if (l == lp) {
            if (hp) used1 = true;
            else used0 = true;
        }
    }
    __syncthreads();
    // P^-1[r][q] = M[piv[r]][rowof[q]]: entry (i, c) of M goes to
    // row rowof[i], column piv[c]
#pragma unroll
    for (int k = 0; k < GJ_COLS; ++k)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = l + 64 * h, c = w + GJ_WAVES * k;
            if (i < b && c < b) Pinv[(size_t)piv[c] * b + rowof[i]] = m[k][h];
        }
}

// Rows k0..k0+b-1 of A: the scaled row block Rp (b x n, ld b), and the pivot
// block's inverse on the diagonal.
__global__ void gj_put_rows_kernel(double* __restrict__ A, int n, int k0, int b, const double* __restrict__ Rp,
                                   const double* __restrict__ Pinv)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)b * n) return;
    const int r = (int)(t % b);
    const int j = (int)(t / b);
    const double v = (j >= k0 && j < k0 + b) ? Pinv[(size_t)(j - k0) * b + r] : Rp[t];
    A[(size_t)j * n + k0 + r] = v;
}

// inv (row-major, type T) = A (column-major), through 32 x 32 LDS tiles
template <typename T>
__global__ void __launch_bounds__(256) transpose_to_type_kernel(const double* __restrict__ A, int n,
                                                                 T* __restrict__ inv)
{
    __shared__ double tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // bx: column of A, by: row of A
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int q = ty; q < 32; q += 8) {
        const int row = by + tx, col = bx + q;
        tile[q][tx] = (row < n && col < n) ? A[(size_t)col * n + row] : 0.0;
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        const int row = by + q, col = bx + tx;
        if (row < n && col < n) inv[(size_t)row * n + col] = (T)tile[tx][q];
    }
}

template <typename T>
__global__ void coarse_pack_kernel(MGLevelDev<T> l, UnknownMap m, int n, T* __restrict__ b)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    int kind, i, j;
    m.node(u, kind, i, j);
    T v;
    if (kind == 0) v = l.RX[(size_t)(j - 1) * (l.ynum + 1) + (i - 1)];
    else if (kind == 1) v = l.RY[(size_t)(j - 1) * l.ynum + (i - 1)];
    else v = (i == 1 && j == 1) ? T(0) : l.RC[(size_t)(j - 1) * (l.ynum - 1) + (i - 1)];
    b[u] = v;
}

template <typename T>
__global__ void coarse_unpack_kernel(MGLevelDev<T> l, UnknownMap m, int n, const T* __restrict__ x)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    int kind, i, j;
    m.node(u, kind, i, j);
    if (kind == 0) MG_VX(l, i, j) = x[u];
    else if (kind == 1) MG_VY(l, i, j) = x[u];
    else MG_PR(l, i, j) = x[u];
}

// x = inv * b, one wavefront per row of the row-major inverse
template <typename T>
__global__ void __launch_bounds__(256) coarse_gemv_kernel(const T* __restrict__ inv, const T* __restrict__ b,
                                                          T* __restrict__ x, int n)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const T* a = inv + (size_t)row * n;
    // four independent partial sums per lane (four row segments of 64 in
    // flight per step), added in a fixed order
    T acc0 = T(0), acc1 = T(0), acc2 = T(0), acc3 = T(0);
    int j = lane;
    for (; j + 192 < n; j += 256) {
        acc0 += a[j] * b[j];
        acc1 += a[j + 64] * b[j + 64];
        acc2 += a[j + 128] * b[j + 128];
        acc3 += a[j + 192] * b[j + 192];
    }
    for (; j < n; j += 64) acc0 += a[j] * b[j];
    T acc = (acc0 + acc1) + (acc2 + acc3);
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
    if (lane == 0) x[row] = acc;
}

}  // namespace

// The coarsest level's operator in column-major A (n x n) with the gauge row.
template <typename T>
static int coarse_assemble(const MGLevelDev<T>& l, const UnknownMap& m, int n, const double* bc_d, double* A,
                           hipStream_t s)
{
    const int64_t nn = (int64_t)n * n;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, A, nn);
    const int64_t pairs = (int64_t)n * CAND;
    hipLaunchKernelGGL(coarse_assemble_kernel<T>, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, s, l, m, n,
                       bc_d[0], bc_d[1], bc_d[2], bc_d[3], A);
    hipLaunchKernelGGL(coarse_gauge_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A, n,
                       m.index(2, 1, 1));
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

// Blocked Gauss-Jordan: A (column-major) becomes A^-1.  work holds the
// column block C (n x NB), the row blocks R and Rp (NB x n) and the pivot
// block inverse (NB x NB); *status is set on a bad block pivot.
static int gauss_jordan_inverse(rocblas_handle h, double* A, int n, double* work, int* status, hipStream_t s)
{
    double* C = work;
    double* R = C + (size_t)n * GJ_NB;
    double* Rp = R + (size_t)n * GJ_NB;
    double* Pinv = Rp + (size_t)n * GJ_NB;
    const double one = 1.0, mone = -1.0, zero = 0.0;
    EBG_HIP(hipMemsetAsync(status, 0, sizeof(int), s));
    for (int k0 = 0; k0 < n; k0 += GJ_NB) {
        const int b = std::min(GJ_NB, n - k0);
        const int64_t nb_n = (int64_t)b * n;
        const unsigned g = (unsigned)((nb_n + 255) / 256);
        EBG_HIP(hipMemcpyAsync(C, A + (size_t)k0 * n, sizeof(double) * nb_n, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(gj_rows_kernel, dim3(g), dim3(256), 0, s, (const double*)A, n, k0, b, R);
        hipLaunchKernelGGL(gj_block_inverse_kernel, dim3(1), dim3(1024), 0, s,
                           (const double*)(C + k0), n, b, Pinv, status);
        EBG_HIP(hipGetLastError());
        // Rp = Pinv R; A -= C Rp; A[:, K] = -C Pinv; rows K = Rp with Pinv on the diagonal block
        EBG_CHECK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, b, n, b, &one, Pinv, b, R, b,
                                &zero, Rp, b) == rocblas_status_success,
                  EBG_ERR_HIP, "coarse solve: rocblas_dgemm (row block) failed");
        EBG_CHECK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, b, &mone, C, n, Rp, b,
                                &one, A, n) == rocblas_status_success,
                  EBG_ERR_HIP, "coarse solve: rocblas_dgemm (update) failed");
        EBG_CHECK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, b, b, &mone, C, n, Pinv, b,
                                &zero, A + (size_t)k0 * n, n) == rocblas_status_success,
                  EBG_ERR_HIP, "coarse solve: rocblas_dgemm (column block) failed");
        hipLaunchKernelGGL(gj_put_rows_kernel, dim3(g), dim3(256), 0, s, A, n, k0, b, (const double*)Rp,
                           (const double*)Pinv);
        EBG_HIP(hipGetLastError());
    }
    return EBG_OK;
}

template <typename T>
int mg_coarse_setup(const MGHierarchy<T>& H, int k, const double* bc_d, double* A, double* work, int* ipiv,
                    int* info, T* inv, void* solver_handle, hipStream_t s)
{
    const MGLevelDev<T>& l = H.lev[k];
    const UnknownMap m = unknown_map(l.xnum, l.ynum);
    const int n = (int)mg_level_unknowns(l.xnum, l.ynum);
    const int64_t nn = (int64_t)n * n;
    int rc;
    rocblas_handle h = (rocblas_handle)solver_handle;
    EBG_CHECK(rocblas_set_stream(h, s) == rocblas_status_success, EBG_ERR_HIP, "coarse solve: rocblas_set_stream");
    if ((rc = coarse_assemble(l, m, n, bc_d, A, s))) return rc;
    if (n >= GJ_MIN_N) {
        if ((rc = gauss_jordan_inverse(h, A, n, work, info, s))) return rc;
        int hstatus = 0;
        EBG_HIP(hipMemcpyAsync(&hstatus, info, sizeof(int), hipMemcpyDeviceToHost, s));
        EBG_HIP(hipStreamSynchronize(s));
        if (hstatus == 0) {
            const dim3 tg((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
            hipLaunchKernelGGL(transpose_to_type_kernel<T>, tg, dim3(256), 0, s, (const double*)A, n, inv);
            EBG_HIP(hipGetLastError());
            return EBG_OK;
        }
        // a block pivot vanished: assemble again and factor with pivoting
        if ((rc = coarse_assemble(l, m, n, bc_d, A, s))) return rc;
    }
    EBG_CHECK(rocsolver_dgetrf(h, n, n, A, n, ipiv, info) == rocblas_status_success, EBG_ERR_HIP,
              "coarse solve: rocsolver_dgetrf failed");
    int hinfo = 0;
    EBG_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
    EBG_HIP(hipStreamSynchronize(s));
    EBG_CHECK(hinfo == 0, EBG_ERR_NUMERIC, "coarse solve: coarsest-level operator is singular");
    // X = A^-T by solving A^T X = I: column-major A^-T is the row-major
    // inverse the GEMV reads
    hipLaunchKernelGGL(identity_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, work, n);
    EBG_CHECK(rocsolver_dgetrs(h, rocblas_operation_transpose, n, n, A, n, ipiv, work, n) == rocblas_status_success,
              EBG_ERR_HIP, "coarse solve: rocsolver_dgetrs failed");
    hipLaunchKernelGGL(to_type_kernel<T>, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s,
                       (const double*)work, inv, nn);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

template <typename T>
void mg_coarse_solve(MGHierarchy<T>& H, const T* bc_d, hipStream_t s, bool zeroed)
{
    const int k = H.nlev() - 1;
    MGLevelDev<T>& l = H.lev[k];
    const UnknownMap m = unknown_map(l.xnum, l.ynum);
    const int n = H.coarse_n;
    const unsigned nbn = (unsigned)((n + 255) / 256);
    if (!zeroed) mg_zero_solution(H, k, s);
    hipLaunchKernelGGL(coarse_pack_kernel<T>, dim3(nbn), dim3(256), 0, s, l, m, n, H.coarse_b);
    hipLaunchKernelGGL(coarse_gemv_kernel<T>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, H.coarse_inv,
                       (const T*)H.coarse_b, H.coarse_x, n);
    hipLaunchKernelGGL(coarse_unpack_kernel<T>, dim3(nbn), dim3(256), 0, s, l, m, n, (const T*)H.coarse_x);
    hipLaunchKernelGGL(mg_bc_kernel<T>, dim3((unsigned)((mg_bc_threads(l.xnum, l.ynum) + 255) / 256)), dim3(256),
                       0, s, l, bc_d[0], bc_d[1], bc_d[2], bc_d[3]);
}

template int mg_coarse_setup<float>(const MGHierarchy<float>&, int, const double*, double*, double*, int*, int*,
                                    float*, void*, hipStream_t);
template int mg_coarse_setup<double>(const MGHierarchy<double>&, int, const double*, double*, double*, int*,
                                     int*, double*, void*, hipStream_t);
template void mg_coarse_solve<float>(MGHierarchy<float>&, const float*, hipStream_t, bool);
template void mg_coarse_solve<double>(MGHierarchy<double>&, const double*, hipStream_t, bool);

}  // namespace ebg
