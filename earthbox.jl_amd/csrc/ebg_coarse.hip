// Direct solve of the coarsest multigrid level of the Stokes preconditioner.
//
// The tiled Gauss-Seidel smoother needs dozens of sweeps on the coarse levels
// and each sweep there is a chain of barrier-bound phases on one or a few
// workgroups.  Replacing the bottom of the hierarchy by an exact solve turns
// that chain into one matrix-vector product: the level's operator is assembled
// column by column from the smoother's own residual arithmetic (so ghost rules,
// mirror factors and steps are exactly the smoother's), inverted once per
// system with rocSOLVER (LU with partial pivoting, then A^T X = I), and
// applied as a dense row-major GEMV, one wavefront per row.
#include "ebg_common.hpp"
#include "ebg_mg.hpp"

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
    T acc = T(0);
    for (int j = lane; j < n; j += 64) acc += a[j] * b[j];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
    if (lane == 0) x[row] = acc;
}

}  // namespace

template <typename T>
int mg_coarse_setup(const MGHierarchy<T>& H, int k, const double* bc_d, double* A, double* work, int* ipiv,
                    int* info, T* inv, void* solver_handle, hipStream_t s)
{
    const MGLevelDev<T>& l = H.lev[k];
    const UnknownMap m = unknown_map(l.xnum, l.ynum);
    const int n = (int)mg_level_unknowns(l.xnum, l.ynum);
    const int64_t nn = (int64_t)n * n;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, A, nn);
    const int64_t pairs = (int64_t)n * CAND;
    hipLaunchKernelGGL(coarse_assemble_kernel<T>, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, s, l, m, n,
                       bc_d[0], bc_d[1], bc_d[2], bc_d[3], A);
    hipLaunchKernelGGL(coarse_gauge_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A, n,
                       m.index(2, 1, 1));
    EBG_HIP(hipGetLastError());
    rocblas_handle h = (rocblas_handle)solver_handle;
    EBG_CHECK(rocblas_set_stream(h, s) == rocblas_status_success, EBG_ERR_HIP, "coarse solve: rocblas_set_stream");
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
void mg_coarse_solve(MGHierarchy<T>& H, const T* bc_d, hipStream_t s)
{
    const int k = H.nlev() - 1;
    MGLevelDev<T>& l = H.lev[k];
    const UnknownMap m = unknown_map(l.xnum, l.ynum);
    const int n = H.coarse_n;
    const unsigned nbn = (unsigned)((n + 255) / 256);
    mg_zero_solution(H, k, s);
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
template void mg_coarse_solve<float>(MGHierarchy<float>&, const float*, hipStream_t);
template void mg_coarse_solve<double>(MGHierarchy<double>&, const double*, hipStream_t);

}  // namespace ebg
