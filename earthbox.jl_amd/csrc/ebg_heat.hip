// Heat equation on the basic grid: right-hand-side terms, the system rows, a
// multigrid-preconditioned flexible GMRES solve in place of the direct solve,
// residuals, the adaptive heat loop and the thermal boundary conditions.
//
// Replaces HeatSolver.calculate_heat_rhs_terms_grid! (heat_solver/HeatSolver.jl:107-118,
// rhs/HeatRhs.jl:5-79), SystemSolver.discretize_and_solve_conductive_heat_equation
// (solver/SystemSolver.jl:32-66: build_system_of_equations, parallel_direct_solver,
// reload_solu_temp), HeatResiduals.compute_residuals_temp!
// (solver/residuals/HeatResiduals.jl:12-162), HeatLoop.run_conductive_heat_solver_loop!
// (heat_loop/HeatLoop.jl:12-78) and ApplyThermalBC.apply_thermal_bc_along_boundaries!
// (interpolation/utils/heat_funcs/ApplyThermalBC.jl:20-44).
//
// The system is never assembled.  Each solve forms the five row coefficients
// of every node once (central, left, right, upper, lower, in the order
// build_system_of_equations emits them; boundary rows hold the unit diagonal
// and one neighbour), so L*T and the right-hand side equal the reference's
// COO triplets applied in order, bit for bit.  The preconditioner is a
// geometric V-cycle on tensor sub-grids (every other node of each axis, the
// last node always kept), the operator rediscretized on each level with the
// same formula from injected conductivity and rho*Cp, red-black Gauss-Seidel
// smoothing, area-weighted restriction of interior residuals, injection on
// the boundary, bilinear prolongation, and a dense inverse on the coarsest
// level (Gauss-Jordan in one workgroup, rebuilt per solve).
#include "ebg_common.hpp"
#include "ebg_krylov.hpp"

#include <cmath>
#include <memory>
#include <vector>

namespace ebg {

namespace {

constexpr int DENSE_MAX = 90;    // coarsest-level unknowns solved by the dense inverse (fits 64 KB of LDS)

__host__ __device__ inline unsigned nbk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

struct HeatLevelDev {
    int nx, ny;
    int64_t n;
    const double *xstp, *ystp, *xstpc, *ystpc;  // nx-1, ny-1, nx-2, ny-2
    const double *kt, *rc;                      // conductivity and rho*Cp at the nodes
    const double *btop2, *bbot2, *bleft2, *bright2;  // boundary neighbour factors
    double* coef;                               // 5 n: C, L, R, U, D
};

// Row coefficients (InternalNodes.jl:56-197, BoundaryNodes.jl:16-251).
__global__ void __launch_bounds__(256) heat_coef_kernel(HeatLevelDev L, double dt)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const int ny = L.ny, nx = L.nx;
    const int i = (int)(t % ny) + 1, j = (int)(t / ny) + 1;
    double c = 0.0, l = 0.0, r = 0.0, u = 0.0, d = 0.0;
    if (j == 1 || j == nx || i == 1 || i == ny) {
        c = 1.0;
        if (i == 1 && 1 < j && j < nx) d = -L.btop2[j - 1];
        else if (i == ny && 1 < j && j < nx) u = -L.bbot2[j - 1];
        else if (j == 1) r = -L.bleft2[i - 1];
        else l = -L.bright2[i - 1];
    } else {
        const double* k = L.kt;
        const double ktC = k[t], ktU = k[t - 1], ktD = k[t + 1], ktL = k[t - ny], ktR = k[t + ny];
        const double dxL = L.xstp[j - 2], dxR = L.xstp[j - 1], dxC = L.xstpc[j - 2];
        const double dyU = L.ystp[i - 2], dyD = L.ystp[i - 1], dyC = L.ystpc[i - 2];
        c = L.rc[t] / dt + ((ktL + ktC) / dxL + (ktC + ktR) / dxR) / 2.0 / dxC +
            ((ktU + ktC) / dyU + (ktC + ktD) / dyD) / 2.0 / dyC;
        l = -(ktL + ktC) / 2.0 / dxL / dxC;
        r = -(ktC + ktR) / 2.0 / dxR / dxC;
        u = -(ktU + ktC) / 2.0 / dyU / dyC;
        d = -(ktC + ktD) / 2.0 / dyD / dyC;
    }
    const int64_t n = L.n;
    L.coef[t] = c;
    L.coef[n + t] = l;
    L.coef[2 * n + t] = r;
    L.coef[3 * n + t] = u;
    L.coef[4 * n + t] = d;
}

// (L x)[t], the row's triplets accumulated from 0.0 in emission order.
__device__ __forceinline__ double heat_row(const HeatLevelDev& L, const double* __restrict__ x, int64_t t)
{
    const int ny = L.ny, nx = L.nx;
    const int i = (int)(t % ny) + 1, j = (int)(t / ny) + 1;
    const int64_t n = L.n;
    const double* cf = L.coef;
    double y = 0.0;
    y += cf[t] * x[t];
    if (j == 1 || j == nx || i == 1 || i == ny) {
        if (i == 1 && 1 < j && j < nx) y += cf[4 * n + t] * x[t + 1];
        else if (i == ny && 1 < j && j < nx) y += cf[3 * n + t] * x[t - 1];
        else if (j == 1) y += cf[2 * n + t] * x[t + ny];
        else y += cf[n + t] * x[t - ny];
    } else {
        y += cf[n + t] * x[t - ny];
        y += cf[2 * n + t] * x[t + ny];
        y += cf[3 * n + t] * x[t - 1];
        y += cf[4 * n + t] * x[t + 1];
    }
    return y;
}

// y = L x (b == nullptr) or y = b - L x
__global__ void __launch_bounds__(256) heat_apply_kernel(HeatLevelDev L, const double* __restrict__ x,
                                                         const double* __restrict__ b, double* __restrict__ y)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const double v = heat_row(L, x, t);
    y[t] = b ? b[t] - v : v;
}

// Right-hand side R (InternalNodes.jl:37-53, BoundaryNodes.jl:16-251).
__global__ void __launch_bounds__(256) heat_rhs_vector_kernel(int nx, int ny, const double* __restrict__ RT,
                                                              const double* __restrict__ tk0,
                                                              const double* __restrict__ rc, double dt,
                                                              ebg_thermal_bc bc, double* __restrict__ R)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny) return;
    const int i = (int)(t % ny) + 1, j = (int)(t / ny) + 1;
    double v;
    if (j == 1 || j == nx || i == 1 || i == ny) {
        if (i == 1 && 1 < j && j < nx) v = bc.btopt[j - 1];
        else if (i == ny && 1 < j && j < nx) v = bc.bbottomt[j - 1];
        else if (j == 1) v = bc.bleftt[i - 1];
        else v = bc.brightt[i - 1];
    } else {
        v = RT[t] + tk0[t] * rc[t] / dt;
    }
    R[t] = v;
}

// One colour of red-black Gauss-Seidel: nodes with (i + j) % 2 == colour.
__global__ void __launch_bounds__(256) heat_gs_kernel(HeatLevelDev L, double* __restrict__ x,
                                                      const double* __restrict__ b, int colour)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    const int i = (int)(t % L.ny), j = (int)(t / L.ny);
    if (((i + j) & 1) != colour) return;
    x[t] += (b[t] - heat_row(L, x, t)) / L.coef[t];
}

struct Transfer {
    int nxf, nyf, nxc, nyc;
    const int *cfx, *cfy;        // fine index of each coarse node (0-based)
    const double *rlx, *rrx, *rly, *rry;  // restriction weights of fine neighbours f-1, f+1
    const int *pcx, *pcy;        // prolongation: coarse left index of each fine node
    const double *ptx, *pty;     // and its weight toward the right coarse node
    const double *xstpc_f, *ystpc_f, *xstpc_c, *ystpc_c;
};

// Coarse residual: area-weighted transpose of the bilinear prolongation over
// the fine interior nodes, divided by the coarse area; injection on the
// coarse boundary.
__global__ void __launch_bounds__(256) heat_restrict_kernel(Transfer T, const double* __restrict__ rf,
                                                            double* __restrict__ rc)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)T.nxc * T.nyc) return;
    const int I = (int)(t % T.nyc), J = (int)(t / T.nyc);
    const int fi = T.cfy[I], fj = T.cfx[J];
    if (I == 0 || J == 0 || I == T.nyc - 1 || J == T.nxc - 1) {
        rc[t] = rf[(int64_t)fj * T.nyf + fi];
        return;
    }
    const double wy[3] = {T.rly[I], 1.0, T.rry[I]}, wx[3] = {T.rlx[J], 1.0, T.rrx[J]};
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const int j0 = fj + b - 1;
        if (wx[b] == 0.0 || j0 < 1 || j0 > T.nxf - 2) continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int i0 = fi + a - 1;
            if (wy[a] == 0.0 || i0 < 1 || i0 > T.nyf - 2) continue;
            s += wy[a] * wx[b] * (T.xstpc_f[j0 - 1] * T.ystpc_f[i0 - 1]) * rf[(int64_t)j0 * T.nyf + i0];
        }
    }
    rc[t] = s / (T.xstpc_c[J - 1] * T.ystpc_c[I - 1]);
}

// x_f += P e_c (bilinear on the tensor grids)
__global__ void __launch_bounds__(256) heat_prolong_kernel(Transfer T, const double* __restrict__ ec,
                                                           double* __restrict__ xf)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)T.nxf * T.nyf) return;
    const int i = (int)(t % T.nyf), j = (int)(t / T.nyf);
    const int ci = T.pcy[i], cj = T.pcx[j];
    const double ty = T.pty[i], tx = T.ptx[j];
    const double* e0 = ec + (int64_t)cj * T.nyc + ci;
    const double* e1 = e0 + T.nyc;
    xf[t] += (1.0 - ty) * (1.0 - tx) * e0[0] + ty * (1.0 - tx) * e0[1] + (1.0 - ty) * tx * e1[0] + ty * tx * e1[1];
}

// Coarse-level node values by injection: conductivity, rho*Cp and the
// boundary neighbour factors.
__global__ void __launch_bounds__(256) heat_inject_kernel(Transfer T, const double* __restrict__ ktf,
                                                          const double* __restrict__ rcf, double* __restrict__ ktc,
                                                          double* __restrict__ rcc)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)T.nxc * T.nyc) return;
    const int I = (int)(t % T.nyc), J = (int)(t / T.nyc);
    const int64_t f = (int64_t)T.cfx[J] * T.nyf + T.cfy[I];
    ktc[t] = ktf[f];
    rcc[t] = rcf[f];
}

__global__ void heat_inject_bc_kernel(Transfer T, const double* __restrict__ top, const double* __restrict__ bot,
                                      const double* __restrict__ left, const double* __restrict__ right,
                                      double* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxc = T.nxc, nyc = T.nyc;
    if (k < nxc) {
        out[k] = top[T.cfx[k]];
        out[nxc + k] = bot[T.cfx[k]];
    }
    if (k < nyc) {
        out[2 * nxc + k] = left[T.cfy[k]];
        out[2 * nxc + nyc + k] = right[T.cfy[k]];
    }
}

// Dense coarsest operator, column-major (A[col * n + row]).
__global__ void __launch_bounds__(256) heat_dense_kernel(HeatLevelDev L, double* __restrict__ A)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = L.n;
    if (t >= n) return;
    for (int64_t c = 0; c < n; ++c) A[c * n + t] = 0.0;
    const int ny = L.ny, nx = L.nx;
    const int i = (int)(t % ny) + 1, j = (int)(t / ny) + 1;
    const double* cf = L.coef;
    A[t * n + t] = cf[t];
    if (j == 1 || j == nx || i == 1 || i == ny) {
        if (i == 1 && 1 < j && j < nx) A[(t + 1) * n + t] = cf[4 * n + t];
        else if (i == ny && 1 < j && j < nx) A[(t - 1) * n + t] = cf[3 * n + t];
        else if (j == 1) A[(t + ny) * n + t] = cf[2 * n + t];
        else A[(t - ny) * n + t] = cf[n + t];
    } else {
        A[(t - ny) * n + t] = cf[n + t];
        A[(t + ny) * n + t] = cf[2 * n + t];
        A[(t - 1) * n + t] = cf[3 * n + t];
        A[(t + 1) * n + t] = cf[4 * n + t];
    }
}

// In-place Gauss-Jordan inverse of a row diagonally dominant matrix (no
// pivoting needed) in one workgroup's LDS: the n x n column-major matrix is
// loaded, inverted with a barrier per pivot step and written back.  Thread q
// scales column q of the pivot row, then eliminates row q.
__global__ void __launch_bounds__(1024) heat_gauss_jordan_kernel(double* __restrict__ A, int n)
{
    extern __shared__ double S[];
    const int tid = threadIdx.x, bs = blockDim.x;
    const int nn = n * n;
    for (int e = tid; e < nn; e += bs) S[e] = A[e];
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const double pinv = 1.0 / S[k * n + k];
        __syncthreads();
        for (int c = tid; c < n; c += bs)
            if (c != k) S[c * n + k] *= pinv;
        __syncthreads();
        for (int r = tid; r < n; r += bs) {
            if (r == k) continue;
            const double f = S[k * n + r];
            for (int c = 0; c < n; ++c)
                if (c != k) S[c * n + r] -= f * S[c * n + k];
            S[k * n + r] = -f * pinv;
        }
        __syncthreads();
        if (tid == 0) S[k * n + k] = pinv;
        __syncthreads();
    }
    for (int e = tid; e < nn; e += bs) A[e] = S[e];
}

// x = Ainv b (column-major inverse)
__global__ void __launch_bounds__(256) heat_dense_solve_kernel(const double* __restrict__ Ainv, int n,
                                                               const double* __restrict__ b, double* __restrict__ x)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (int c = 0; c < n; ++c) s += Ainv[(int64_t)c * n + r] * b[c];
    x[r] = s;
}

__global__ void diff_kernel(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                            double* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = a[t] - b[t];
}

__global__ void heat_fill_kernel(double* a, int64_t n, double v)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// per-block max |a - b| into partial[blockIdx.x]
__global__ void __launch_bounds__(256) max_abs_diff_kernel(const double* __restrict__ a,
                                                           const double* __restrict__ b, int64_t n,
                                                           double* __restrict__ partial)
{
    __shared__ double sh[256];
    double m = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const double d = fabs(a[t] - b[t]);
        m = (d > m || d != d) ? d : m;   // a NaN propagates
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const double o = sh[threadIdx.x + off];
            if (o > sh[threadIdx.x] || o != o) sh[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// HeatRhs (rhs/HeatRhs.jl:5-79); squares as products
__global__ void __launch_bounds__(256) heat_rhs_kernel(int nx, int ny, ebg_heat_rhs_in in, double* __restrict__ RT)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny) return;
    const int i = (int)(t % ny) + 1, j = (int)(t / ny) + 1;
    double v = in.hr1[t];
    if (i >= 2 && i <= ny - 1 && j >= 2 && j <= nx - 1) {
        if (in.iuse_adiabatic_heating == 1) {
            const int64_t vx = (int64_t)(j - 1) * (ny + 1) + (i - 1), vy = (int64_t)(j - 1) * ny + (i - 1);
            v += in.ha1[t] * in.rho1[t] *
                 (in.gravity_x * (in.vx1[vx] + in.vx1[vx + 1]) + in.gravity_y * (in.vy1[vy] + in.vy1[vy + ny])) / 2.0;
        }
        if (in.iuse_shear_heating == 1) {
            const double s = in.sxy2[t];
            v += s * s / in.etas1[t];
            const int np = ny - 1;
            const int64_t p00 = (int64_t)(j - 2) * np + (i - 2), p01 = p00 + np;
            const double a = in.sxx2[p00], b = in.sxx2[p00 + 1], c = in.sxx2[p01], d = in.sxx2[p01 + 1];
            v += (a * a / in.etan1[p00] + b * b / in.etan1[p00 + 1] + c * c / in.etan1[p01] +
                  d * d / in.etan1[p01 + 1]) / 4.0;
        }
    }
    RT[t] = v;
}

// HeatResiduals.compute_residuals_temp!
__global__ void __launch_bounds__(256) heat_residuals_kernel(int nx, int ny, const double* __restrict__ xstp,
                                                             const double* __restrict__ ystp,
                                                             const double* __restrict__ xstpc,
                                                             const double* __restrict__ ystpc, ebg_heat_system s,
                                                             const double* __restrict__ tk2, double* __restrict__ rest)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny) return;
    const int i = (int)(t % ny) + 1, j = (int)(t / ny) + 1;
    if (j == 1 || j == nx || i == 1 || i == ny) {
        rest[t] = 0.0;
        return;
    }
    const double* kt = s.kt1;
    double r = s.RT1[t] + -s.rhocp1[t] * (tk2[t] - s.tk0[t]) / s.timestep;
    r += ((kt[t] + kt[t + ny]) * (tk2[t + ny] - tk2[t]) / xstp[j - 1] -
          (kt[t - ny] + kt[t]) * (tk2[t] - tk2[t - ny]) / xstp[j - 2]) / 2.0 / xstpc[j - 2];
    r += ((kt[t] + kt[t + 1]) * (tk2[t + 1] - tk2[t]) / ystp[i - 1] -
          (kt[t - 1] + kt[t]) * (tk2[t] - tk2[t - 1]) / ystp[i - 2]) / 2.0 / ystpc[i - 2];
    rest[t] = r;
}

// ApplyThermalBC: top and bottom rows (inner columns), then the side columns
// (every row, after the rows: the corners read the updated neighbours)
__global__ void thermal_bc_rows_kernel(int nx, int ny, ebg_thermal_bc bc, double* __restrict__ tk)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 2;
    if (j > nx - 1) return;
    const int64_t c = (int64_t)(j - 1) * ny;
    tk[c] = bc.btopt[j - 1] + bc.btopt[nx + j - 1] * tk[c + 1];
    tk[c + ny - 1] = bc.bbottomt[j - 1] + bc.bbottomt[nx + j - 1] * tk[c + ny - 2];
}

__global__ void thermal_bc_sides_kernel(int nx, int ny, ebg_thermal_bc bc, double* __restrict__ tk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > ny) return;
    tk[i - 1] = bc.bleftt[i - 1] + bc.bleftt[ny + i - 1] * tk[ny + i - 1];
    const int64_t r = (int64_t)(nx - 1) * ny + (i - 1);
    tk[r] = bc.brightt[i - 1] + bc.brightt[ny + i - 1] * tk[r - ny];
}

struct Level {
    int nx, ny;
    int64_t n;
    std::vector<double> gx, gy;
    DevBuf steps;   // xstp, ystp, xstpc, ystpc packed
    DevBuf kt, rc, b2, coef, x, b, r;
    // transfer to the next coarser level
    DevBuf tint, tdbl;
    Transfer T{};
    const double *xstp() { return steps.as<double>(); }
};

struct HeatState {
    uint64_t generation = ~0ull;
    std::vector<std::unique_ptr<Level>> lev;
    DevBuf dense;                 // coarsest inverse
    DevBuf R, w, V, Z, partial, hdev, ydev, nrm, red;
};

std::vector<int> coarse_index(int n)
{
    std::vector<int> c;
    for (int k = 0; k < n; k += 2) c.push_back(k);
    if (c.back() != n - 1) c.push_back(n - 1);
    return c;
}

int upload(DevBuf& d, const void* src, size_t bytes, hipStream_t s)
{
    int rc = d.ensure(bytes);
    if (rc) return rc;
    EBG_HIP(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, s));
    return EBG_OK;
}

// steps of a tensor axis: basic steps and control-volume widths
// (x_{k+1} - x_{k-1}) / 2, the pressure-grid steps
void axis_steps(const std::vector<double>& g, std::vector<double>& stp, std::vector<double>& stpc)
{
    const size_t n = g.size();
    stp.resize(n - 1);
    stpc.resize(n >= 2 ? n - 2 : 0);
    for (size_t k = 0; k + 1 < n; ++k) stp[k] = g[k + 1] - g[k];
    for (size_t k = 0; k + 2 < n; ++k) stpc[k] = (g[k + 2] - g[k]) / 2.0;
}

int build_hierarchy(ebg_ctx* ctx, HeatState& H)
{
    const Grid2D& g = ctx->grid;
    if (H.generation == g.generation && !H.lev.empty()) return EBG_OK;
    hipStream_t s = ctx->stream;
    H.lev.clear();
    std::vector<double> gx = g.h_gridx_b, gy = g.h_gridy_b;
    int rc;
    while (true) {
        auto L = std::make_unique<Level>();
        L->nx = (int)gx.size();
        L->ny = (int)gy.size();
        L->n = (int64_t)L->nx * L->ny;
        L->gx = gx;
        L->gy = gy;
        std::vector<double> sx, sxc, sy, syc;
        axis_steps(gx, sx, sxc);
        axis_steps(gy, sy, syc);
        std::vector<double> packed;
        packed.insert(packed.end(), sx.begin(), sx.end());
        packed.insert(packed.end(), sy.begin(), sy.end());
        packed.insert(packed.end(), sxc.begin(), sxc.end());
        packed.insert(packed.end(), syc.begin(), syc.end());
        if ((rc = upload(L->steps, packed.data(), packed.size() * 8, s))) return rc;
        const int64_t n = L->n;
        if ((rc = L->coef.ensure(5 * n * 8)) || (rc = L->r.ensure(n * 8))) return rc;
        const bool last = n <= DENSE_MAX;
        if (!H.lev.empty()) {  // coarse levels own their node values and vectors
            if ((rc = L->kt.ensure(n * 8)) || (rc = L->rc.ensure(n * 8)) || (rc = L->x.ensure(n * 8)) ||
                (rc = L->b.ensure(n * 8)) || (rc = L->b2.ensure((2 * L->nx + 2 * L->ny) * 8)))
                return rc;
        }
        H.lev.push_back(std::move(L));
        if (last) break;
        Level& F = *H.lev.back();
        // coarsen each axis that still has more than 3 nodes
        std::vector<int> cx, cy;
        if (F.nx > 3) cx = coarse_index(F.nx);
        else for (int k = 0; k < F.nx; ++k) cx.push_back(k);
        if (F.ny > 3) cy = coarse_index(F.ny);
        else for (int k = 0; k < F.ny; ++k) cy.push_back(k);
        std::vector<double> ngx(cx.size()), ngy(cy.size());
        for (size_t k = 0; k < cx.size(); ++k) ngx[k] = gx[cx[k]];
        for (size_t k = 0; k < cy.size(); ++k) ngy[k] = gy[cy[k]];
        // restriction weights: fine neighbours f-1 and f+1 of coarse node I
        auto rweights = [](const std::vector<double>& gf, const std::vector<int>& ci, std::vector<double>& wl,
                           std::vector<double>& wr) {
            const size_t nc = ci.size();
            wl.assign(nc, 0.0);
            wr.assign(nc, 0.0);
            for (size_t I = 0; I < nc; ++I) {
                const int f = ci[I];
                if (I > 0 && ci[I - 1] < f - 1)
                    wl[I] = (gf[f - 1] - gf[ci[I - 1]]) / (gf[f] - gf[ci[I - 1]]);
                if (I + 1 < nc && ci[I + 1] > f + 1)
                    wr[I] = (gf[ci[I + 1]] - gf[f + 1]) / (gf[ci[I + 1]] - gf[f]);
            }
        };
        // prolongation: coarse left index and weight of each fine node
        auto pweights = [](const std::vector<double>& gf, const std::vector<double>& gc, std::vector<int>& c0,
                           std::vector<double>& tw) {
            const int nf = (int)gf.size(), nc = (int)gc.size();
            c0.resize(nf);
            tw.resize(nf);
            int k = 0;
            for (int f = 0; f < nf; ++f) {
                while (k + 1 < nc - 1 && gc[k + 1] <= gf[f]) ++k;
                const int kk = nc >= 2 ? std::min(k, nc - 2) : 0;
                c0[f] = kk;
                tw[f] = nc >= 2 ? (gf[f] - gc[kk]) / (gc[kk + 1] - gc[kk]) : 0.0;
            }
        };
        std::vector<double> rlx, rrx, rly, rry, ptx, pty;
        std::vector<int> pcx, pcy;
        rweights(gx, cx, rlx, rrx);
        rweights(gy, cy, rly, rry);
        pweights(gx, ngx, pcx, ptx);
        pweights(gy, ngy, pcy, pty);
        std::vector<int> ti;
        ti.insert(ti.end(), cx.begin(), cx.end());
        ti.insert(ti.end(), cy.begin(), cy.end());
        ti.insert(ti.end(), pcx.begin(), pcx.end());
        ti.insert(ti.end(), pcy.begin(), pcy.end());
        std::vector<double> td;
        td.insert(td.end(), rlx.begin(), rlx.end());
        td.insert(td.end(), rrx.begin(), rrx.end());
        td.insert(td.end(), rly.begin(), rly.end());
        td.insert(td.end(), rry.begin(), rry.end());
        td.insert(td.end(), ptx.begin(), ptx.end());
        td.insert(td.end(), pty.begin(), pty.end());
        if ((rc = upload(F.tint, ti.data(), ti.size() * 4, s)) || (rc = upload(F.tdbl, td.data(), td.size() * 8, s)))
            return rc;
        const int* ip = F.tint.as<int>();
        const double* dp = F.tdbl.as<double>();
        const int nxc = (int)cx.size(), nyc = (int)cy.size();
        Transfer& T = F.T;
        T.nxf = F.nx;
        T.nyf = F.ny;
        T.nxc = nxc;
        T.nyc = nyc;
        T.cfx = ip;
        T.cfy = ip + nxc;
        T.pcx = ip + nxc + nyc;
        T.pcy = ip + nxc + nyc + F.nx;
        T.rlx = dp;
        T.rrx = dp + nxc;
        T.rly = dp + 2 * nxc;
        T.rry = dp + 2 * nxc + nyc;
        T.ptx = dp + 2 * nxc + 2 * nyc;
        T.pty = dp + 2 * nxc + 2 * nyc + F.nx;
        gx = ngx;
        gy = ngy;
    }
    // fine and coarse control-volume widths of every transfer
    for (size_t l = 0; l + 1 < H.lev.size(); ++l) {
        Level& F = *H.lev[l];
        Level& C = *H.lev[l + 1];
        const double* fs = F.steps.as<double>();
        const double* cs = C.steps.as<double>();
        F.T.xstpc_f = fs + (F.nx - 1) + (F.ny - 1);
        F.T.ystpc_f = F.T.xstpc_f + (F.nx - 2);
        F.T.xstpc_c = cs + (C.nx - 1) + (C.ny - 1);
        F.T.ystpc_c = F.T.xstpc_c + (C.nx - 2);
    }
    const int64_t nd = H.lev.back()->n;
    if ((rc = H.dense.ensure((size_t)nd * nd * 8))) return rc;
    H.generation = g.generation;
    return EBG_OK;
}

HeatLevelDev level_dev(Level& L, const double* kt, const double* rc, const double* b2top, const double* b2bot,
                       const double* b2left, const double* b2right)
{
    HeatLevelDev d{};
    d.nx = L.nx;
    d.ny = L.ny;
    d.n = L.n;
    const double* st = L.steps.as<double>();
    d.xstp = st;
    d.ystp = st + (L.nx - 1);
    d.xstpc = st + (L.nx - 1) + (L.ny - 1);
    d.ystpc = d.xstpc + (L.nx - 2);
    d.kt = kt;
    d.rc = rc;
    d.btop2 = b2top;
    d.bbot2 = b2bot;
    d.bleft2 = b2left;
    d.bright2 = b2right;
    d.coef = L.coef.as<double>();
    return d;
}

// Coefficients on every level for one system (conductivity, rho*Cp and
// boundary factors injected down the hierarchy), and the dense coarsest inverse.
int prepare_levels(ebg_ctx* ctx, HeatState& H, const ebg_heat_system& sys, std::vector<HeatLevelDev>& D)
{
    hipStream_t s = ctx->stream;
    const int nx = (int)ctx->grid.xnum, ny = (int)ctx->grid.ynum;
    D.clear();
    Level& L0 = *H.lev[0];
    D.push_back(level_dev(L0, sys.kt1, sys.rhocp1, sys.bc.btopt + nx, sys.bc.bbottomt + nx, sys.bc.bleftt + ny,
                          sys.bc.brightt + ny));
    for (size_t l = 1; l < H.lev.size(); ++l) {
        Level& F = *H.lev[l - 1];
        Level& C = *H.lev[l];
        const HeatLevelDev& fd = D.back();
        hipLaunchKernelGGL(heat_inject_kernel, dim3(nbk(C.n)), dim3(256), 0, s, F.T, fd.kt, fd.rc,
                           C.kt.as<double>(), C.rc.as<double>());
        hipLaunchKernelGGL(heat_inject_bc_kernel, dim3(nbk(std::max(C.nx, C.ny))), dim3(256), 0, s, F.T, fd.btop2,
                           fd.bbot2, fd.bleft2, fd.bright2, C.b2.as<double>());
        const double* b2 = C.b2.as<double>();
        D.push_back(level_dev(C, C.kt.as<double>(), C.rc.as<double>(), b2, b2 + C.nx, b2 + 2 * C.nx,
                              b2 + 2 * C.nx + C.ny));
    }
    for (auto& d : D) hipLaunchKernelGGL(heat_coef_kernel, dim3(nbk(d.n)), dim3(256), 0, s, d, sys.timestep);
    const HeatLevelDev& dc = D.back();
    double* A = H.dense.as<double>();
    hipLaunchKernelGGL(heat_dense_kernel, dim3(nbk(dc.n)), dim3(256), 0, s, dc, A);
    hipLaunchKernelGGL(heat_gauss_jordan_kernel, dim3(1), dim3(128), (size_t)dc.n * dc.n * 8, s, A, (int)dc.n);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

// z = V-cycle(v) from level 0, zero initial guess
void vcycle(HeatState& H, const std::vector<HeatLevelDev>& D, const double* v, double* z, int pre, int post,
            hipStream_t s)
{
    const int nl = (int)D.size();
    for (int l = 0; l < nl; ++l) {
        const HeatLevelDev& d = D[l];
        Level& L = *H.lev[l];
        const double* b = l == 0 ? v : L.b.as<double>();
        double* x = l == 0 ? z : L.x.as<double>();
        if (l == nl - 1) {
            hipLaunchKernelGGL(heat_dense_solve_kernel, dim3(nbk(d.n)), dim3(256), 0, s,
                               (const double*)H.dense.as<double>(), (int)d.n, b, x);
            break;
        }
        hipLaunchKernelGGL(heat_fill_kernel, dim3(nbk(d.n)), dim3(256), 0, s, x, d.n, 0.0);
        for (int k = 0; k < pre; ++k) {
            hipLaunchKernelGGL(heat_gs_kernel, dim3(nbk(d.n)), dim3(256), 0, s, d, x, b, 0);
            hipLaunchKernelGGL(heat_gs_kernel, dim3(nbk(d.n)), dim3(256), 0, s, d, x, b, 1);
        }
        hipLaunchKernelGGL(heat_apply_kernel, dim3(nbk(d.n)), dim3(256), 0, s, d, (const double*)x, b,
                           L.r.as<double>());
        hipLaunchKernelGGL(heat_restrict_kernel, dim3(nbk(H.lev[l + 1]->n)), dim3(256), 0, s, L.T,
                           (const double*)L.r.as<double>(), H.lev[l + 1]->b.as<double>());
    }
    for (int l = nl - 2; l >= 0; --l) {
        const HeatLevelDev& d = D[l];
        Level& L = *H.lev[l];
        const double* b = l == 0 ? v : L.b.as<double>();
        double* x = l == 0 ? z : L.x.as<double>();
        hipLaunchKernelGGL(heat_prolong_kernel, dim3(nbk(d.n)), dim3(256), 0, s, L.T,
                           (const double*)H.lev[l + 1]->x.as<double>(), x);
        for (int k = 0; k < post; ++k) {
            hipLaunchKernelGGL(heat_gs_kernel, dim3(nbk(d.n)), dim3(256), 0, s, d, x, b, 1);
            hipLaunchKernelGGL(heat_gs_kernel, dim3(nbk(d.n)), dim3(256), 0, s, d, x, b, 0);
        }
    }
}

void multidot(HeatState& H, const double* w, const double* V, int64_t ldv, int k, int64_t n, double* out,
              hipStream_t s)
{
    hipLaunchKernelGGL(multidot_chunks_kernel, dim3(RED_BLOCKS), dim3(256), 0, s, w, V, ldv, k, n,
                       H.partial.as<double>(), (const int*)nullptr);
    hipLaunchKernelGGL(reduce_columns_kernel, dim3((k + 3) / 4), dim3(256), 0, s, H.partial.as<double>(),
                       RED_BLOCKS, k, out, 0, (const int*)nullptr);
}

ebg_heat_solve_opts default_opts()
{
    ebg_heat_solve_opts o{};
    o.restart = 40;
    o.max_iterations = 400;
    o.rtol = 1e-12;
    o.pre_sweeps = 2;
    o.post_sweeps = 2;
    return o;
}

ebg_heat_solve_opts resolve_opts(const ebg_heat_solve_opts* in)
{
    ebg_heat_solve_opts o = default_opts();
    if (!in) return o;
    if (in->restart > 0) o.restart = in->restart;
    if (in->max_iterations > 0) o.max_iterations = in->max_iterations;
    if (in->rtol > 0) o.rtol = in->rtol;
    if (in->pre_sweeps > 0) o.pre_sweeps = in->pre_sweeps;
    if (in->post_sweeps > 0) o.post_sweeps = in->post_sweeps;
    return o;
}

// Right-preconditioned flexible GMRES on L T = R from T = 0 (the direct
// solve's result does not depend on a guess; the zero start keeps the solve
// a function of the system alone).
int heat_fgmres(ebg_ctx* ctx, HeatState& H, const ebg_heat_system& sys, const ebg_heat_solve_opts& o,
                double* x, ebg_solve_stats* stats)
{
    hipStream_t s = ctx->stream;
    const int nx = (int)ctx->grid.xnum, ny = (int)ctx->grid.ynum;
    const int64_t n = (int64_t)nx * ny;
    const int m = o.restart;
    int rc;
    if ((rc = H.R.ensure(n * 8)) || (rc = H.w.ensure(n * 8)) || (rc = H.V.ensure((size_t)(m + 1) * n * 8)) ||
        (rc = H.Z.ensure((size_t)m * n * 8)) || (rc = H.partial.ensure((size_t)RED_BLOCKS * MAXK * 8)) ||
        (rc = H.hdev.ensure(MAXK * 8 * 2)) || (rc = H.ydev.ensure(MAXK * 8)) || (rc = H.nrm.ensure(64)))
        return rc;
    std::vector<HeatLevelDev> D;
    if ((rc = prepare_levels(ctx, H, sys, D))) return rc;
    const HeatLevelDev& d0 = D[0];
    double* R = H.R.as<double>();
    double* w = H.w.as<double>();
    double* V = H.V.as<double>();
    double* Z = H.Z.as<double>();
    double* hd = H.hdev.as<double>();
    double* nrm = H.nrm.as<double>();
    hipLaunchKernelGGL(heat_rhs_vector_kernel, dim3(nbk(n)), dim3(256), 0, s, nx, ny, sys.RT1, sys.tk0, sys.rhocp1,
                       sys.timestep, sys.bc, R);
    hipLaunchKernelGGL(heat_fill_kernel, dim3(nbk(n)), dim3(256), 0, s, x, n, 0.0);
    std::vector<double> Hm((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), hbuf(2 * MAXK);
    multidot(H, R, R, 0, 1, n, nrm, s);
    double bn2 = 0;
    EBG_HIP(hipMemcpyAsync(&bn2, nrm, 8, hipMemcpyDeviceToHost, s));
    EBG_HIP(hipStreamSynchronize(s));
    const double bnorm = std::sqrt(bn2);
    if (!std::isfinite(bnorm)) {
        set_error("ebg_heat_solve: right-hand side is not finite");
        return EBG_ERR_NUMERIC;
    }
    int total = 0, restarts = 0;
    double relres = 0.0, relres0 = -1.0;
    if (bnorm == 0) {
        if (stats) { stats->iterations = 0; stats->restarts = 0; stats->relres = 0; stats->relres_initial = 0; }
        return EBG_OK;
    }
    while (true) {
        hipLaunchKernelGGL(heat_apply_kernel, dim3(nbk(n)), dim3(256), 0, s, d0, (const double*)x, (const double*)R, w);
        multidot(H, w, w, 0, 1, n, nrm, s);
        double beta2;
        EBG_HIP(hipMemcpyAsync(&beta2, nrm, 8, hipMemcpyDeviceToHost, s));
        EBG_HIP(hipStreamSynchronize(s));
        const double beta = std::sqrt(beta2);
        relres = beta / bnorm;
        if (relres0 < 0) relres0 = relres;
        if (!std::isfinite(relres)) {
            set_error("ebg_heat_solve: residual became non-finite");
            return EBG_ERR_NUMERIC;
        }
        if (relres <= o.rtol || total >= o.max_iterations) break;
        EBG_HIP(hipMemcpyAsync(nrm, &beta, 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(scale_kernel, dim3(nbk(n)), dim3(256), 0, s, V, w, nrm, n);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        for (; k < m && total < o.max_iterations; ++k) {
            double* vk = V + (int64_t)k * n;
            double* zk = Z + (int64_t)k * n;
            vcycle(H, D, vk, zk, o.pre_sweeps, o.post_sweeps, s);
            hipLaunchKernelGGL(heat_apply_kernel, dim3(nbk(n)), dim3(256), 0, s, d0, (const double*)zk,
                               (const double*)nullptr, w);
            // classical Gram-Schmidt, a second pass when the first cancelled
            // more than half of the squared norm
            multidot(H, w, w, 0, 1, n, nrm + 1, s);
            multidot(H, w, V, n, k + 1, n, hd, s);
            hipLaunchKernelGGL(update_kernel, dim3(nbk(n)), dim3(256), 0, s, w, V, n, hd, k + 1, n);
            multidot(H, w, w, 0, 1, n, nrm, s);
            double wn[2];
            EBG_HIP(hipMemcpyAsync(hbuf.data(), hd, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, s));
            EBG_HIP(hipMemcpyAsync(wn, nrm, sizeof(wn), hipMemcpyDeviceToHost, s));
            EBG_HIP(hipStreamSynchronize(s));
            double wn2 = wn[0];
            for (int q = 0; q <= k; ++q) Hm[(size_t)k * (m + 1) + q] = hbuf[q];
            if (!(wn2 > 0.5 * wn[1])) {
                multidot(H, w, V, n, k + 1, n, hd + MAXK, s);
                hipLaunchKernelGGL(update_kernel, dim3(nbk(n)), dim3(256), 0, s, w, V, n, hd + MAXK, k + 1, n);
                multidot(H, w, w, 0, 1, n, nrm, s);
                EBG_HIP(hipMemcpyAsync(hbuf.data() + MAXK, hd + MAXK, sizeof(double) * (k + 1),
                                       hipMemcpyDeviceToHost, s));
                EBG_HIP(hipMemcpyAsync(&wn2, nrm, 8, hipMemcpyDeviceToHost, s));
                EBG_HIP(hipStreamSynchronize(s));
                for (int q = 0; q <= k; ++q) Hm[(size_t)k * (m + 1) + q] += hbuf[MAXK + q];
            }
            const double hn = std::sqrt(wn2);
            Hm[(size_t)k * (m + 1) + k + 1] = hn;
            if (!std::isfinite(hn)) {
                set_error("ebg_heat_solve: preconditioned Krylov vector is non-finite");
                return EBG_ERR_NUMERIC;
            }
            if (hn > 0) {
                EBG_HIP(hipMemcpyAsync(nrm, &hn, 8, hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(scale_kernel, dim3(nbk(n)), dim3(256), 0, s, V + (int64_t)(k + 1) * n, w, nrm,
                                   n);
            }
            double* col = &Hm[(size_t)k * (m + 1)];
            for (int q = 0; q < k; ++q) {
                const double t = cs[q] * col[q] + sn[q] * col[q + 1];
                col[q + 1] = -sn[q] * col[q] + cs[q] * col[q + 1];
                col[q] = t;
            }
            const double dd = std::hypot(col[k], col[k + 1]);
            cs[k] = col[k] / dd;
            sn[k] = col[k + 1] / dd;
            col[k] = dd;
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
            for (int p = q + 1; p < k; ++p) v -= Hm[(size_t)p * (m + 1) + q] * y[p];
            y[q] = v / Hm[(size_t)q * (m + 1) + q];
        }
        EBG_HIP(hipMemcpyAsync(H.ydev.p, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(combine_kernel, dim3(nbk(n)), dim3(256), 0, s, x, Z, n, H.ydev.as<double>(), k, n);
        ++restarts;
    }
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

HeatState* state(ebg_ctx* ctx)
{
    if (!ctx->heat) ctx->heat = new HeatState();
    return (HeatState*)ctx->heat;
}

int check_system(ebg_ctx* ctx, const ebg_heat_system* sys, const char* who)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, std::string(who) + ": no grid");
    EBG_CHECK(sys && sys->RT1 && sys->tk0 && sys->rhocp1 && sys->kt1, EBG_ERR_ARG,
              std::string(who) + ": NULL system arrays");
    EBG_CHECK(sys->bc.btopt && sys->bc.bbottomt && sys->bc.bleftt && sys->bc.brightt, EBG_ERR_ARG,
              std::string(who) + ": NULL thermal boundary arrays");
    EBG_CHECK(sys->timestep > 0 && std::isfinite(sys->timestep), EBG_ERR_ARG,
              std::string(who) + ": timestep must be > 0");
    EBG_CHECK(ctx->grid.xnum >= 3 && ctx->grid.ynum >= 3, EBG_ERR_ARG,
              std::string(who) + ": the grid needs at least 3 x 3 nodes");
    return EBG_OK;
}

int max_abs_diff(ebg_ctx* ctx, HeatState& H, const double* a, const double* b, int64_t n, double* out)
{
    hipStream_t s = ctx->stream;
    int rc;
    if ((rc = H.red.ensure(RED_BLOCKS * 8))) return rc;
    const unsigned blocks = (unsigned)std::min<int64_t>(RED_BLOCKS, nbk(n));
    hipLaunchKernelGGL(max_abs_diff_kernel, dim3(blocks), dim3(256), 0, s, a, b, n, H.red.as<double>());
    std::vector<double> part(blocks);
    EBG_HIP(hipMemcpyAsync(part.data(), H.red.p, blocks * 8, hipMemcpyDeviceToHost, s));
    EBG_HIP(hipStreamSynchronize(s));
    double m = 0.0;
    for (double v : part) m = (v > m || v != v) ? v : m;
    *out = m;
    return EBG_OK;
}

}  // namespace

void heat_release(ebg_ctx* ctx)
{
    if (ctx && ctx->heat) {
        delete (HeatState*)ctx->heat;
        ctx->heat = nullptr;
    }
}

}  // namespace ebg

using namespace ebg;

extern "C" {

int ebg_apply_thermal_bc(ebg_ctx* ctx, const ebg_thermal_bc* bc, double* tk1)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_apply_thermal_bc: no grid");
    EBG_CHECK(bc && tk1 && bc->btopt && bc->bbottomt && bc->bleftt && bc->brightt, EBG_ERR_ARG,
              "ebg_apply_thermal_bc: NULL arguments");
    const int nx = (int)ctx->grid.xnum, ny = (int)ctx->grid.ynum;
    hipLaunchKernelGGL(thermal_bc_rows_kernel, dim3(nbk(nx)), dim3(256), 0, ctx->stream, nx, ny, *bc, tk1);
    hipLaunchKernelGGL(thermal_bc_sides_kernel, dim3(nbk(ny)), dim3(256), 0, ctx->stream, nx, ny, *bc, tk1);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_heat_rhs(ebg_ctx* ctx, const ebg_heat_rhs_in* in, double* RT1)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_heat_rhs: no grid");
    EBG_CHECK(in && RT1 && in->hr1, EBG_ERR_ARG, "ebg_heat_rhs: NULL arguments");
    EBG_CHECK(in->iuse_adiabatic_heating != 1 || (in->ha1 && in->rho1 && in->vx1 && in->vy1), EBG_ERR_ARG,
              "ebg_heat_rhs: adiabatic heating needs ha1, rho1, vx1 and vy1");
    EBG_CHECK(in->iuse_shear_heating != 1 || (in->sxy2 && in->etas1 && in->sxx2 && in->etan1), EBG_ERR_ARG,
              "ebg_heat_rhs: shear heating needs sxy2, etas1, sxx2 and etan1");
    const int nx = (int)ctx->grid.xnum, ny = (int)ctx->grid.ynum;
    hipLaunchKernelGGL(heat_rhs_kernel, dim3(nbk((int64_t)nx * ny)), dim3(256), 0, ctx->stream, nx, ny, *in, RT1);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_heat_system_apply(ebg_ctx* ctx, const ebg_heat_system* sys, const double* T, double* LT, double* R)
{
    int rc = check_system(ctx, sys, "ebg_heat_system_apply");
    if (rc) return rc;
    EBG_CHECK(!LT || T, EBG_ERR_ARG, "ebg_heat_system_apply: LT needs T");
    HeatState& H = *state(ctx);
    if ((rc = build_hierarchy(ctx, H))) return rc;
    hipStream_t s = ctx->stream;
    const int nx = (int)ctx->grid.xnum, ny = (int)ctx->grid.ynum;
    const int64_t n = (int64_t)nx * ny;
    HeatLevelDev d = level_dev(*H.lev[0], sys->kt1, sys->rhocp1, sys->bc.btopt + nx, sys->bc.bbottomt + nx,
                               sys->bc.bleftt + ny, sys->bc.brightt + ny);
    hipLaunchKernelGGL(heat_coef_kernel, dim3(nbk(n)), dim3(256), 0, s, d, sys->timestep);
    if (LT) hipLaunchKernelGGL(heat_apply_kernel, dim3(nbk(n)), dim3(256), 0, s, d, T, (const double*)nullptr, LT);
    if (R)
        hipLaunchKernelGGL(heat_rhs_vector_kernel, dim3(nbk(n)), dim3(256), 0, s, nx, ny, sys->RT1, sys->tk0,
                           sys->rhocp1, sys->timestep, sys->bc, R);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_heat_solve(ebg_ctx* ctx, const ebg_heat_system* sys, double* tk2, const ebg_heat_solve_opts* opts,
                   ebg_solve_stats* stats)
{
    int rc = check_system(ctx, sys, "ebg_heat_solve");
    if (rc) return rc;
    EBG_CHECK(tk2, EBG_ERR_ARG, "ebg_heat_solve: NULL tk2");
    const ebg_heat_solve_opts o = resolve_opts(opts);
    EBG_CHECK(o.restart >= 1 && o.restart <= MAXK - 2, EBG_ERR_ARG,
              "ebg_heat_solve: restart must be 1.." + std::to_string(MAXK - 2));
    HeatState& H = *state(ctx);
    if ((rc = build_hierarchy(ctx, H))) return rc;
    return heat_fgmres(ctx, H, *sys, o, tk2, stats);
}

int ebg_heat_residuals(ebg_ctx* ctx, const ebg_heat_system* sys, const double* tk2, double* rest)
{
    int rc = check_system(ctx, sys, "ebg_heat_residuals");
    if (rc) return rc;
    EBG_CHECK(tk2 && rest, EBG_ERR_ARG, "ebg_heat_residuals: NULL arguments");
    HeatState& H = *state(ctx);
    if ((rc = build_hierarchy(ctx, H))) return rc;
    const int nx = (int)ctx->grid.xnum, ny = (int)ctx->grid.ynum;
    Level& L = *H.lev[0];
    const double* st = L.steps.as<double>();
    const double *xstp = st, *ystp = st + (nx - 1), *xstpc = st + (nx - 1) + (ny - 1), *ystpc = xstpc + (nx - 2);
    hipLaunchKernelGGL(heat_residuals_kernel, dim3(nbk((int64_t)nx * ny)), dim3(256), 0, ctx->stream, nx, ny, xstp,
                       ystp, xstpc, ystpc, *sys, tk2, rest);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_run_conductive_heat_solver_loop(ebg_ctx* ctx, const ebg_heat_loop_in* in, ebg_heat_loop_out* out,
                                        const ebg_heat_solve_opts* opts, ebg_heat_loop_stats* stats)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_run_conductive_heat_solver_loop: no grid");
    EBG_CHECK(in && out && out->tk0 && out->tk2 && out->rest, EBG_ERR_ARG,
              "ebg_run_conductive_heat_solver_loop: NULL arguments");
    EBG_CHECK(in->timestep > 0 && std::isfinite(in->timestep) && in->max_temp_change > 0, EBG_ERR_ARG,
              "ebg_run_conductive_heat_solver_loop: timestep and max_temp_change must be > 0");
    EBG_CHECK(in->max_substeps >= 0, EBG_ERR_ARG, "ebg_run_conductive_heat_solver_loop: max_substeps < 0");
    const int max_sub = in->max_substeps > 0 ? in->max_substeps : 100000;
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->grid.xnum * ctx->grid.ynum;
    HeatState& H = *state(ctx);
    ebg_heat_system sys{};
    sys.RT1 = in->RT1;
    sys.tk0 = out->tk0;
    sys.rhocp1 = in->rhocp1;
    sys.kt1 = in->kt1;
    sys.bc = in->bc;
    ebg_heat_loop_stats st{};
    double dt_loop = in->timestep, total = 0.0;
    int rc = EBG_OK;
    // HeatLoop.jl:46-70
    while (total < in->timestep) {
        EBG_CHECK(st.substeps < max_sub, EBG_ERR_NUMERIC,
                  "ebg_run_conductive_heat_solver_loop: more than " + std::to_string(max_sub) + " sub-steps");
        ebg_solve_stats ss{};
        sys.timestep = dt_loop;
        rc = ebg_heat_solve(ctx, &sys, out->tk2, opts, &ss);
        if (rc < 0) return rc;
        st.iterations += ss.iterations;
        st.last_relres = ss.relres;
        st.not_converged += rc == EBG_NOT_CONVERGED;
        ++st.solves;
        if ((rc = ebg_heat_residuals(ctx, &sys, out->tk2, out->rest))) return rc;
        double dtkmax;
        if ((rc = max_abs_diff(ctx, H, out->tk2, out->tk0, n, &dtkmax))) return rc;
        EBG_CHECK(std::isfinite(dtkmax), EBG_ERR_NUMERIC,
                  "ebg_run_conductive_heat_solver_loop: temperature change is not finite");
        if (dtkmax > in->max_temp_change) {
            dt_loop = dt_loop * in->max_temp_change / dtkmax;
            sys.timestep = dt_loop;
            rc = ebg_heat_solve(ctx, &sys, out->tk2, opts, &ss);
            if (rc < 0) return rc;
            st.iterations += ss.iterations;
            st.last_relres = ss.relres;
            st.not_converged += rc == EBG_NOT_CONVERGED;
            ++st.solves;
            if ((rc = ebg_heat_residuals(ctx, &sys, out->tk2, out->rest))) return rc;
        }
        st.timestep_last = dt_loop;
        total = total + dt_loop;
        dt_loop = std::min(dt_loop, in->timestep - total);
        EBG_HIP(hipMemcpyAsync(out->tk0, out->tk2, n * 8, hipMemcpyDeviceToDevice, s));
        ++st.substeps;
    }
    st.timestep_sum = total;
    if (out->dtk1) {
        EBG_CHECK(in->tk1, EBG_ERR_ARG, "ebg_run_conductive_heat_solver_loop: dtk1 needs tk1");
        hipLaunchKernelGGL(diff_kernel, dim3(nbk(n)), dim3(256), 0, s, (const double*)out->tk2, in->tk1, n,
                           out->dtk1);
        EBG_HIP(hipGetLastError());
    }
    if (stats) *stats = st;
    return st.not_converged ? EBG_NOT_CONVERGED : EBG_OK;
}

}  // extern "C"
