// Geometric multigrid for the 3-D staggered Stokes-continuity equations (the
// reference's experimental 3-D solver, src/stokes_continuity_solver/multigrid/),
// as HIP kernels templated on the working precision T (double: plain FP64 and
// the reference-order API; float: the mixed-precision preconditioner).
//
// Layout: every level array is column-major with y fastest, then x, then z
// (Julia's Array{Float64,3}); shapes (y, x, z) as the reference's LevelData:
//   vx (ny+1, nx, nz+1), vy (ny, nx+1, nz+1), vz (ny+1, nx+1, nz),
//   pr / etan (ny-1, nx-1, nz-1), etaxy (ny, nx, nz-1), etaxz (ny-1, nx, nz),
//   etayz (ny, nx-1, nz).
//
//  - smoother: Gauss-Seidel v += dR/diag*theta_s and p += etan*dRc*theta_c
//    (SolveStokes3dOpt.jl:12-157), coloured by (i+j+k) parity and phased per
//    colour as vx, vy, vz, p so every launch is race-free and deterministic;
//  - free-slip ghost nodes (BoundaryConditions.jl:92-174) from a per-level
//    table that the host derives by running the reference's visiting order
//    symbolically, applied as gather-then-scatter so reads see the pre-pass
//    values exactly as the serial loops do;
//  - residuals (Residuals.jl:10-560), pressure normalisation
//    (PressureUpdate.jl:5-21);
//  - restriction (Restriction.jl:64-376), viscosity restriction
//    (ViscosityRestriction.jl:55-230) and prolongation (Prolongation.jl:60-140)
//    in gather form: every coarse node sums its fine contributors in the order
//    the reference's serial scatter adds them, so FP64 results equal the loops.
#pragma once
#include <hip/hip_runtime.h>
#include <array>
#include <cstdint>
#include <vector>

namespace ebg {

__host__ __device__ __forceinline__ size_t i3(int n1, int n2, int i, int j, int k)
{
    return ((size_t)(k - 1) * (size_t)n2 + (size_t)(j - 1)) * (size_t)n1 + (size_t)(i - 1);
}

template <typename T>
struct MG3LevelDev {
    int xnum, ynum, znum;
    const T *xstp_b, *ystp_b, *zstp_b, *ystp_vx, *zstp_vx, *xstp_vy;
    T *etan, *etaxy, *etaxz, *etayz;
    T *vx, *vy, *vz, *pr;
    T *RX, *RY, *RZ, *RC;
    T *rx, *ry, *rz, *rc;
    // ghost-node table (host-derived): per entry the node index and the
    // source index (-1: zero) within one velocity array; entries [0, nbx)
    // belong to vx, [nbx, nbx+nby) to vy, the rest to vz
    const int *bc_dst, *bc_src;
    int nbx, nby, nbz;
    T* bc_tmp;

    __device__ T& VX(int i, int j, int k) const { return vx[i3(ynum + 1, xnum, i, j, k)]; }
    __device__ T& VY(int i, int j, int k) const { return vy[i3(ynum, xnum + 1, i, j, k)]; }
    __device__ T& VZ(int i, int j, int k) const { return vz[i3(ynum + 1, xnum + 1, i, j, k)]; }
    __device__ T& PR(int i, int j, int k) const { return pr[i3(ynum - 1, xnum - 1, i, j, k)]; }
    __device__ T EN(int i, int j, int k) const { return etan[i3(ynum - 1, xnum - 1, i, j, k)]; }
    __device__ T EXY(int i, int j, int k) const { return etaxy[i3(ynum, xnum, i, j, k)]; }
    __device__ T EXZ(int i, int j, int k) const { return etaxz[i3(ynum - 1, xnum, i, j, k)]; }
    __device__ T EYZ(int i, int j, int k) const { return etayz[i3(ynum, xnum - 1, i, j, k)]; }
    __host__ __device__ size_t nvx() const { return (size_t)(ynum + 1) * xnum * (znum + 1); }
    __host__ __device__ size_t nvy() const { return (size_t)ynum * (xnum + 1) * (znum + 1); }
    __host__ __device__ size_t nvz() const { return (size_t)(ynum + 1) * (xnum + 1) * znum; }
    __host__ __device__ size_t npr() const { return (size_t)(ynum - 1) * (xnum - 1) * (znum - 1); }
};

// x-Stokes residual and diagonal at vx(i, j, k) (Residuals.jl:145-225).
// ZR: the right-hand side is taken as zero (operator application).
template <typename T, bool ZR>
__device__ __forceinline__ T vx_res3(const MG3LevelDev<T>& l, int i, int j, int k, T* coef)
{
    const T dxL = l.xstp_b[j - 2], dxR = l.xstp_b[j - 1], dxC = l.xstp_vy[j - 1];
    const T dyU = l.ystp_vx[i - 2], dyD = l.ystp_vx[i - 1], dyC = l.ystp_b[i - 2];
    const T dzF = l.zstp_vx[k - 1], dzB = l.zstp_vx[k - 2], dzC = l.zstp_b[k - 2];
    const T nR = l.EN(i - 1, j, k - 1), nL = l.EN(i - 1, j - 1, k - 1);
    const T xyU = l.EXY(i - 1, j, k - 1), xyD = l.EXY(i, j, k - 1);
    const T xzF = l.EXZ(i - 1, j, k), xzB = l.EXZ(i - 1, j, k - 1);
    const T pL = l.PR(i - 1, j - 1, k - 1), pR = l.PR(i - 1, j, k - 1);
    const T vxL = l.VX(i, j - 1, k), vxC = l.VX(i, j, k), vxR = l.VX(i, j + 1, k);
    const T vxU = l.VX(i - 1, j, k), vxD = l.VX(i + 1, j, k);
    const T vxF = l.VX(i, j, k + 1), vxB = l.VX(i, j, k - 1);
    const T vyUL = l.VY(i - 1, j, k), vyUR = l.VY(i - 1, j + 1, k);
    const T vyLL = l.VY(i, j, k), vyLR = l.VY(i, j + 1, k);
    const T vzRF = l.VZ(i, j + 1, k), vzLF = l.VZ(i, j, k);
    const T vzRB = l.VZ(i, j + 1, k - 1), vzLB = l.VZ(i, j, k - 1);
    const T dp_dx = (pR - pL) / dxC;
    const T dsxx = T(2) * (nR * (vxR - vxC) / dxR - nL * (vxC - vxL) / dxL) / dxC;
    const T dsxy = (xyD * ((vxD - vxC) / dyD + (vyLR - vyLL) / dxC) -
                    xyU * ((vxC - vxU) / dyU + (vyUR - vyUL) / dxC)) / dyC;
    const T dsxz = (xzF * ((vzRF - vzLF) / dxC + (vxF - vxC) / dzF) -
                    xzB * ((vzRB - vzLB) / dxC + (vxC - vxB) / dzB)) / dzC;
    const T rhs = ZR ? T(0) : l.RX[i3(l.ynum + 1, l.xnum, i, j, k)];
    if (coef)
        *coef = T(-2) * (nR / dxR / dxC + nL / dxL / dxC) - (xyD / dyD / dyC + xyU / dyU / dyC) -
                (xzF / dzF / dzC + xzB / dzB / dzC);
    return rhs + dp_dx - dsxx - dsxy - dsxz;
}

// y-Stokes residual and diagonal at vy(i, j, k) (Residuals.jl:289-371)
template <typename T, bool ZR>
__device__ __forceinline__ T vy_res3(const MG3LevelDev<T>& l, int i, int j, int k, T* coef)
{
    const T dxL = l.xstp_b[j - 2], dxR = l.xstp_vy[j - 1], dxC = l.xstp_vy[j - 2];
    const T dyU = l.ystp_b[i - 2], dyD = l.ystp_b[i - 1], dyC = l.ystp_vx[i - 1];
    const T dzF = l.zstp_vx[k - 1], dzB = l.zstp_vx[k - 2], dzC = l.zstp_b[k - 2];
    const T pD = l.PR(i, j - 1, k - 1), pU = l.PR(i - 1, j - 1, k - 1);
    const T nD = l.EN(i, j - 1, k - 1), nU = l.EN(i - 1, j - 1, k - 1);
    const T xyR = l.EXY(i, j, k - 1), xyL = l.EXY(i, j - 1, k - 1);
    const T yzF = l.EYZ(i, j - 1, k), yzB = l.EYZ(i, j - 1, k - 1);
    const T vyU = l.VY(i - 1, j, k), vyC = l.VY(i, j, k), vyD = l.VY(i + 1, j, k);
    const T vyR = l.VY(i, j + 1, k), vyL = l.VY(i, j - 1, k);
    const T vyF = l.VY(i, j, k + 1), vyB = l.VY(i, j, k - 1);
    const T vxLR = l.VX(i + 1, j, k), vxUR = l.VX(i, j, k);
    const T vxLL = l.VX(i + 1, j - 1, k), vxUL = l.VX(i, j - 1, k);
    const T vzDF = l.VZ(i + 1, j, k), vzUF = l.VZ(i, j, k);
    const T vzDB = l.VZ(i + 1, j, k - 1), vzUB = l.VZ(i, j, k - 1);
    const T dp_dy = (pD - pU) / dyC;
    const T dsyy = T(2) * (nD * (vyD - vyC) / dyD - nU * (vyC - vyU) / dyU) / dyC;
    const T dsyx = (xyR * ((vyR - vyC) / dxR + (vxLR - vxUR) / dyC) -
                    xyL * ((vyC - vyL) / dxL + (vxLL - vxUL) / dyC)) / dxC;
    const T dsyz = (yzF * ((vyF - vyC) / dzF + (vzDF - vzUF) / dyC) -
                    yzB * ((vyC - vyB) / dzB + (vzDB - vzUB) / dyC)) / dzC;
    const T rhs = ZR ? T(0) : l.RY[i3(l.ynum, l.xnum + 1, i, j, k)];
    if (coef)
        *coef = T(-2) * (nD / dyD / dyC + nU / dyU / dyC) - (xyR / dxR / dxC + xyL / dxL / dxC) -
                (yzF / dzF / dzC + yzB / dzB / dzC);
    return rhs + dp_dy - dsyy - dsyx - dsyz;
}

// z-Stokes residual and diagonal at vz(i, j, k) (Residuals.jl:435-521)
template <typename T, bool ZR>
__device__ __forceinline__ T vz_res3(const MG3LevelDev<T>& l, int i, int j, int k, T* coef)
{
    const T dzR = l.zstp_b[k - 1], dzL = l.zstp_b[k - 2], dzC = l.zstp_vx[k - 1];
    const T dxF = l.xstp_vy[j - 1], dxB = l.xstp_vy[j - 2], dxC = l.xstp_b[j - 2];
    const T dyU = l.ystp_vx[i - 2], dyD = l.ystp_vx[i - 1], dyC = l.ystp_b[i - 2];
    const T pF = l.PR(i - 1, j - 1, k), pB = l.PR(i - 1, j - 1, k - 1);
    const T nF = l.EN(i - 1, j - 1, k), nB = l.EN(i - 1, j - 1, k - 1);
    const T xzR = l.EXZ(i - 1, j, k), xzL = l.EXZ(i - 1, j - 1, k);
    const T yzU = l.EYZ(i - 1, j - 1, k), yzD = l.EYZ(i, j - 1, k);
    const T vxRF = l.VX(i, j, k + 1), vxLF = l.VX(i, j, k);
    const T vxRB = l.VX(i, j - 1, k + 1), vxLB = l.VX(i, j - 1, k);
    const T vyLR = l.VY(i, j, k + 1), vyLL = l.VY(i, j, k);
    const T vyUR = l.VY(i - 1, j, k + 1), vyUL = l.VY(i - 1, j, k);
    const T vzL = l.VZ(i, j, k - 1), vzC = l.VZ(i, j, k), vzR = l.VZ(i, j, k + 1);
    const T vzF = l.VZ(i, j + 1, k), vzB = l.VZ(i, j - 1, k);
    const T vzD = l.VZ(i + 1, j, k), vzU = l.VZ(i - 1, j, k);
    const T dp_dz = (pF - pB) / dzC;
    const T dszz = T(2) * (nF * (vzR - vzC) / dzR - nB * (vzC - vzL) / dzL) / dzC;
    const T dszx = (xzR * ((vxRF - vxLF) / dzC + (vzF - vzC) / dxF) -
                    xzL * ((vxRB - vxLB) / dzC + (vzC - vzB) / dxB)) / dxC;
    const T dszy = (yzD * ((vzD - vzC) / dyD + (vyLR - vyLL) / dzC) -
                    yzU * ((vzC - vzU) / dyU + (vyUR - vyUL) / dzC)) / dyC;
    const T rhs = ZR ? T(0) : l.RZ[i3(l.ynum + 1, l.xnum + 1, i, j, k)];
    if (coef)
        *coef = T(-2) * (nF / dzR / dzC + nB / dzL / dzC) - (xzR / dxF / dxC + xzL / dxB / dxC) -
                (yzD / dyD / dyC + yzU / dyU / dyC);
    return rhs + dp_dz - dszz - dszx - dszy;
}

// continuity residual at pr(i, j, k) (Residuals.jl:525-560)
template <typename T, bool ZR>
__device__ __forceinline__ T pr_res3(const MG3LevelDev<T>& l, int i, int j, int k)
{
    const T dxC = l.xstp_b[j - 1], dyC = l.ystp_b[i - 1], dzC = l.zstp_b[k - 1];
    const T div = (l.VX(i + 1, j + 1, k + 1) - l.VX(i + 1, j, k + 1)) / dxC +
                  (l.VY(i + 1, j + 1, k + 1) - l.VY(i, j + 1, k + 1)) / dyC +
                  (l.VZ(i + 1, j + 1, k + 1) - l.VZ(i + 1, j + 1, k)) / dzC;
    const T rhs = ZR ? T(0) : l.RC[i3(l.ynum - 1, l.xnum - 1, i, j, k)];
    return rhs - div;
}

// Fine->coarse transfer of one axis and one node kind: for each fine node its
// left coarse node (1-based) and normalised distance; for each coarse node
// the contiguous fine range [lo, hi] (1-based, empty when lo > hi) that adds
// to it within the loop bounds of the operation.
template <typename T>
struct Axis3Dev {
    const int* idx;
    const T* d;
};
struct Range3Dev {
    const int *lo, *hi;
};
enum { K3_B = 0, K3_V = 1, K3_P = 2, K3_S = 3 };
// restriction ranges (fine loop bounds of Restriction.jl:130-158) per node kind:
// R3_V [2, n], R3_B [2, n-1], R3_P [2, n-1]; viscosity ranges: all fine nodes
enum { R3_V = 0, R3_B = 1, R3_P = 2, F3_B = 3, F3_S = 4, F3_P = 5, NR3 = 6 };

template <typename T>
struct Transfer3Dev {
    Axis3Dev<T> ax[3][4];   // [axis y/x/z][kind]
    Range3Dev rg[3][NR3];   // [axis][range]
};

// ---------------------------------------------------------------- kernels
template <typename T>
__global__ void __launch_bounds__(256) mg3_bc_gather_kernel(MG3LevelDev<T> l)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = l.nbx + l.nby + l.nbz;
    if (t >= n) return;
    const T* base = t < l.nbx ? l.vx : (t < l.nbx + l.nby ? l.vy : l.vz);
    const int s = l.bc_src[t];
    l.bc_tmp[t] = s >= 0 ? base[s] : T(0);
}
template <typename T>
__global__ void __launch_bounds__(256) mg3_bc_scatter_kernel(MG3LevelDev<T> l)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = l.nbx + l.nby + l.nbz;
    if (t >= n) return;
    T* base = t < l.nbx ? l.vx : (t < l.nbx + l.nby ? l.vy : l.vz);
    base[l.bc_dst[t]] = l.bc_tmp[t];
}

// interior index box of one phase (0 vx, 1 vy, 2 vz, 3 p)
struct Box3 {
    int ilo, ihi, jlo, jhi, klo, khi;
};
template <typename T>
__host__ __device__ inline Box3 phase_box(const MG3LevelDev<T>& l, int phase)
{
    const int X = l.xnum, Y = l.ynum, Z = l.znum;
    switch (phase) {
        case 0: return {2, Y, 2, X - 1, 2, Z};
        case 1: return {2, Y - 1, 2, X, 2, Z};
        case 2: return {2, Y, 2, X, 2, Z - 1};
        default: return {1, Y - 1, 1, X - 1, 1, Z - 1};
    }
}
// threads of one colour: ceil(ni/2) per row
inline int64_t phase_threads(const Box3& b)
{
    const int64_t nh = (b.ihi - b.ilo + 2) / 2;
    return nh * (b.jhi - b.jlo + 1) * (int64_t)(b.khi - b.klo + 1);
}

template <typename T, int PHASE>
__global__ void __launch_bounds__(256) mg3_sweep_kernel(MG3LevelDev<T> l, Box3 b, int color, T theta_s,
                                                        T theta_c)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nh = (b.ihi - b.ilo + 2) / 2, nj = b.jhi - b.jlo + 1;
    const int64_t total = (int64_t)nh * nj * (b.khi - b.klo + 1);
    if (t >= total) return;
    const int ii = (int)(t % nh);
    const int64_t r = t / nh;
    const int j = b.jlo + (int)(r % nj), k = b.klo + (int)(r / nj);
    const int i = b.ilo + 2 * ii + ((color + b.ilo + j + k) & 1);
    if (i > b.ihi) return;
    if (PHASE == 0) {
        T coef;
        const T dR = vx_res3<T, false>(l, i, j, k, &coef);
        l.VX(i, j, k) += dR / coef * theta_s;
    } else if (PHASE == 1) {
        T coef;
        const T dR = vy_res3<T, false>(l, i, j, k, &coef);
        l.VY(i, j, k) += dR / coef * theta_s;
    } else if (PHASE == 2) {
        T coef;
        const T dR = vz_res3<T, false>(l, i, j, k, &coef);
        l.VZ(i, j, k) += dR / coef * theta_s;
    } else {
        const T dR = pr_res3<T, false>(l, i, j, k);
        l.PR(i, j, k) += l.EN(i, j, k) * dR * theta_c;
    }
}

// OP false: residuals of the level (zeros on ghost nodes).
// OP true: A x of the outer solver: negated residuals with zero right-hand
// side and the pressure gauge p(1,1,1)/etan(1,1,1) in the continuity entry of
// cell (1,1,1).
template <typename T, bool OP>
__global__ void __launch_bounds__(256) mg3_residual_kernel(MG3LevelDev<T> l)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int X = l.xnum, Y = l.ynum, Z = l.znum;
    const int ni = Y + 1, nj = X + 1;
    if (t >= (int64_t)ni * nj * (Z + 1)) return;
    const int i = 1 + (int)(t % ni);
    const int64_t r = t / ni;
    const int j = 1 + (int)(r % nj), k = 1 + (int)(r / nj);
    const T sg = OP ? T(-1) : T(1);
    if (j <= X) {
        T v = T(0);
        if (i >= 2 && i <= Y && j >= 2 && j <= X - 1 && k >= 2 && k <= Z) v = sg * vx_res3<T, OP>(l, i, j, k, nullptr);
        l.rx[i3(Y + 1, X, i, j, k)] = v;
    }
    if (i <= Y) {
        T v = T(0);
        if (i >= 2 && i <= Y - 1 && j >= 2 && j <= X && k >= 2 && k <= Z) v = sg * vy_res3<T, OP>(l, i, j, k, nullptr);
        l.ry[i3(Y, X + 1, i, j, k)] = v;
    }
    if (k <= Z) {
        T v = T(0);
        if (i >= 2 && i <= Y && j >= 2 && j <= X && k >= 2 && k <= Z - 1) v = sg * vz_res3<T, OP>(l, i, j, k, nullptr);
        l.rz[i3(Y + 1, X + 1, i, j, k)] = v;
    }
    if (i <= Y - 1 && j <= X - 1 && k <= Z - 1) {
        T v;
        if (OP && i == 1 && j == 1 && k == 1) v = l.pr[0] / l.etan[0];
        else v = sg * pr_res3<T, OP>(l, i, j, k);
        l.rc[i3(Y - 1, X - 1, i, j, k)] = v;
    }
}

// A whole smoother call (optional zeroing of the solution, nsweeps x [ghost
// pass; four phases per colour], pressure normalisation, residuals) for a
// level small enough that kernel launches would dominate: one workgroup,
// a barrier between dependent steps.  Same arithmetic and order as the
// multi-launch path, so results are identical.
template <typename T>
__global__ void __launch_bounds__(1024) mg3_smoother_small_kernel(MG3LevelDev<T> l, int nsweeps, T theta_s,
                                                                  T theta_c, T prnorm, int zero_first)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    if (zero_first) {
        for (size_t t = tid; t < l.nvx(); t += nt) l.vx[t] = T(0);
        for (size_t t = tid; t < l.nvy(); t += nt) l.vy[t] = T(0);
        for (size_t t = tid; t < l.nvz(); t += nt) l.vz[t] = T(0);
        for (size_t t = tid; t < l.npr(); t += nt) l.pr[t] = T(0);
        __syncthreads();
    }
    const int nb = l.nbx + l.nby + l.nbz;
    for (int sw = 0; sw < nsweeps; ++sw) {
        for (int t = tid; t < nb; t += nt) {
            const T* base = t < l.nbx ? l.vx : (t < l.nbx + l.nby ? l.vy : l.vz);
            const int s = l.bc_src[t];
            l.bc_tmp[t] = s >= 0 ? base[s] : T(0);
        }
        __syncthreads();
        for (int t = tid; t < nb; t += nt) {
            T* base = t < l.nbx ? l.vx : (t < l.nbx + l.nby ? l.vy : l.vz);
            base[l.bc_dst[t]] = l.bc_tmp[t];
        }
        __syncthreads();
        for (int color = 0; color <= 1; ++color) {
            for (int phase = 0; phase < 4; ++phase) {
                const Box3 b = phase_box(l, phase);
                const int nh = (b.ihi - b.ilo + 2) / 2, nj = b.jhi - b.jlo + 1;
                const int total = nh * nj * (b.khi - b.klo + 1);
                for (int t = tid; t < total; t += nt) {
                    const int ii = t % nh, r = t / nh;
                    const int j = b.jlo + r % nj, k = b.klo + r / nj;
                    const int i = b.ilo + 2 * ii + ((color + b.ilo + j + k) & 1);
                    if (i > b.ihi) continue;
                    T coef;
                    if (phase == 0) {
                        const T dR = vx_res3<T, false>(l, i, j, k, &coef);
                        l.VX(i, j, k) += dR / coef * theta_s;
                    } else if (phase == 1) {
                        const T dR = vy_res3<T, false>(l, i, j, k, &coef);
                        l.VY(i, j, k) += dR / coef * theta_s;
                    } else if (phase == 2) {
                        const T dR = vz_res3<T, false>(l, i, j, k, &coef);
                        l.VZ(i, j, k) += dR / coef * theta_s;
                    } else {
                        const T dR = pr_res3<T, false>(l, i, j, k);
                        l.PR(i, j, k) += l.EN(i, j, k) * dR * theta_c;
                    }
                }
                __syncthreads();
            }
        }
    }
    const T dp = prnorm - l.pr[0];
    __syncthreads();
    for (size_t t = tid; t < l.npr(); t += nt) l.pr[t] += dp;
    __syncthreads();
    const int X = l.xnum, Y = l.ynum, Z = l.znum;
    const int ni = Y + 1, njj = X + 1, total = ni * njj * (Z + 1);
    for (int t = tid; t < total; t += nt) {
        const int i = 1 + t % ni, r = t / ni;
        const int j = 1 + r % njj, k = 1 + r / njj;
        if (j <= X) {
            T v = T(0);
            if (i >= 2 && i <= Y && j >= 2 && j <= X - 1 && k >= 2 && k <= Z) v = vx_res3<T, false>(l, i, j, k, nullptr);
            l.rx[i3(Y + 1, X, i, j, k)] = v;
        }
        if (i <= Y) {
            T v = T(0);
            if (i >= 2 && i <= Y - 1 && j >= 2 && j <= X && k >= 2 && k <= Z) v = vy_res3<T, false>(l, i, j, k, nullptr);
            l.ry[i3(Y, X + 1, i, j, k)] = v;
        }
        if (k <= Z) {
            T v = T(0);
            if (i >= 2 && i <= Y && j >= 2 && j <= X && k >= 2 && k <= Z - 1) v = vz_res3<T, false>(l, i, j, k, nullptr);
            l.rz[i3(Y + 1, X + 1, i, j, k)] = v;
        }
        if (i <= Y - 1 && j <= X - 1 && k <= Z - 1) l.rc[i3(Y - 1, X - 1, i, j, k)] = pr_res3<T, false>(l, i, j, k);
    }
}

template <typename T>
__global__ void mg3_pressure_shift_kernel(const T* pr, T prnorm, T* scratch)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) scratch[0] = prnorm - pr[0];
}
template <typename T>
__global__ void mg3_pressure_add_kernel(T* pr, int64_t n, const T* scratch)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) pr[t] += scratch[0];
}

// weight of fine node f (1-based) for coarse node c on one axis
template <typename T>
__device__ __forceinline__ T axis_w(const Axis3Dev<T>& a, int f, int c)
{
    const T d = a.d[f - 1];
    return a.idx[f - 1] == c ? T(1) - d : d;
}

// Gather restriction of one residual component (KIND 0 vx, 1 vy, 2 vz, 3
// continuity scaled by the fine etan); fine nodes in the order k, j, i.
template <typename T, int KIND>
__global__ void __launch_bounds__(256) mg3_restrict_kernel(MG3LevelDev<T> f, MG3LevelDev<T> c,
                                                           Transfer3Dev<T> m)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int yc = c.ynum, xc = c.xnum, zc = c.znum;
    int n1, n2, n3, ky, kx, kz, ry, rx, rz;
    if (KIND == 0) { n1 = yc + 1; n2 = xc; n3 = zc + 1; ky = K3_V; kx = K3_B; kz = K3_V; ry = R3_V; rx = R3_B; rz = R3_V; }
    else if (KIND == 1) { n1 = yc; n2 = xc + 1; n3 = zc + 1; ky = K3_B; kx = K3_V; kz = K3_V; ry = R3_B; rx = R3_V; rz = R3_V; }
    else if (KIND == 2) { n1 = yc + 1; n2 = xc + 1; n3 = zc; ky = K3_V; kx = K3_V; kz = K3_B; ry = R3_V; rx = R3_V; rz = R3_B; }
    else { n1 = yc - 1; n2 = xc - 1; n3 = zc - 1; ky = K3_P; kx = K3_P; kz = K3_P; ry = R3_P; rx = R3_P; rz = R3_P; }
    if (t >= (int64_t)n1 * n2 * n3) return;
    const int ic = 1 + (int)(t % n1);
    const int64_t r = t / n1;
    const int jc = 1 + (int)(r % n2), kc = 1 + (int)(r / n2);
    bool interior;
    if (KIND == 0) interior = ic > 1 && ic < yc + 1 && jc > 1 && jc < xc && kc > 1 && kc < zc + 1;
    else if (KIND == 1) interior = ic > 1 && ic < yc && jc > 1 && jc < xc + 1 && kc > 1 && kc < zc + 1;
    else if (KIND == 2) interior = ic > 1 && ic < yc + 1 && jc > 1 && jc < xc + 1 && kc > 1 && kc < zc;
    else interior = true;
    T num = T(0), den = T(0);
    if (interior) {
        const Axis3Dev<T> Ay = m.ax[0][ky], Ax = m.ax[1][kx], Az = m.ax[2][kz];
        const int ilo = m.rg[0][ry].lo[ic - 1], ihi = m.rg[0][ry].hi[ic - 1];
        const int jlo = m.rg[1][rx].lo[jc - 1], jhi = m.rg[1][rx].hi[jc - 1];
        const int klo = m.rg[2][rz].lo[kc - 1], khi = m.rg[2][rz].hi[kc - 1];
        for (int k = klo; k <= khi; ++k) {
            const T wz = axis_w(Az, k, kc);
            for (int j = jlo; j <= jhi; ++j) {
                const T wx = axis_w(Ax, j, jc);
                for (int i = ilo; i <= ihi; ++i) {
                    const T w = wx * axis_w(Ay, i, ic) * wz;
                    T val;
                    if (KIND == 0) val = f.rx[i3(f.ynum + 1, f.xnum, i, j, k)];
                    else if (KIND == 1) val = f.ry[i3(f.ynum, f.xnum + 1, i, j, k)];
                    else if (KIND == 2) val = f.rz[i3(f.ynum + 1, f.xnum + 1, i, j, k)];
                    else val = f.EN(i, j, k) * f.rc[i3(f.ynum - 1, f.xnum - 1, i, j, k)];
                    num += w * val;
                    den += w;
                }
            }
        }
    }
    const size_t o = i3(n1, n2, ic, jc, kc);
    if (KIND == 0) c.RX[o] = (interior && den != T(0)) ? num / den : T(0);
    else if (KIND == 1) c.RY[o] = (interior && den != T(0)) ? num / den : T(0);
    else if (KIND == 2) c.RZ[o] = (interior && den != T(0)) ? num / den : T(0);
    else c.RC[o] = den != T(0) ? num / den / c.EN(ic, jc, kc) : T(0);
}

// Gather viscosity restriction (KIND 0 etaxy, 1 etaxz, 2 etayz, 3 etan);
// fine nodes in the order i, j, k.
template <typename T, int KIND>
__global__ void __launch_bounds__(256) mg3_visc_restrict_kernel(MG3LevelDev<T> f, MG3LevelDev<T> c,
                                                                Transfer3Dev<T> m)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int yc = c.ynum, xc = c.xnum, zc = c.znum;
    int n1, n2, n3, ky, kx, kz, ry, rx, rz;
    if (KIND == 0) { n1 = yc; n2 = xc; n3 = zc - 1; ky = K3_B; kx = K3_B; kz = K3_S; ry = F3_B; rx = F3_B; rz = F3_S; }
    else if (KIND == 1) { n1 = yc - 1; n2 = xc; n3 = zc; ky = K3_S; kx = K3_B; kz = K3_B; ry = F3_S; rx = F3_B; rz = F3_B; }
    else if (KIND == 2) { n1 = yc; n2 = xc - 1; n3 = zc; ky = K3_B; kx = K3_S; kz = K3_B; ry = F3_B; rx = F3_S; rz = F3_B; }
    else { n1 = yc - 1; n2 = xc - 1; n3 = zc - 1; ky = K3_P; kx = K3_P; kz = K3_P; ry = F3_P; rx = F3_P; rz = F3_P; }
    if (t >= (int64_t)n1 * n2 * n3) return;
    const int ic = 1 + (int)(t % n1);
    const int64_t r = t / n1;
    const int jc = 1 + (int)(r % n2), kc = 1 + (int)(r / n2);
    const Axis3Dev<T> Ay = m.ax[0][ky], Ax = m.ax[1][kx], Az = m.ax[2][kz];
    const int ilo = m.rg[0][ry].lo[ic - 1], ihi = m.rg[0][ry].hi[ic - 1];
    const int jlo = m.rg[1][rx].lo[jc - 1], jhi = m.rg[1][rx].hi[jc - 1];
    const int klo = m.rg[2][rz].lo[kc - 1], khi = m.rg[2][rz].hi[kc - 1];
    T num = T(0), den = T(0);
    for (int i = ilo; i <= ihi; ++i) {
        const T wy = axis_w(Ay, i, ic);
        for (int j = jlo; j <= jhi; ++j) {
            const T wxy = axis_w(Ax, j, jc) * wy;
            for (int k = klo; k <= khi; ++k) {
                const T w = wxy * axis_w(Az, k, kc);
                T val;
                if (KIND == 0) val = f.EXY(i, j, k);
                else if (KIND == 1) val = f.EXZ(i, j, k);
                else if (KIND == 2) val = f.EYZ(i, j, k);
                else val = f.EN(i, j, k);
                num += w * val;
                den += w;
            }
        }
    }
    const T out = den != T(0) ? num / den : num;
    const size_t o = i3(n1, n2, ic, jc, kc);
    if (KIND == 0) c.etaxy[o] = out;
    else if (KIND == 1) c.etaxz[o] = out;
    else if (KIND == 2) c.etayz[o] = out;
    else c.etan[o] = out;
}

// trilinear coarse -> fine value (TriLinearInterpolation.jl:72-107): eight
// products added one by one in the reference's corner order
template <typename T>
__device__ __forceinline__ T c2f3(const T* a, int n1, int n2, int iy, int jx, int kz, T dy, T dx, T dz)
{
    const T omx = T(1) - dx, omy = T(1) - dy, omz = T(1) - dz;
    const size_t s1 = (size_t)n1, s2 = (size_t)n1 * n2;
    const T* p = a + i3(n1, n2, iy, jx, kz);
    T v = T(0);
    v += omx * omy * omz * p[0];
    v += omx * omy * dz * p[s2];
    v += omx * dy * omz * p[1];
    v += omx * dy * dz * p[1 + s2];
    v += dx * omy * omz * p[s1];
    v += dx * omy * dz * p[s1 + s2];
    v += dx * dy * omz * p[s1 + 1];
    v += dx * dy * dz * p[s1 + 1 + s2];
    return v;
}

template <typename T>
__global__ void __launch_bounds__(256) mg3_prolong_kernel(MG3LevelDev<T> c, MG3LevelDev<T> f,
                                                          Transfer3Dev<T> m, T relax_v, T relax_p)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int X = f.xnum, Y = f.ynum, Z = f.znum;
    const int ni = Y + 1, nj = X + 1;
    if (t >= (int64_t)ni * nj * (Z + 1)) return;
    const int i = 1 + (int)(t % ni);
    const int64_t r = t / ni;
    const int j = 1 + (int)(r % nj), k = 1 + (int)(r / nj);
    const Axis3Dev<T>(&A)[3][4] = m.ax;
#define MG3_C2F(arr, n1, n2, ky, kx, kz)                                                         \
    c2f3(arr, n1, n2, A[0][ky].idx[i - 1], A[1][kx].idx[j - 1], A[2][kz].idx[k - 1], A[0][ky].d[i - 1], \
         A[1][kx].d[j - 1], A[2][kz].d[k - 1])
    if (j <= X) f.VX(i, j, k) += MG3_C2F(c.vx, c.ynum + 1, c.xnum, K3_V, K3_B, K3_V) * relax_v;
    if (i <= Y) f.VY(i, j, k) += MG3_C2F(c.vy, c.ynum, c.xnum + 1, K3_B, K3_V, K3_V) * relax_v;
    if (k <= Z) f.VZ(i, j, k) += MG3_C2F(c.vz, c.ynum + 1, c.xnum + 1, K3_V, K3_V, K3_B) * relax_v;
    if (i <= Y - 1 && j <= X - 1 && k <= Z - 1)
        f.PR(i, j, k) += MG3_C2F(c.pr, c.ynum - 1, c.xnum - 1, K3_P, K3_P, K3_P) * relax_p;
#undef MG3_C2F
}

// ------------------------------------------------------- host-side hierarchy
// One axis of a level grid by node kind (coordinates) and the steps the
// stencil reads (basic and staggered).
struct HostAxis3 {
    int n;
    std::vector<double> b, stp_b, v, stp_v, p, s;
    const std::vector<double>& kind(int k) const { return k == K3_B ? b : k == K3_V ? v : k == K3_P ? p : s; }
};
HostAxis3 make_uniform_axis(int n, double size);
std::vector<std::array<int, 3>> mg3_level_sizes(int xnum, int ynum, int znum, int max_levels);

template <typename T>
struct MG3Hierarchy {
    std::vector<std::array<HostAxis3, 3>> axes;  // per level: x, y, z
    std::vector<MG3LevelDev<T>> lev;
    std::vector<Transfer3Dev<T>> xfer;           // xfer[k]: level k -> k+1
    std::vector<void*> allocations;
    T* scratch = nullptr;
    // levels with at most this many nodes ((ynum+1)(xnum+1)(znum+1)) smooth
    // in one single-workgroup launch (launch overhead dominates there)
    int64_t fused_nodes = 2000;
    int nlev() const { return (int)lev.size(); }
    ~MG3Hierarchy();
};

// clamp_weights: clamp normalised transfer distances to [0, 1] so boundary
// nodes of the non-nested pressure and shear-centre grids interpolate instead
// of extrapolating (the reference's maps extrapolate; the solver's
// preconditioner clamps)
template <typename T>
int mg3_build(MG3Hierarchy<T>& H, int xnum, int ynum, int znum, double xsize, double ysize, double zsize,
              int nlevels, bool clamp_weights);
template <typename T>
void mg3_set_bc(const MG3LevelDev<T>& l, hipStream_t s);
template <typename T>
void mg3_restrict_viscosity(MG3Hierarchy<T>& H, hipStream_t s);
template <typename T>
void mg3_smoother(MG3Hierarchy<T>& H, int k, int nsweeps, T theta_s, T theta_c, T prnorm, hipStream_t s);
template <typename T>
void mg3_restrict(MG3Hierarchy<T>& H, int k, hipStream_t s);
template <typename T>
void mg3_prolong(MG3Hierarchy<T>& H, int k, T relax_v, T relax_p, hipStream_t s);
template <typename T>
void mg3_zero_solution(MG3Hierarchy<T>& H, int k, hipStream_t s);
template <typename T>
void mg3_vcycle(MG3Hierarchy<T>& H, const std::vector<int>& smoothing, T theta_s, T theta_c, T relax_v,
                T relax_p, T prnorm, hipStream_t s);
template <typename T>
void mg3_apply(const MG3LevelDev<T>& op, hipStream_t s);

}  // namespace ebg
