// Host side of the 3-D multigrid hierarchy (level planning, uniform level
// grids, fine->coarse maps with gather ranges, the ghost-node tables, device
// allocation, smoother and V-cycle drivers) and the C ABI of the reference's
// experimental 3-D multigrid (multigrid_solver/MultigridSolver.jl:113-135 with
// the level, smoother, restriction and prolongation modules it drives), FP64.
#include "ebg_common.hpp"
#include "ebg_mg.hpp"
#include "ebg_mg3d.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <memory>
#include <numeric>

namespace ebg {

static inline unsigned nb3(int64_t n) { return (unsigned)((n + 255) / 256); }

// One uniform axis (Initialization.jl:47-74, VxGridTools.jl:40-99,
// PressureGridTools.jl:34-86) plus the shear-centre nodes of the etaxy z-axis,
// etaxz y-axis and etayz x-axis (SxyGridTools.jl:30-40, SxzGridTools.jl:28-40,
// SyzGridTools.jl:28-40).
HostAxis3 make_uniform_axis(int n, double size)
{
    HostAxis3 a;
    a.n = n;
    basic_axis(n, size / (double)(n - 1), a.b, a.stp_b);
    staggered_axis(n, a.b, a.stp_b, a.v, a.stp_v);
    pressure_axis(n, a.b, a.p);
    a.s.assign(n - 1, 0.0);
    a.s[0] = a.b[0] + a.stp_b[0] / 2.0;
    a.s[n - 2] = a.b[n - 1] - a.stp_b[n - 2] / 2.0;
    for (int i = 2; i <= n - 2; ++i) a.s[i - 1] = a.b[i - 1] + a.stp_b[i - 1] / 2.0;
    return a;
}

// multigrid/level_manager/LevelManager.jl:411-435
std::vector<std::array<int, 3>> mg3_level_sizes(int xnum, int ynum, int znum, int max_levels)
{
    std::vector<std::array<int, 3>> s{{xnum, ynum, znum}};
    while (std::min({s.back()[0], s.back()[1], s.back()[2]}) > 4) {
        const int xn = (s.back()[0] + 1) / 2, yn = (s.back()[1] + 1) / 2, zn = (s.back()[2] + 1) / 2;
        if (std::min({xn, yn, zn}) < 4 || (int)s.size() >= max_levels) break;
        s.push_back({xn, yn, zn});
    }
    return s;
}

// Ghost-node table of one velocity component: the reference's six face loops
// (BoundaryConditions.jl:92-174) with the free-slip rules of
// VxBoundaryConditions.jl:6-88, VyBoundaryConditions.jl:6-86 and
// VzBoundaryConditions.jl:6-87, run on symbols instead of values: sym[n] is
// the node whose pre-pass value node n holds (-1: zero).  Every boundary node
// is visited once, so the table (node, final symbol) reproduces the in-place
// serial pass when applied as gather-then-scatter.
static void bc_table(int xnum, int ynum, int znum, int comp, std::vector<int>& dst, std::vector<int>& src)
{
    int n1, n2, n3;
    if (comp == 0) { n1 = ynum + 1; n2 = xnum; n3 = znum + 1; }
    else if (comp == 1) { n1 = ynum; n2 = xnum + 1; n3 = znum + 1; }
    else { n1 = ynum + 1; n2 = xnum + 1; n3 = znum; }
    std::vector<int> sym((size_t)n1 * n2 * n3);
    std::iota(sym.begin(), sym.end(), 0);
    auto I = [&](int i, int j, int k) { return (size_t)i3(n1, n2, i, j, k); };
    dst.clear();
    auto set = [&](int i, int j, int k) {
        int v = -1;
        if (comp == 0) {
            if (i == 1) v = sym[I(i + 1, j, k)];
            if (i == n1) v = sym[I(i - 1, j, k)];
            if (k == 1) v = sym[I(i, j, k + 1)];
            if (k == n3) v = sym[I(i, j, k - 1)];
        } else if (comp == 1) {
            if (j == 1) v = sym[I(i, j + 1, k)];
            if (j == n2) v = sym[I(i, j - 1, k)];
            if (k == 1) v = sym[I(i, j, k + 1)];
            if (k == n3) v = sym[I(i, j, k - 1)];
        } else {
            if (i == 1) v = sym[I(i + 1, j, k)];
            if (i == n1) v = sym[I(i - 1, j, k)];
            if (j == 1) v = sym[I(i, j + 1, k)];
            if (j == n2) v = sym[I(i, j - 1, k)];
        }
        sym[I(i, j, k)] = v;
        dst.push_back((int)I(i, j, k));
    };
    for (int j = 1; j <= n2; ++j)
        for (int i = 1; i <= n1; ++i) { set(i, j, 1); set(i, j, n3); }
    for (int k = 2; k <= n3 - 1; ++k)
        for (int j = 1; j <= n2; ++j) { set(1, j, k); set(n1, j, k); }
    for (int k = 2; k <= n3 - 1; ++k)
        for (int i = 2; i <= n1 - 1; ++i) { set(i, 1, k); set(i, n2, k); }
    src.resize(dst.size());
    for (size_t e = 0; e < dst.size(); ++e) src[e] = sym[(size_t)dst[e]];
}

template <typename T>
MG3Hierarchy<T>::~MG3Hierarchy()
{
    for (void* p : allocations) (void)hipFree(p);
}

template <typename T>
static int alloc3(MG3Hierarchy<T>& H, size_t n, T** out)
{
    void* p = nullptr;
    EBG_HIP(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    H.allocations.push_back(p);
    EBG_HIP(hipMemset(p, 0, (n ? n : 1) * sizeof(T)));
    *out = (T*)p;
    return EBG_OK;
}

template <typename T, typename S>
static int upload3(MG3Hierarchy<T>& H, const std::vector<S>& v, const S** out)
{
    void* p = nullptr;
    const size_t bytes = (v.empty() ? 1 : v.size()) * sizeof(S);
    EBG_HIP(hipMalloc(&p, bytes));
    H.allocations.push_back(p);
    if (!v.empty()) EBG_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(S), hipMemcpyHostToDevice));
    *out = (const S*)p;
    return EBG_OK;
}

template <typename T>
static int upload_steps(MG3Hierarchy<T>& H, const std::vector<double>& v, const T** out)
{
    std::vector<T> t(v.begin(), v.end());
    return upload3<T, T>(H, t, out);
}

template <typename T>
int mg3_build(MG3Hierarchy<T>& H, int xnum, int ynum, int znum, double xsize, double ysize, double zsize,
              int nlevels, bool clamp_weights)
{
    auto sizes = mg3_level_sizes(xnum, ynum, znum, 10);
    if (nlevels > 0 && nlevels < (int)sizes.size()) sizes.resize(nlevels);
    const int nl = (int)sizes.size();
    H.axes.resize(nl);
    H.lev.resize(nl);
    int rc;
    for (int k = 0; k < nl; ++k) {
        const int xn = sizes[k][0], yn = sizes[k][1], zn = sizes[k][2];
        H.axes[k] = {make_uniform_axis(xn, xsize), make_uniform_axis(yn, ysize), make_uniform_axis(zn, zsize)};
        MG3LevelDev<T>& l = H.lev[k];
        l = MG3LevelDev<T>{};
        l.xnum = xn;
        l.ynum = yn;
        l.znum = zn;
        const HostAxis3 &ax = H.axes[k][0], &ay = H.axes[k][1], &az = H.axes[k][2];
        if ((rc = upload_steps(H, ax.stp_b, &l.xstp_b)) || (rc = upload_steps(H, ay.stp_b, &l.ystp_b)) ||
            (rc = upload_steps(H, az.stp_b, &l.zstp_b)) || (rc = upload_steps(H, ay.stp_v, &l.ystp_vx)) ||
            (rc = upload_steps(H, az.stp_v, &l.zstp_vx)) || (rc = upload_steps(H, ax.stp_v, &l.xstp_vy)))
            return rc;
        const size_t nvx = l.nvx(), nvy = l.nvy(), nvz = l.nvz(), npr = l.npr();
        const size_t nxy = (size_t)yn * xn * (zn - 1), nxz = (size_t)(yn - 1) * xn * zn,
                     nyz = (size_t)yn * (xn - 1) * zn;
        if ((rc = alloc3(H, npr, &l.etan)) || (rc = alloc3(H, nxy, &l.etaxy)) || (rc = alloc3(H, nxz, &l.etaxz)) ||
            (rc = alloc3(H, nyz, &l.etayz)) || (rc = alloc3(H, nvx, &l.vx)) || (rc = alloc3(H, nvy, &l.vy)) ||
            (rc = alloc3(H, nvz, &l.vz)) || (rc = alloc3(H, npr, &l.pr)) || (rc = alloc3(H, nvx, &l.RX)) ||
            (rc = alloc3(H, nvy, &l.RY)) || (rc = alloc3(H, nvz, &l.RZ)) || (rc = alloc3(H, npr, &l.RC)) ||
            (rc = alloc3(H, nvx, &l.rx)) || (rc = alloc3(H, nvy, &l.ry)) || (rc = alloc3(H, nvz, &l.rz)) ||
            (rc = alloc3(H, npr, &l.rc)))
            return rc;
        std::vector<int> dst, src, d1, s1;
        for (int c = 0; c < 3; ++c) {
            bc_table(xn, yn, zn, c, d1, s1);
            (c == 0 ? l.nbx : c == 1 ? l.nby : l.nbz) = (int)d1.size();
            dst.insert(dst.end(), d1.begin(), d1.end());
            src.insert(src.end(), s1.begin(), s1.end());
        }
        if ((rc = upload3<T, int>(H, dst, &l.bc_dst)) || (rc = upload3<T, int>(H, src, &l.bc_src)) ||
            (rc = alloc3(H, dst.size(), &l.bc_tmp)))
            return rc;
    }
    H.xfer.resize(nl > 0 ? nl - 1 : 0);
    for (int k = 0; k + 1 < nl; ++k) {
        Transfer3Dev<T>& t = H.xfer[k];
        for (int a = 0; a < 3; ++a) {
            // transfer axes are ordered y, x, z (the array index order); the
            // level axes are stored x, y, z
            const int h = a == 0 ? 1 : (a == 1 ? 0 : 2);
            const HostAxis3& f = H.axes[k][h];
            const HostAxis3& c = H.axes[k + 1][h];
            const int nf = f.n;
            for (int kind = 0; kind < 4; ++kind) {
                auto m = axis_map<T>(f.kind(kind), c.kind(kind), 1, 0, clamp_weights);
                if ((rc = upload3<T, int>(H, m.idx, &t.ax[a][kind].idx)) ||
                    (rc = upload3<T, T>(H, m.d, &t.ax[a][kind].d)))
                    return rc;
            }
            // fine loop bounds per range (kind, first, last)
            const int spec[NR3][3] = {{K3_V, 2, nf}, {K3_B, 2, nf - 1}, {K3_P, 2, nf - 1},
                                      {K3_B, 1, nf}, {K3_S, 1, nf - 1}, {K3_P, 1, nf - 1}};
            for (int r = 0; r < NR3; ++r) {
                auto m = axis_map<T>(f.kind(spec[r][0]), c.kind(spec[r][0]), spec[r][1], spec[r][2], false);
                if ((rc = upload3<T, int>(H, m.lo, &t.rg[a][r].lo)) || (rc = upload3<T, int>(H, m.hi, &t.rg[a][r].hi)))
                    return rc;
            }
        }
    }
    if ((rc = alloc3(H, 8, &H.scratch))) return rc;
    return EBG_OK;
}

template <typename T>
void mg3_set_bc(const MG3LevelDev<T>& l, hipStream_t s)
{
    const int n = l.nbx + l.nby + l.nbz;
    hipLaunchKernelGGL(mg3_bc_gather_kernel<T>, dim3(nb3(n)), dim3(256), 0, s, l);
    hipLaunchKernelGGL(mg3_bc_scatter_kernel<T>, dim3(nb3(n)), dim3(256), 0, s, l);
}

template <typename T>
void mg3_restrict_viscosity(MG3Hierarchy<T>& H, hipStream_t s)
{
    for (int k = 0; k + 1 < H.nlev(); ++k) {
        const MG3LevelDev<T>&f = H.lev[k], &c = H.lev[k + 1];
        const int64_t yc = c.ynum, xc = c.xnum, zc = c.znum;
        hipLaunchKernelGGL((mg3_visc_restrict_kernel<T, 0>), dim3(nb3(yc * xc * (zc - 1))), dim3(256), 0, s, f, c, H.xfer[k]);
        hipLaunchKernelGGL((mg3_visc_restrict_kernel<T, 1>), dim3(nb3((yc - 1) * xc * zc)), dim3(256), 0, s, f, c, H.xfer[k]);
        hipLaunchKernelGGL((mg3_visc_restrict_kernel<T, 2>), dim3(nb3(yc * (xc - 1) * zc)), dim3(256), 0, s, f, c, H.xfer[k]);
        hipLaunchKernelGGL((mg3_visc_restrict_kernel<T, 3>), dim3(nb3((yc - 1) * (xc - 1) * (zc - 1))), dim3(256), 0, s,
                           f, c, H.xfer[k]);
    }
}

template <typename T, int PHASE>
static void launch_phase(const MG3LevelDev<T>& l, int color, T ts, T tc, hipStream_t s)
{
    const Box3 b = phase_box(l, PHASE);
    const int64_t n = phase_threads(b);
    if (n <= 0) return;
    hipLaunchKernelGGL((mg3_sweep_kernel<T, PHASE>), dim3(nb3(n)), dim3(256), 0, s, l, b, color, ts, tc);
}

template <typename T>
static void residuals(const MG3LevelDev<T>& l, hipStream_t s)
{
    const int64_t n = (int64_t)(l.ynum + 1) * (l.xnum + 1) * (l.znum + 1);
    hipLaunchKernelGGL((mg3_residual_kernel<T, false>), dim3(nb3(n)), dim3(256), 0, s, l);
}

template <typename T>
void mg3_zero_solution(MG3Hierarchy<T>& H, int k, hipStream_t s)
{
    const MG3LevelDev<T>& l = H.lev[k];
    hipLaunchKernelGGL(fill_kernel<T>, dim3(nb3((int64_t)l.nvx())), dim3(256), 0, s, l.vx, (int64_t)l.nvx(), T(0));
    hipLaunchKernelGGL(fill_kernel<T>, dim3(nb3((int64_t)l.nvy())), dim3(256), 0, s, l.vy, (int64_t)l.nvy(), T(0));
    hipLaunchKernelGGL(fill_kernel<T>, dim3(nb3((int64_t)l.nvz())), dim3(256), 0, s, l.vz, (int64_t)l.nvz(), T(0));
    hipLaunchKernelGGL(fill_kernel<T>, dim3(nb3((int64_t)l.npr())), dim3(256), 0, s, l.pr, (int64_t)l.npr(), T(0));
}

// smoother/Smoother.jl:50-64 with StokesSolverGaussSeidel.jl:16-28; levels
// of at most H.fused_nodes nodes run as one single-workgroup launch
template <typename T>
static void smoother_impl(MG3Hierarchy<T>& H, int k, int nsweeps, T theta_s, T theta_c, T prnorm, bool zero_first,
                          hipStream_t s)
{
    const MG3LevelDev<T>& l = H.lev[k];
    const int64_t nodes = (int64_t)(l.ynum + 1) * (l.xnum + 1) * (l.znum + 1);
    if (nodes <= H.fused_nodes) {
        hipLaunchKernelGGL(mg3_smoother_small_kernel<T>, dim3(1), dim3(1024), 0, s, l, nsweeps, theta_s, theta_c,
                           prnorm, zero_first ? 1 : 0);
        return;
    }
    if (zero_first) mg3_zero_solution(H, k, s);
    for (int sw = 0; sw < nsweeps; ++sw) {
        mg3_set_bc(l, s);
        for (int color = 0; color <= 1; ++color) {
            launch_phase<T, 0>(l, color, theta_s, theta_c, s);
            launch_phase<T, 1>(l, color, theta_s, theta_c, s);
            launch_phase<T, 2>(l, color, theta_s, theta_c, s);
            launch_phase<T, 3>(l, color, theta_s, theta_c, s);
        }
    }
    const int64_t np = (int64_t)l.npr();
    hipLaunchKernelGGL(mg3_pressure_shift_kernel<T>, dim3(1), dim3(64), 0, s, l.pr, prnorm, H.scratch);
    hipLaunchKernelGGL(mg3_pressure_add_kernel<T>, dim3(nb3(np)), dim3(256), 0, s, l.pr, np, (const T*)H.scratch);
    residuals(l, s);
}

template <typename T>
void mg3_smoother(MG3Hierarchy<T>& H, int k, int nsweeps, T theta_s, T theta_c, T prnorm, hipStream_t s)
{
    smoother_impl(H, k, nsweeps, theta_s, theta_c, prnorm, false, s);
}

// restriction/Restriction.jl:64-376
template <typename T>
void mg3_restrict(MG3Hierarchy<T>& H, int k, hipStream_t s)
{
    const MG3LevelDev<T>&f = H.lev[k], &c = H.lev[k + 1];
    hipLaunchKernelGGL((mg3_restrict_kernel<T, 0>), dim3(nb3((int64_t)c.nvx())), dim3(256), 0, s, f, c, H.xfer[k]);
    hipLaunchKernelGGL((mg3_restrict_kernel<T, 1>), dim3(nb3((int64_t)c.nvy())), dim3(256), 0, s, f, c, H.xfer[k]);
    hipLaunchKernelGGL((mg3_restrict_kernel<T, 2>), dim3(nb3((int64_t)c.nvz())), dim3(256), 0, s, f, c, H.xfer[k]);
    hipLaunchKernelGGL((mg3_restrict_kernel<T, 3>), dim3(nb3((int64_t)c.npr())), dim3(256), 0, s, f, c, H.xfer[k]);
}

// prolongation/Prolongation.jl:60-140 with SmoothAndProlongate.jl:39-56:
// level k's correction added to level k-1
template <typename T>
void mg3_prolong(MG3Hierarchy<T>& H, int k, T relax_v, T relax_p, hipStream_t s)
{
    const MG3LevelDev<T>&c = H.lev[k], &f = H.lev[k - 1];
    const int64_t n = (int64_t)(f.ynum + 1) * (f.xnum + 1) * (f.znum + 1);
    hipLaunchKernelGGL(mg3_prolong_kernel<T>, dim3(nb3(n)), dim3(256), 0, s, c, f, H.xfer[k - 1], relax_v, relax_p);
}

// vcycle/core/SmoothAndRestrict.jl:15-46 then SmoothAndProlongate.jl:14-60
template <typename T>
void mg3_vcycle(MG3Hierarchy<T>& H, const std::vector<int>& sm, T theta_s, T theta_c, T relax_v, T relax_p,
                T prnorm, hipStream_t s)
{
    const int nl = H.nlev();
    for (int n = 0; n < nl; ++n) {
        smoother_impl(H, n, sm[n], theta_s, theta_c, prnorm, n > 0, s);
        if (n < nl - 1) mg3_restrict(H, n, s);
    }
    for (int n = nl - 1; n >= 0; --n) {
        if (n == 0) {
            mg3_smoother(H, 0, sm[0], theta_s, theta_c, prnorm, s);
        } else {
            mg3_smoother(H, n, sm[n], theta_s, theta_c, T(0), s);
            mg3_prolong(H, n, relax_v, relax_p, s);
        }
    }
}

// A x of the outer solver on a level view whose vx..pr hold x and rx..rc
// receive A x (ghost slots of x are refreshed in place first)
template <typename T>
void mg3_apply(const MG3LevelDev<T>& op, hipStream_t s)
{
    mg3_set_bc(op, s);
    const int64_t n = (int64_t)(op.ynum + 1) * (op.xnum + 1) * (op.znum + 1);
    hipLaunchKernelGGL((mg3_residual_kernel<T, true>), dim3(nb3(n)), dim3(256), 0, s, op);
}

#define EBG_MG3_INST(T)                                                                                   \
    template struct MG3Hierarchy<T>;                                                                      \
    template int mg3_build<T>(MG3Hierarchy<T>&, int, int, int, double, double, double, int, bool);       \
    template void mg3_set_bc<T>(const MG3LevelDev<T>&, hipStream_t);                                     \
    template void mg3_restrict_viscosity<T>(MG3Hierarchy<T>&, hipStream_t);                              \
    template void mg3_smoother<T>(MG3Hierarchy<T>&, int, int, T, T, T, hipStream_t);                     \
    template void mg3_restrict<T>(MG3Hierarchy<T>&, int, hipStream_t);                                   \
    template void mg3_prolong<T>(MG3Hierarchy<T>&, int, T, T, hipStream_t);                              \
    template void mg3_zero_solution<T>(MG3Hierarchy<T>&, int, hipStream_t);                              \
    template void mg3_vcycle<T>(MG3Hierarchy<T>&, const std::vector<int>&, T, T, T, T, T, hipStream_t); \
    template void mg3_apply<T>(const MG3LevelDev<T>&, hipStream_t);
EBG_MG3_INST(double)
EBG_MG3_INST(float)

// ------------------------------------------------- reference 3-D multigrid ABI
namespace {
struct MG3Ref {
    MG3Hierarchy<double> H;
    // original (unscaled) viscosity per level for the continuation and the
    // reduction partials (4 * MG_RUN_BLOCKS); allocated by the first ebg_mg3d_run
    std::vector<std::array<double*, 4>> orig;
    double* partial = nullptr;
    // the global level 0 of the multimulti driver: vx, vy, vz, pr, RX, RY, RZ,
    // RC, rx, ry, rz, rc (allocated by the first multimulti run)
    std::array<double*, 12> lev0{};
    ~MG3Ref()
    {
        for (auto& a : orig)
            for (double* p : a) (void)hipFree(p);
        if (partial) (void)hipFree(partial);
        for (double* p : lev0)
            if (p) (void)hipFree(p);
    }
};

__global__ void mm3_add_kernel(double* __restrict__ a, const double* __restrict__ b, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] += b[t];
}

// per-block sums of the squared scaled interior residuals of the finest level
// (CalculateMeanResiduals.jl:14-70): part[4*b + {x, y, z, c}]
__global__ void __launch_bounds__(256) mean_residual_partial_kernel(MG3LevelDev<double> l, double inv_ss,
                                                                    double inv_cc, double* __restrict__ part)
{
    __shared__ double sh[4][256];
    const int X = l.xnum, Y = l.ynum, Z = l.znum;
    const int64_t total = (int64_t)Y * X * Z;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(t % Y);
        const int64_t r = t / Y;
        const int j = 1 + (int)(r % X), k = 1 + (int)(r / X);
        if (i > 1 && j > 1 && k > 1 && j < X) {
            const double v = l.rx[i3(Y + 1, X, i, j, k)] * inv_ss;
            s[0] += v * v;
        }
        if (i > 1 && j > 1 && k > 1 && i < Y) {
            const double v = l.ry[i3(Y, X + 1, i, j, k)] * inv_ss;
            s[1] += v * v;
        }
        if (i > 1 && j > 1 && k > 1 && k < Z) {
            const double v = l.rz[i3(Y + 1, X + 1, i, j, k)] * inv_ss;
            s[2] += v * v;
        }
        if (j < X && i < Y && k < Z) {
            const double v = l.rc[i3(Y - 1, X - 1, i, j, k)] * inv_cc;
            s[3] += v * v;
        }
    }
    for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

MG3Ref* get3(ebg_ctx* ctx) { return ctx ? (MG3Ref*)ctx->mg3_ref : nullptr; }

double* field3(MG3LevelDev<double>& l, int which, size_t* n)
{
    const size_t ny = l.ynum, nx = l.xnum, nz = l.znum;
    switch (which) {
        case EBG_MG3_ETAN: *n = l.npr(); return l.etan;
        case EBG_MG3_ETAXY: *n = ny * nx * (nz - 1); return l.etaxy;
        case EBG_MG3_ETAXZ: *n = (ny - 1) * nx * nz; return l.etaxz;
        case EBG_MG3_ETAYZ: *n = ny * (nx - 1) * nz; return l.etayz;
        case EBG_MG3_VX: *n = l.nvx(); return l.vx;
        case EBG_MG3_VY: *n = l.nvy(); return l.vy;
        case EBG_MG3_VZ: *n = l.nvz(); return l.vz;
        case EBG_MG3_PR: *n = l.npr(); return l.pr;
        case EBG_MG3_RX: *n = l.nvx(); return l.RX;
        case EBG_MG3_RY: *n = l.nvy(); return l.RY;
        case EBG_MG3_RZ: *n = l.nvz(); return l.RZ;
        case EBG_MG3_RC: *n = l.npr(); return l.RC;
        case EBG_MG3_RESX: *n = l.nvx(); return l.rx;
        case EBG_MG3_RESY: *n = l.nvy(); return l.ry;
        case EBG_MG3_RESZ: *n = l.nvz(); return l.rz;
        case EBG_MG3_RESC: *n = l.npr(); return l.rc;
        default: *n = 0; return nullptr;
    }
}
}  // namespace

int mg3ref_release(ebg_ctx* ctx)
{
    if (ctx && ctx->mg3_ref) {
        delete (MG3Ref*)ctx->mg3_ref;
        ctx->mg3_ref = nullptr;
    }
    return EBG_OK;
}

}  // namespace ebg

using namespace ebg;

extern "C" {

int ebg_mg3d_create(ebg_ctx* ctx, int64_t xnum, int64_t ynum, int64_t znum, double xsize, double ysize,
                    double zsize, int nlevels, int* levels_out)
{
    EBG_CHECK(ctx, EBG_ERR_ARG, "ebg_mg3d_create: ctx is NULL");
    EBG_CHECK(xnum >= 4 && ynum >= 4 && znum >= 4, EBG_ERR_ARG, "ebg_mg3d_create: need xnum, ynum, znum >= 4");
    EBG_CHECK((double)(ynum + 1) * (double)(xnum + 1) * (double)(znum + 1) < 2.0e9, EBG_ERR_ARG,
              "ebg_mg3d_create: level arrays must hold fewer than 2^31 entries");
    EBG_CHECK(xsize > 0 && ysize > 0 && zsize > 0, EBG_ERR_ARG, "ebg_mg3d_create: sizes must be > 0");
    EBG_HIP(hipSetDevice(ctx->device));
    EBG_HIP(hipStreamSynchronize(ctx->stream));
    mg3ref_release(ctx);
    auto m = std::make_unique<MG3Ref>();
    int rc = mg3_build(m->H, (int)xnum, (int)ynum, (int)znum, xsize, ysize, zsize, nlevels, false);
    if (rc) return rc;
    if (levels_out) *levels_out = m->H.nlev();
    ctx->mg3_ref = m.release();
    return EBG_OK;
}

int ebg_mg3d_set_fused_limit(ebg_ctx* ctx, int64_t nodes)
{
    MG3Ref* m = get3(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg3d_set_fused_limit: no hierarchy");
    EBG_CHECK(nodes >= 0, EBG_ERR_ARG, "ebg_mg3d_set_fused_limit: nodes must be >= 0");
    m->H.fused_nodes = nodes;
    return EBG_OK;
}

int ebg_mg3d_level_shape(ebg_ctx* ctx, int level, int64_t* xnum, int64_t* ynum, int64_t* znum)
{
    MG3Ref* m = get3(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg3d_level_shape: no hierarchy");
    EBG_CHECK(level >= 0 && level < m->H.nlev(), EBG_ERR_ARG, "ebg_mg3d_level_shape: bad level");
    *xnum = m->H.lev[level].xnum;
    *ynum = m->H.lev[level].ynum;
    *znum = m->H.lev[level].znum;
    return EBG_OK;
}

int ebg_mg3d_copy_field(ebg_ctx* ctx, int level, int which, double* dst, const double* src)
{
    MG3Ref* m = get3(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg3d_copy_field: no hierarchy");
    EBG_CHECK(level >= 0 && level < m->H.nlev(), EBG_ERR_ARG, "ebg_mg3d_copy_field: bad level");
    EBG_CHECK((dst == nullptr) != (src == nullptr), EBG_ERR_ARG,
              "ebg_mg3d_copy_field: give exactly one of dst (read) or src (write)");
    size_t n = 0;
    double* f = field3(m->H.lev[level], which, &n);
    EBG_CHECK(f, EBG_ERR_ARG, "ebg_mg3d_copy_field: unknown field id");
    if (dst) EBG_HIP(hipMemcpyAsync(dst, f, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    else EBG_HIP(hipMemcpyAsync(f, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return EBG_OK;
}

int ebg_mg3d_restrict_viscosity(ebg_ctx* ctx)
{
    MG3Ref* m = get3(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg3d_restrict_viscosity: no hierarchy");
    mg3_restrict_viscosity(m->H, ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_mg3d_set_bc(ebg_ctx* ctx, int level)
{
    MG3Ref* m = get3(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg3d_set_bc: no hierarchy");
    EBG_CHECK(level >= 0 && level < m->H.nlev(), EBG_ERR_ARG, "ebg_mg3d_set_bc: bad level");
    mg3_set_bc(m->H.lev[level], ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_mg3d_smoother(ebg_ctx* ctx, int level, int nsweeps, double theta_s, double theta_c, double prnorm)
{
    MG3Ref* m = get3(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg3d_smoother: no hierarchy");
    EBG_CHECK(level >= 0 && level < m->H.nlev() && nsweeps >= 0, EBG_ERR_ARG,
              "ebg_mg3d_smoother: bad level or sweep count");
    mg3_smoother(m->H, level, nsweeps, theta_s, theta_c, prnorm, ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_mg3d_restrict(ebg_ctx* ctx, int level)
{
    MG3Ref* m = get3(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg3d_restrict: no hierarchy");
    EBG_CHECK(level >= 0 && level + 1 < m->H.nlev(), EBG_ERR_ARG, "ebg_mg3d_restrict: bad level");
    mg3_restrict(m->H, level, ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_mg3d_prolong(ebg_ctx* ctx, int level, double relax_v, double relax_p)
{
    MG3Ref* m = get3(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg3d_prolong: no hierarchy");
    EBG_CHECK(level >= 1 && level < m->H.nlev(), EBG_ERR_ARG, "ebg_mg3d_prolong: bad level");
    mg3_prolong(m->H, level, relax_v, relax_p, ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_mg3d_vcycles(ebg_ctx* ctx, int ncycles, int finest_sweeps, int multiplier, int cap, double theta_s,
                     double theta_c, double relax_v, double relax_p, double prnorm)
{
    MG3Ref* m = get3(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg3d_vcycles: no hierarchy");
    EBG_CHECK(ncycles >= 0 && finest_sweeps >= 0 && multiplier >= 1, EBG_ERR_ARG,
              "ebg_mg3d_vcycles: bad cycle parameters");
    auto sm = mg_smoothing_counts(m->H.nlev(), finest_sweeps, multiplier, cap);
    for (int c = 0; c < ncycles; ++c) mg3_vcycle(m->H, sm, theta_s, theta_c, relax_v, relax_p, prnorm, ctx->stream);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

// MultigridSolver.run_multigrid_solver (MultigridSolver.jl:113-135) with
// use_multimulti = false: viscosity restriction, saved originals, viscosity
// continuation from constant viscosity (MultigridDataManager.jl:349-373,
// ViscosityScaling.jl:11-24, 41-60, 81-110), then V-cycles with mean principal residuals
// (CalculateMeanResiduals.jl:14-70) and the convergence test
// (MultigridVCycle.jl:79-135, 390-412, 436-452).
static int mg3d_run_impl(ebg_ctx* ctx, const ebg_mg_run_params* prm, bool multimulti, int* cycles_done,
                         int* converged, double* history, int* global_updates, double* global_history,
                         int global_capacity)
{
    MG3Ref* m = get3(ctx);
    EBG_CHECK(m, EBG_ERR_STATE, "ebg_mg3d_run: no hierarchy");
    EBG_CHECK(prm, EBG_ERR_ARG, "ebg_mg3d_run: params is NULL");
    const ebg_mg_run_params& p = *prm;
    EBG_CHECK(p.nvcycles >= 0 && p.finest_sweeps >= 0 && p.sweep_multiplier >= 1 && p.nviscosity_jumps > 0 &&
                  p.stokesscale != 0 && p.continscale != 0,
              EBG_ERR_ARG, "ebg_mg3d_run: bad parameters");
    hipStream_t s = ctx->stream;
    MG3Hierarchy<double>& H = m->H;
    const int nl = H.nlev();
    auto nvisc = [&](int k, int f) -> int64_t {
        const int64_t y = H.lev[k].ynum, x = H.lev[k].xnum, z = H.lev[k].znum;
        switch (f) {
            case 0: return (y - 1) * (x - 1) * (z - 1);
            case 1: return y * x * (z - 1);
            case 2: return (y - 1) * x * z;
            default: return y * (x - 1) * z;
        }
    };
    auto field = [&](int k, int f) {
        const MG3LevelDev<double>& l = H.lev[k];
        return f == 0 ? l.etan : f == 1 ? l.etaxy : f == 2 ? l.etaxz : l.etayz;
    };
    if (!m->partial) {   // first run (or a failed first allocation): allocate everything
        for (auto& a : m->orig)
            for (double*& q : a) {
                if (q) (void)hipFree(q);
                q = nullptr;
            }
        m->orig.assign(nl, {nullptr, nullptr, nullptr, nullptr});
        for (int k = 0; k < nl; ++k)
            for (int f = 0; f < 4; ++f) EBG_HIP(hipMalloc(&m->orig[k][f], (size_t)nvisc(k, f) * sizeof(double)));
        EBG_HIP(hipMalloc(&m->partial, (size_t)MG_RUN_BLOCKS * 4 * sizeof(double)));
    }
    MG3LevelDev<double>& l1 = H.lev[0];
    const int64_t sizes[4] = {(int64_t)l1.nvx(), (int64_t)l1.nvy(), (int64_t)l1.nvz(), (int64_t)l1.npr()};
    MG3LevelDev<double> l0v = l1;   // level 0 view
    auto global_residual = [&]() {
        // Smoother.calculate_residuals_without_smoothing on level 0, then
        // level 1's right-hand sides (Smoother.jl:101-141)
        mg3_set_bc(l0v, s);
        hipLaunchKernelGGL(mg3_pressure_shift_kernel<double>, dim3(1), dim3(64), 0, s, l0v.pr, p.pressure_bc,
                           H.scratch);
        hipLaunchKernelGGL(mg3_pressure_add_kernel<double>, dim3(nb3(sizes[3])), dim3(256), 0, s, l0v.pr, sizes[3],
                           (const double*)H.scratch);
        residuals(l0v, s);
        const double* src[4] = {l0v.rx, l0v.ry, l0v.rz, l0v.rc};
        double* dst[4] = {l1.RX, l1.RY, l1.RZ, l1.RC};
        for (int q = 0; q < 4; ++q)
            (void)hipMemcpyAsync(dst[q], src[q], sizes[q] * sizeof(double), hipMemcpyDeviceToDevice, s);
    };
    if (multimulti) {
        if (!m->lev0[0])
            for (int q = 0; q < 12; ++q) EBG_HIP(hipMalloc(&m->lev0[q], (size_t)sizes[q % 4] * sizeof(double)));
        double** v0[12] = {&l0v.vx, &l0v.vy, &l0v.vz, &l0v.pr, &l0v.RX, &l0v.RY, &l0v.RZ, &l0v.RC,
                           &l0v.rx, &l0v.ry, &l0v.rz, &l0v.rc};
        for (int q = 0; q < 12; ++q) *v0[q] = m->lev0[q];
        // the caller's initial guess and right-hand sides belong to level 0
        double* from[8] = {l1.vx, l1.vy, l1.vz, l1.pr, l1.RX, l1.RY, l1.RZ, l1.RC};
        for (int q = 0; q < 8; ++q)
            EBG_HIP(hipMemcpyAsync(*v0[q], from[q], sizes[q % 4] * sizeof(double), hipMemcpyDeviceToDevice, s));
        mg3_zero_solution(H, 0, s);
        global_residual();
        EBG_HIP(hipGetLastError());
    }
    int rc;
    // set_initial_viscosity_scaling_parameters: range of the finest level
    double vmin = INFINITY, vmax = -INFINITY;
    for (int f = 0; f < 4; ++f)
        if ((rc = mg_device_range(field(0, f), nvisc(0, f), m->partial, s, vmin, vmax))) return rc;
    EBG_CHECK(vmin > 0.0 && std::isfinite(vmax), EBG_ERR_ARG, "ebg_mg3d_run: viscosity must be positive and finite");
    ViscRamp ramp;
    ramp.init(vmin, vmax, p.nviscosity_jumps);
    const double max_start = vmin;
    // interpolate_viscosity_to_coarser_levels! then save_original_viscosity_on_all_layers!
    mg3_restrict_viscosity(H, s);
    double omin = INFINITY, omax = -INFINITY;
    for (int k = 0; k < nl; ++k)
        for (int f = 0; f < 4; ++f) {
            EBG_HIP(hipMemcpyAsync(m->orig[k][f], field(k, f), (size_t)nvisc(k, f) * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
            if ((rc = mg_device_range(m->orig[k][f], nvisc(k, f), m->partial, s, omin, omax))) return rc;
        }
    // level 0 keeps the original viscosity (update_level0_viscosity_using_original!)
    l0v.etan = m->orig[0][0];
    l0v.etaxy = m->orig[0][1];
    l0v.etaxz = m->orig[0][2];
    l0v.etayz = m->orig[0][3];
    // The restriction's boundary weights extrapolate, so a rough finest-level
    // field can restrict to negative coarse viscosity; the reference's power
    // law then raises a DomainError (negative base, fractional exponent),
    // reported here as an error before that scaling.
    auto scale_all = [&]() -> int {
        const double coef = ramp.coef();
        EBG_CHECK(omin >= 0.0 || coef == std::floor(coef), EBG_ERR_ARG,
                  "ebg_mg3d_run: restricted viscosity is negative on a coarse level (non-integer power)");
        for (int k = 0; k < nl; ++k)
            for (int f = 0; f < 4; ++f)
                mg_scale_viscosity(m->orig[k][f], field(k, f), nvisc(k, f), ramp.min_cur, 1.0 / ramp.min_o, coef, s);
        return EBG_OK;
    };
    if (ramp.below_full() && (rc = scale_all())) return rc;   // rescale_viscosity!
    const auto sm = mg_smoothing_counts(nl, p.finest_sweeps, p.sweep_multiplier, p.sweep_cap);
    const MG3LevelDev<double>& l0 = H.lev[0];   // level 1 of the reference (the finest of the hierarchy)
    const int64_t ny = l0.ynum, nx = l0.xnum, nz = l0.znum;
    const double cnt[4] = {(double)((ny - 1) * (nx - 2) * (nz - 1)), (double)((ny - 2) * (nx - 1) * (nz - 1)),
                           (double)((ny - 1) * (nx - 1) * (nz - 2)), (double)((ny - 1) * (nx - 1) * (nz - 1))};
    double part[4 * MG_RUN_BLOCKS];
    auto mean_residuals = [&](const MG3LevelDev<double>& l, double* res) -> int {
        hipLaunchKernelGGL(mean_residual_partial_kernel, dim3(MG_RUN_BLOCKS), dim3(256), 0, s, l,
                           1.0 / p.stokesscale, 1.0 / p.continscale, m->partial);
        EBG_HIP(hipMemcpyAsync(part, m->partial, sizeof(part), hipMemcpyDeviceToHost, s));
        EBG_HIP(hipStreamSynchronize(s));
        for (int q = 0; q < 4; ++q) {
            double sum = 0.0;
            for (int b = 0; b < MG_RUN_BLOCKS; ++b) sum += part[4 * b + q];
            res[q] = std::log10(std::sqrt(sum / cnt[q]));
        }
        return EBG_OK;
    };
    int jump = 1, overstep = 0, iglobal = 1, done = 0, conv = 0;
    double glast[4] = {0.0, 0.0, 0.0, 0.0};
    for (int ivcycle = 1; ivcycle <= p.nvcycles; ++ivcycle) {
        mg3_vcycle(H, sm, p.theta_s, p.theta_c, p.relax_v, p.relax_p, p.pressure_bc, s);
        double res[4];
        if ((rc = mean_residuals(l0, res))) return rc;
        if (history)
            for (int q = 0; q < 4; ++q) history[4 * (ivcycle - 1) + q] = res[q];
        done = ivcycle;
        if (multimulti) {   // rescale_viscosity_for_levels_using_multimulti! (ViscosityScaling.jl:133-168)
            ++jump;
            if (jump > p.nvcycles_viscosity_jump) {
                jump = 1;
                ramp.max_cur = ramp.max_cur * ramp.max_factor;
                if (ramp.max_cur >= ramp.max_o) ++overstep;
                if (ramp.max_cur > ramp.max_o) ramp.max_cur = ramp.max_o;
                if (overstep > 1) {
                    ramp.max_cur = max_start;
                    overstep = 0;
                    double* a1[4] = {l1.vx, l1.vy, l1.vz, l1.pr};
                    double* a0[4] = {l0v.vx, l0v.vy, l0v.vz, l0v.pr};
                    for (int q = 0; q < 4; ++q)
                        hipLaunchKernelGGL(mm3_add_kernel, dim3(nb3(sizes[q])), dim3(256), 0, s, a0[q],
                                           (const double*)a1[q], sizes[q]);
                    global_residual();
                    if ((rc = mean_residuals(l0v, glast))) return rc;
                    if (global_history && iglobal <= global_capacity)
                        for (int q = 0; q < 4; ++q) global_history[4 * (iglobal - 1) + q] = glast[q];
                    mg3_zero_solution(H, 0, s);
                    ++iglobal;
                }
                if ((rc = scale_all())) return rc;
            }
            double gmax = 1e38;   // calculate_max_global_residual_3d
            if (iglobal - 1 > 1) {
                gmax = 0.0;
                for (int q = 0; q < 4; ++q) {
                    const double v = std::fabs(std::pow(10.0, glast[q]));
                    if (std::isnan(v)) { gmax = NAN; break; }
                    gmax = std::max(gmax, v);
                }
            }
            if (gmax < p.convergence_criterion) {
                conv = 1;
                break;
            }
            continue;
        }
        // rescale_viscosity_for_levels! (3-D): restot sums the x-Stokes
        // residual four times, as the reference does
        ++jump;
        const double rx10 = std::pow(10.0, res[0]);
        const double restot = rx10 + rx10 + rx10 + rx10;
        if ((jump > p.nvcycles_viscosity_jump || restot < p.resmax) && ramp.below_full()) {
            jump = 1;
            ramp.jump();
            if ((rc = scale_all())) return rc;
        }
        // is_converged (non-multimulti)
        if (mg_max_principal_residual(res, 4, ivcycle) < p.convergence_criterion &&
            ivcycle > p.nvcycles_viscosity_jump * p.nviscosity_jumps) {
            conv = 1;
            break;
        }
    }
    if (multimulti) {   // the global solution is level 0's
        double* a1[4] = {l1.vx, l1.vy, l1.vz, l1.pr};
        const double* a0[4] = {l0v.vx, l0v.vy, l0v.vz, l0v.pr};
        for (int q = 0; q < 4; ++q)
            EBG_HIP(hipMemcpyAsync(a1[q], a0[q], sizes[q] * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    EBG_HIP(hipGetLastError());
    if (cycles_done) *cycles_done = done;
    if (converged) *converged = conv;
    if (global_updates) *global_updates = iglobal - 1;
    return EBG_OK;
}

int ebg_mg3d_run(ebg_ctx* ctx, const ebg_mg_run_params* prm, int* cycles_done, int* converged, double* history)
{
    return mg3d_run_impl(ctx, prm, false, cycles_done, converged, history, nullptr, nullptr, 0);
}

int ebg_mg3d_run_multimulti(ebg_ctx* ctx, const ebg_mg_run_params* prm, int* cycles_done, int* converged,
                            double* history, int* global_updates, double* global_history, int global_capacity)
{
    EBG_CHECK(global_capacity >= 0, EBG_ERR_ARG, "ebg_mg3d_run_multimulti: global_capacity < 0");
    return mg3d_run_impl(ctx, prm, true, cycles_done, converged, history, global_updates, global_history,
                         global_capacity);
}

}  // extern "C"
