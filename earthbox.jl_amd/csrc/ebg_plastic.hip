// Nodal plastic yielding: one step of the reference's Picard loop over nodes.
//
// Replaces PlasticFailureNodes.update_nodes_for_plastic_yielding
// (rheology/plastic_failure/nodes/PlasticFailureNodes.jl:24-319) with
// CheckYield.check_yield_is_applicable, YieldStress.calc_yield_stress,
// FluidPressure.calculate_fluid_pressure (plastic_failure/utils/) and
// MobileWall.in_mobile_wall_transient / in_plate_extension_transient
// (utils/MobileWall.jl:106-122, 279-329).  One thread per basic node; the
// yield count and the sum of squared yield errors are reduced in a fixed
// order (per-block partials, then one block), so repeated calls agree.
#include "ebg_common.hpp"
#include "ebg_yield.hpp"

#include <algorithm>
#include <mutex>

#include <cmath>

namespace ebg {

namespace {

constexpr int PL_THREADS = 256;
constexpr int PL_MAX_BLOCKS = 1024;

struct PlArgs {
    int xnum, ynum;
    const double *gridx_b, *gridy_b;
    ebg_plastic_nodes_in in;
    double *etavp_new, *plastic_strain_rate, *plastic_yield_new, *yield_error;
    double* partial_err;       // [nblocks]
    long long* partial_count;  // [nblocks]
};

// PlasticFailureNodes.jl:206-248 (interpolate_from_pressure_grid_to_basic_grid)
__device__ double pressure_to_basic(int i, int j, int ynum, int xnum, const double* __restrict__ a)
{
    const int ld = ynum - 1;
#define PA(ii, jj) a[(size_t)((jj) - 1) * ld + ((ii) - 1)]
    if (1 < i && i < ynum && 1 < j && j < xnum) return (PA(i - 1, j - 1) + PA(i, j - 1) + PA(i - 1, j) + PA(i, j)) / 4.0;
    if (i == 1 && j == 1) return PA(i, j);
    if (i == ynum && j == 1) return PA(i - 1, j);
    if (i == ynum && j == xnum) return PA(i - 1, j - 1);
    if (i == 1 && j == xnum) return PA(i, j - 1);
    if (j == 1) return (PA(i, j) + PA(i - 1, j)) / 2.0;
    if (j == xnum) return (PA(i - 1, j - 1) + PA(i, j - 1)) / 2.0;
    if (i == 1) return (PA(i, j) + PA(i, j - 1)) / 2.0;
    if (i == ynum) return (PA(i - 1, j) + PA(i - 1, j - 1)) / 2.0;
#undef PA
    return 0.0;
}

__global__ void __launch_bounds__(PL_THREADS) plastic_nodes_kernel(PlArgs a)
{
    __shared__ double serr[PL_THREADS];
    __shared__ long long scnt[PL_THREADS];
    const int ny = a.ynum, nx = a.xnum;
    const int64_t nn = (int64_t)nx * ny;
    const ebg_plastic_nodes_in& in = a.in;
    const double dt = in.timestep;
    const double pi_term = M_PI / 180.0;
    double err = 0.0;
    long long cnt = 0;
    for (int64_t t = (int64_t)blockIdx.x * PL_THREADS + threadIdx.x; t < nn; t += (int64_t)gridDim.x * PL_THREADS) {
        const int i = (int)(t % ny) + 1, j = (int)(t / ny) + 1;
        const double yield_marker = in.plastic_yield[t];
        const double eta_old = in.etas1[t];
        const double eta_flow = in.eta_flow[t];
        const double mu = in.mus1[t];
        const double cohesion = in.cohesion_grid[t];
        const double fric = sin(in.fric_degrees_grid[t] * pi_term);
        bool applicable = in.ntimestep > 0 && (cohesion > 0.0 || fric > 0.0);
        const double xb = a.gridx_b[j - 1], yb = a.gridy_b[i - 1];
        if (in.no_yielding_in_mobile_wall) {
            const double xl = in.x_left_mobile_wall + in.velocity_internal_x * in.timesum;
            const double xr = in.x_right_mobile_wall + in.velocity_internal_x * in.timesum;
            if (xl <= xb && xb <= xr && in.y_top_mobile_wall <= yb && yb <= in.y_bottom_mobile_wall)
                applicable = false;
        }
        if (in.no_yielding_in_plate_extension) {
            const double xl = in.x_left_mobile_wall + in.velocity_internal_x * in.timesum;
            const double xmin = xl - in.plate_extension_width, xmax = xl;
            const double ymin = in.y_bottom_mobile_wall - in.plate_extension_thickness, ymax = in.y_bottom_mobile_wall;
            if (xmin <= xb && xb <= xmax && ymin <= yb && yb <= ymax) applicable = false;
        }
        double ev = eta_flow, py = 0.0, psr = 0.0;
        if (applicable) {
            const double pressure = pressure_to_basic(i, j, ny, nx, in.pr1);
            const double sxx = pressure_to_basic(i, j, ny, nx, in.sxx2);
            const double sxy = in.sxy2[t];
            const double sii = sqrt(sxy * sxy + sxx * sxx);
            const double sii_e = sii * (mu * dt + eta_old) / eta_old;
            const double sy = yield_stress(cohesion, fric, pressure, in.iuse_fluid_pressure_for_yield, yb, in.y_sealevel);
            const bool at_yield = yield_marker > 0;
            bool counted = false;
            if (at_yield) counted = true;
            if (sy < sii_e) {
                double eta_new = dt * mu * sy / (sii_e - sy);
                if (eta_new < eta_flow) {
                    py = 1.0;
                    eta_new = pow(eta_new, 1.0 - 0.0) * pow(eta_old, 0.0);
                    eta_new = jl_max(eta_new, in.viscosity_min);
                    eta_new = jl_min(eta_new, in.viscosity_max);
                    ev = eta_new;
                    psr = 0.5 * sii / eta_new - 0.5 * sii / eta_flow;
                    counted = true;
                }
            }
            if (counted) {
                const double e = sii - sy;
                a.yield_error[t] = e;
                err += e * e;
                ++cnt;
            }
        }
        a.etavp_new[t] = ev;
        a.plastic_yield_new[t] = py;
        a.plastic_strain_rate[t] = psr;
    }
    serr[threadIdx.x] = err;
    scnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = PL_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            serr[threadIdx.x] += serr[threadIdx.x + s];
            scnt[threadIdx.x] += scnt[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.partial_err[blockIdx.x] = serr[0];
        a.partial_count[blockIdx.x] = scnt[0];
    }
}

// one block: the partials in block order, then sqrt(sum / count)
__global__ void __launch_bounds__(PL_THREADS) plastic_finish_kernel(const double* __restrict__ perr,
                                                                    const long long* __restrict__ pcnt, int nb,
                                                                    double* __restrict__ result)
{
    __shared__ double serr[PL_THREADS];
    __shared__ long long scnt[PL_THREADS];
    double e = 0.0;
    long long c = 0;
    for (int k = threadIdx.x; k < nb; k += PL_THREADS) {
        e += perr[k];
        c += pcnt[k];
    }
    serr[threadIdx.x] = e;
    scnt[threadIdx.x] = c;
    __syncthreads();
    for (int s = PL_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            serr[threadIdx.x] += serr[threadIdx.x + s];
            scnt[threadIdx.x] += scnt[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        result[0] = (double)scnt[0];
        result[1] = scnt[0] > 0 ? sqrt(serr[0] / (double)scnt[0]) : 0.0;
    }
}

}  // namespace

// HarmonicAverage.normal_viscosity_from_harmonic_avg_shear_viscosity! with
// CheckPlasticYield.plastic_yielding (tolerance 0) and MathTools.harmonic4
// (factors 0, summed in the reference's order)
__global__ void harmonic_normal_viscosity_kernel(int xnum, int ynum, const double* __restrict__ etas1,
                                                 const double* __restrict__ plastics,
                                                 const double* __restrict__ plasticn,
                                                 const double* __restrict__ pyield, double* __restrict__ etan1)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = ynum;
    if (c >= (int64_t)(xnum - 1) * (ynum - 1)) return;
    const int i = (int)(c % (ny - 1)) + 1, j = (int)(c / (ny - 1)) + 1;
#define BA(a, ii, jj) (a)[(size_t)((jj) - 1) * ny + ((ii) - 1)]
    bool y = plasticn[c] > 0.0;
    if (BA(plastics, i, j) > 0.0 || BA(plastics, i, j + 1) > 0.0 || BA(plastics, i + 1, j) > 0.0 ||
        BA(plastics, i + 1, j + 1) > 0.0)
        y = true;
    if (BA(pyield, i, j) > 0.0 || BA(pyield, i, j + 1) > 0.0 || BA(pyield, i + 1, j) > 0.0 ||
        BA(pyield, i + 1, j + 1) > 0.0)
        y = true;
    if (!y) return;
    const double vinv = 1.0 / BA(etas1, i, j) + 1.0 / BA(etas1, i, j + 1) + 1.0 / BA(etas1, i + 1, j) +
                        1.0 / BA(etas1, i + 1, j + 1) + 0.0 + 0.0 + 0.0 + 0.0;
    const double den = 4.0 + 0.0 + 0.0 + 0.0 + 0.0;
    if (vinv != 0.0 && den != 0.0) etan1[c] = 1.0 / (vinv / den);
#undef BA
}

// PlasticStrainRateInterpolation.interpolate_at_pressure_nodes!
__global__ void plastic_rate_pressure_kernel(int xnum, int ynum, const double* __restrict__ eb,
                                             double* __restrict__ ep)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = ynum;
    if (c >= (int64_t)(xnum - 1) * (ynum - 1)) return;
    const int i = (int)(c % (ny - 1)) + 1, j = (int)(c / (ny - 1)) + 1;
#define BA(ii, jj) eb[(size_t)((jj) - 1) * ny + ((ii) - 1)]
    ep[c] = (BA(i, j) + BA(i, j + 1) + BA(i + 1, j) + BA(i + 1, j + 1)) / 4.0;
#undef BA
}

// per-context scratch: nb partial sums, nb partial counts, two results
struct PlasticScratch {
    double* buf = nullptr;
    ~PlasticScratch()
    {
        if (buf) (void)hipFree(buf);
    }
};

void plastic_release(ebg_ctx* ctx)
{
    if (ctx && ctx->plastic) {
        delete (PlasticScratch*)ctx->plastic;
        ctx->plastic = nullptr;
    }
}

}  // namespace ebg

using namespace ebg;

namespace ebg {

// Scratch of the allocation-free variant (ebg_update_nodes_for_plastic_yielding_noalloc):
// PL_MAX_BLOCKS partial sums of squared yield errors, as many partial counts
// (int64 in double-sized slots) and the two results (count, RMS): 2,050
// doubles, 16.4 KB per device, held until the process ends.  The node loop is
// grid-strided over at most PL_MAX_BLOCKS blocks, so this size covers any grid.
constexpr int PL_SCRATCH_DOUBLES = 2 * PL_MAX_BLOCKS + 2;
static __device__ __align__(256) double scratch[PL_SCRATCH_DOUBLES];

// Device address of `scratch` per device, taken once per device with that
// device current (each device holds its own copy of the array at its own
// address; in host code the array's name is not a device address).
constexpr int PL_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[PL_MAX_DEVICES] = {};

int plastic_scratch_setup(int device)
{
    EBG_CHECK(device >= 0 && device < PL_MAX_DEVICES, EBG_ERR_ARG, "plastic scratch: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return EBG_OK;
    void* p = nullptr;
    EBG_HIP(hipSetDevice(device));
    EBG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double*>(p);
    return EBG_OK;
}

static double* plastic_scratch_ptr(int device)
{
    if (device < 0 || device >= PL_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

// The node pass and the fixed-order reduction on buf (PL_SCRATCH_DOUBLES
// doubles: partial sums, partial counts, results), then the count and RMS back
// to the host.
static int plastic_yielding_run(ebg_ctx* ctx, const char* name, const ebg_plastic_nodes_in* in,
                                ebg_plastic_nodes_out* out, double* buf)
{
    const Grid2D& g = ctx->grid;
    hipStream_t s = ctx->stream;
    const int64_t nn = g.xnum * g.ynum;
    const int nb = (int)std::min<int64_t>(PL_MAX_BLOCKS, (nn + PL_THREADS - 1) / PL_THREADS);
    EBG_CHECK(nb >= 1 && 2 * nb + 2 <= PL_SCRATCH_DOUBLES, EBG_ERR_ARG,
              std::string(name) + ": reduction does not fit the scratch");
    PlArgs a{};
    a.xnum = (int)g.xnum;
    a.ynum = (int)g.ynum;
    a.gridx_b = g.gridx_b;
    a.gridy_b = g.gridy_b;
    a.in = *in;
    a.etavp_new = out->etavp_new;
    a.plastic_strain_rate = out->plastic_strain_rate;
    a.plastic_yield_new = out->plastic_yield_new;
    a.yield_error = out->yield_error;
    a.partial_err = buf;
    a.partial_count = (long long*)(buf + PL_MAX_BLOCKS);
    double* result = buf + 2 * PL_MAX_BLOCKS;
    hipLaunchKernelGGL(plastic_nodes_kernel, dim3(nb), dim3(PL_THREADS), 0, s, a);
    hipLaunchKernelGGL(plastic_finish_kernel, dim3(1), dim3(PL_THREADS), 0, s, (const double*)a.partial_err,
                       (const long long*)a.partial_count, nb, result);
    EBG_HIP(hipGetLastError());
    double h[2];
    EBG_HIP(hipMemcpyAsync(h, result, sizeof(h), hipMemcpyDeviceToHost, s));
    EBG_HIP(hipStreamSynchronize(s));
    out->nnode_yield = (int64_t)h[0];
    out->global_yield_error = h[1];
    return EBG_OK;
}

static int plastic_check_args(ebg_ctx* ctx, const char* name, const ebg_plastic_nodes_in* in,
                              const ebg_plastic_nodes_out* out)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, std::string(name) + ": no grid");
    EBG_CHECK(in && out, EBG_ERR_ARG, std::string(name) + ": NULL arguments");
    EBG_CHECK(in->plastic_yield && in->etas1 && in->eta_flow && in->mus1 && in->cohesion_grid &&
                  in->fric_degrees_grid && in->sxy2 && in->sxx2 && in->pr1,
              EBG_ERR_ARG, std::string(name) + ": NULL input field");
    EBG_CHECK(out->etavp_new && out->plastic_strain_rate && out->plastic_yield_new && out->yield_error, EBG_ERR_ARG,
              std::string(name) + ": NULL output field");
    return EBG_OK;
}

}  // namespace ebg

extern "C" int ebg_update_nodes_for_plastic_yielding(ebg_ctx* ctx, const ebg_plastic_nodes_in* in,
                                                     ebg_plastic_nodes_out* out)
{
    static const char* name = "ebg_update_nodes_for_plastic_yielding";
    int rc = plastic_check_args(ctx, name, in, out);
    if (rc) return rc;
    if (!ctx->plastic) {
        PlasticScratch* ps = new PlasticScratch();
        if (hipMalloc(&ps->buf, sizeof(double) * PL_SCRATCH_DOUBLES) != hipSuccess) {
            delete ps;
            set_error("ebg_update_nodes_for_plastic_yielding: hipMalloc failed");
            return EBG_ERR_HIP;
        }
        ctx->plastic = ps;
    }
    return plastic_yielding_run(ctx, name, in, out, ((PlasticScratch*)ctx->plastic)->buf);
}

extern "C" int ebg_update_nodes_for_plastic_yielding_noalloc(ebg_ctx* ctx, const ebg_plastic_nodes_in* in,
                                                             ebg_plastic_nodes_out* out)
{
    static const char* name = "ebg_update_nodes_for_plastic_yielding_noalloc";
    int rc = plastic_check_args(ctx, name, in, out);
    if (rc) return rc;
    double* buf = plastic_scratch_ptr(ctx->device);
    EBG_CHECK(buf != nullptr, EBG_ERR_STATE,
              std::string(name) + ": scratch address not set up for this device (ebg_ctx_create failed?)");
    return plastic_yielding_run(ctx, name, in, out, buf);
}

extern "C" int ebg_harmonic_normal_viscosity(ebg_ctx* ctx, const double* etas1, const double* plastics,
                                             const double* plasticn, const double* plastic_yield, double* etan1)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_harmonic_normal_viscosity: no grid");
    EBG_CHECK(etas1 && plastics && plasticn && plastic_yield && etan1, EBG_ERR_ARG,
              "ebg_harmonic_normal_viscosity: NULL field");
    const Grid2D& g = ctx->grid;
    const int64_t np = (g.xnum - 1) * (g.ynum - 1);
    hipLaunchKernelGGL(harmonic_normal_viscosity_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream,
                       (int)g.xnum, (int)g.ynum, etas1, plastics, plasticn, plastic_yield, etan1);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

extern "C" int ebg_plastic_strain_rate_at_pressure_nodes(ebg_ctx* ctx, const double* eii_plastic_basic,
                                                         double* eii_plastic_pressure)
{
    EBG_CHECK(ctx && ctx->has_grid, EBG_ERR_STATE, "ebg_plastic_strain_rate_at_pressure_nodes: no grid");
    EBG_CHECK(eii_plastic_basic && eii_plastic_pressure, EBG_ERR_ARG,
              "ebg_plastic_strain_rate_at_pressure_nodes: NULL field");
    const Grid2D& g = ctx->grid;
    const int64_t np = (g.xnum - 1) * (g.ynum - 1);
    hipLaunchKernelGGL(plastic_rate_pressure_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream,
                       (int)g.xnum, (int)g.ynum, eii_plastic_basic, eii_plastic_pressure);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}
