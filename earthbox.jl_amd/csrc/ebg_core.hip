// Context, memory helpers, error text and grid upload for the C ABI.
#include "ebg_common.hpp"
#include "ebg_learned.hpp"

#include <cmath>
#include <cstring>

namespace ebg {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int grid_release(Grid2D& g)
{
    double** ptrs[] = {&g.gridx_b, &g.gridy_b, &g.xstp_b, &g.ystp_b,
                       &g.gridx_vy, &g.xstp_vy, &g.gridy_vx, &g.ystp_vx};
    for (double** p : ptrs) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    AxisCell** cells[] = {&g.cells_x, &g.cells_y};
    for (AxisCell** p : cells) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    return EBG_OK;
}

// Per-cell axis records for the marker gathers.  The reciprocals are IEEE
// divisions on the host, i.e. correctly rounded, which the device division
// rn_div() requires.  stag/stag_stp are the staggered axis the velocity remap
// chooses from (gridx_vy/xstp_vy for x, gridy_vx/ystp_vx for y).
static int upload_cells(AxisCell** dst, int64_t n, const double* g, const double* stp,
                        const double* stag, const double* stag_stp)
{
    std::vector<AxisCell> h((size_t)(n - 1));
    for (int64_t k = 0; k + 1 < n; ++k) {
        AxisCell& c = h[(size_t)k];
        c.lo = g[k];
        c.hi = g[k + 1];
        c.h = stp[k];
        c.rh = 1.0 / stp[k];
        c.sa = stag[k];
        c.sha = stag_stp[k];
        c.rsha = 1.0 / stag_stp[k];
        c.sb = stag[k + 1];
        c.shb = stag_stp[k + 1];
        c.rshb = 1.0 / stag_stp[k + 1];
    }
    EBG_HIP(hipMalloc(dst, sizeof(AxisCell) * h.size()));
    EBG_HIP(hipMemcpy(*dst, h.data(), sizeof(AxisCell) * h.size(), hipMemcpyHostToDevice));
    return EBG_OK;
}

static int upload(double** dst, const double* src, int64_t n, std::vector<double>& host)
{
    host.assign(src, src + n);
    EBG_HIP(hipMalloc(dst, sizeof(double) * (size_t)n));
    EBG_HIP(hipMemcpy(*dst, src, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
    return EBG_OK;
}

// An axis is "uniform" when every node lies within a quarter cell of
// x0 + (i-1)*h; the O(1) cell guess is then off by at most one and the exact
// comparison fix-up restores searchsortedlast semantics.
static int axis_uniform(const double* g, int64_t n, double* x0, double* inv_h)
{
    double h = (g[n - 1] - g[0]) / (double)(n - 1);
    *x0 = g[0];
    *inv_h = 1.0 / h;
    if (!(h > 0)) return 0;
    for (int64_t i = 0; i < n; ++i)
        if (std::fabs(g[i] - (g[0] + (double)i * h)) > 0.25 * h) return 0;
    return 1;
}

// Basic nodes within 1e-9 of a step of x0 + (i-1)*h and staggered nodes
// within 1e-9 of a step of x0 + (i-3/2)*h (including both ghost nodes): the
// tolerance-mode RK4 then computes cells and weights from these nominal
// positions, which moves a weight by at most 1e-9 of the cell.
static int axis_nominal(const double* g, int64_t n, const double* stag)
{
    const double h = (g[n - 1] - g[0]) / (double)(n - 1);
    if (!(h > 0)) return 0;
    const double tol = 1e-9 * h;
    for (int64_t i = 0; i < n; ++i)
        if (!(std::fabs(g[i] - (g[0] + (double)i * h)) <= tol)) return 0;
    for (int64_t i = 0; i <= n; ++i)
        if (!(std::fabs(stag[i] - (g[0] + ((double)i - 0.5) * h)) <= tol)) return 0;
    return 1;
}

int mg3ref_release(ebg_ctx* ctx);    // ebg_mg3d.hip
int stokes3d_release(ebg_ctx* ctx);  // ebg_stokes3d.hip

}  // namespace ebg

using namespace ebg;

extern "C" {

const char* ebg_last_error(void) { return g_last_error.c_str(); }

int ebg_ctx_create(ebg_ctx** out, int device)
{
    EBG_CHECK(out != nullptr, EBG_ERR_ARG, "ebg_ctx_create: out is NULL");
    int ndev = 0;
    EBG_HIP(hipGetDeviceCount(&ndev));
    EBG_CHECK(device >= 0 && device < ndev, EBG_ERR_ARG, "ebg_ctx_create: bad device index");
    EBG_HIP(hipSetDevice(device));
    ebg_ctx* c = new ebg_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        return EBG_ERR_HIP;
    }
    c->own_stream = true;
    // device address of the allocation-free entry point's scratch, once per device
    int rc = plastic_scratch_setup(device);
    if (rc != EBG_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return EBG_OK;
}

int ebg_stokes_release(ebg_ctx* ctx);
int ebg_scatter_release(ebg_ctx* ctx);
int ebg_mgref_release(ebg_ctx* ctx);

int ebg_ctx_destroy(ebg_ctx* ctx)
{
    if (!ctx) return EBG_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ebg_stokes_release(ctx);
    ebg_scatter_release(ctx);
    ebg_mgref_release(ctx);
    mg3ref_release(ctx);
    stokes3d_release(ctx);
    cnn_release_ctx(ctx);
    plastic_release(ctx);
    heat_release(ctx);
    subgrid_release(ctx);
    grid_release(ctx->grid);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return EBG_OK;
}

int ebg_ctx_set_stream(ebg_ctx* ctx, void* hip_stream)
{
    EBG_CHECK(ctx != nullptr, EBG_ERR_ARG, "ebg_ctx_set_stream: ctx is NULL");
    // NULL selects the default (null) stream, the one an external runtime such
    // as torch uses unless told otherwise
    if (ctx->own_stream && ctx->stream) {
        EBG_HIP(hipStreamSynchronize(ctx->stream));
        EBG_HIP(hipStreamDestroy(ctx->stream));
    }
    ctx->stream = (hipStream_t)hip_stream;
    ctx->own_stream = false;
    return EBG_OK;
}

int ebg_ctx_synchronize(ebg_ctx* ctx)
{
    EBG_CHECK(ctx != nullptr, EBG_ERR_ARG, "ebg_ctx_synchronize: ctx is NULL");
    EBG_HIP(hipStreamSynchronize(ctx->stream));
    return EBG_OK;
}

int ebg_device_alloc(void** ptr, size_t bytes)
{
    EBG_CHECK(ptr != nullptr, EBG_ERR_ARG, "ebg_device_alloc: ptr is NULL");
    EBG_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return EBG_OK;
}

int ebg_device_free(void* ptr)
{
    if (ptr) EBG_HIP(hipFree(ptr));
    return EBG_OK;
}

int ebg_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    EBG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return EBG_OK;
}

int ebg_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    EBG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return EBG_OK;
}

int ebg_grid2d_set(ebg_ctx* ctx, int64_t xnum, int64_t ynum,
                   const double* gridx_b, const double* gridy_b,
                   const double* xstp_b, const double* ystp_b,
                   const double* gridx_vy, const double* xstp_vy,
                   const double* gridy_vx, const double* ystp_vx,
                   double xstpavg, double ystpavg)
{
    EBG_CHECK(ctx != nullptr, EBG_ERR_ARG, "ebg_grid2d_set: ctx is NULL");
    EBG_CHECK(xnum >= 3 && ynum >= 3, EBG_ERR_ARG, "ebg_grid2d_set: need xnum, ynum >= 3");
    EBG_CHECK(xnum < (1 << 30) && ynum < (1 << 30), EBG_ERR_ARG, "ebg_grid2d_set: grid too large");
    EBG_CHECK(gridx_b && gridy_b && xstp_b && ystp_b && gridx_vy && xstp_vy && gridy_vx && ystp_vx,
              EBG_ERR_ARG, "ebg_grid2d_set: NULL grid array");
    for (int64_t i = 1; i < xnum; ++i)
        EBG_CHECK(gridx_b[i] > gridx_b[i - 1], EBG_ERR_ARG, "ebg_grid2d_set: gridx_b not increasing");
    for (int64_t i = 1; i < ynum; ++i)
        EBG_CHECK(gridy_b[i] > gridy_b[i - 1], EBG_ERR_ARG, "ebg_grid2d_set: gridy_b not increasing");
    // the marker kernels' exact division needs every step in [2^-100, 2^100]
    auto steps_ok = [](const double* s, int64_t n) {
        for (int64_t i = 0; i < n; ++i)
            if (!(s[i] >= 0x1p-100 && s[i] <= 0x1p100)) return false;
        return true;
    };
    EBG_CHECK(steps_ok(xstp_b, xnum - 1) && steps_ok(ystp_b, ynum - 1) && steps_ok(xstp_vy, xnum) &&
                  steps_ok(ystp_vx, ynum),
              EBG_ERR_ARG, "ebg_grid2d_set: grid steps must lie in [2^-100, 2^100]");
    EBG_HIP(hipSetDevice(ctx->device));
    EBG_HIP(hipStreamSynchronize(ctx->stream));
    Grid2D& g = ctx->grid;
    grid_release(g);
    g.xnum = xnum;
    g.ynum = ynum;
    g.xstpavg = xstpavg;
    g.ystpavg = ystpavg;
    int rc;
    if ((rc = upload(&g.gridx_b, gridx_b, xnum, g.h_gridx_b))) return rc;
    if ((rc = upload(&g.gridy_b, gridy_b, ynum, g.h_gridy_b))) return rc;
    if ((rc = upload(&g.xstp_b, xstp_b, xnum - 1, g.h_xstp_b))) return rc;
    if ((rc = upload(&g.ystp_b, ystp_b, ynum - 1, g.h_ystp_b))) return rc;
    if ((rc = upload(&g.gridx_vy, gridx_vy, xnum + 1, g.h_gridx_vy))) return rc;
    if ((rc = upload(&g.xstp_vy, xstp_vy, xnum, g.h_xstp_vy))) return rc;
    if ((rc = upload(&g.gridy_vx, gridy_vx, ynum + 1, g.h_gridy_vx))) return rc;
    if ((rc = upload(&g.ystp_vx, ystp_vx, ynum, g.h_ystp_vx))) return rc;
    if ((rc = upload_cells(&g.cells_x, xnum, gridx_b, xstp_b, gridx_vy, xstp_vy))) return rc;
    if ((rc = upload_cells(&g.cells_y, ynum, gridy_b, ystp_b, gridy_vx, ystp_vx))) return rc;
    g.xmin = gridx_b[0];
    g.xmax = gridx_b[xnum - 1];
    g.ymin = gridy_b[0];
    g.ymax = gridy_b[ynum - 1];
    g.uniform_x = axis_uniform(gridx_b, xnum, &g.x0, &g.inv_hx);
    g.uniform_y = axis_uniform(gridy_b, ynum, &g.y0, &g.inv_hy);
    g.nominal_x = g.uniform_x && axis_nominal(gridx_b, xnum, gridx_vy);
    g.nominal_y = g.uniform_y && axis_nominal(gridy_b, ynum, gridy_vx);
    static uint64_t generations = 0;
    g.generation = ++generations;
    ctx->has_grid = true;
    return EBG_OK;
}

}  // extern "C"
