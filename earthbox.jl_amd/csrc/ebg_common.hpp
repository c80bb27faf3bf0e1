// Shared plumbing for the EarthBox MI355X hot-path library: error capture,
// the per-context stream, and device copies of the staggered-grid axes.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "earthbox_gpu.h"

namespace ebg {

// Thread-local last error text for ebg_last_error(); the library never aborts.
void set_error(const std::string& msg);

#define EBG_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            ::ebg::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return EBG_ERR_HIP;                                              \
        }                                                                    \
    } while (0)

#define EBG_CHECK(cond, code, msg)        \
    do {                                  \
        if (!(cond)) {                    \
            ::ebg::set_error(msg);        \
            return (code);                \
        }                                 \
    } while (0)

// One basic cell [lo, hi) of a grid axis together with everything the marker
// gathers read for a point inside it: the basic step h and its correctly
// rounded reciprocal, and the two staggered nodes that the velocity-grid remap
// chooses between (staggered index c-1 when x <= sb, else c), with their steps
// and reciprocals.  Record k (0-based) describes the 1-based basic cell k+1.
struct alignas(16) AxisCell {
    double lo, hi;
    double h, rh;
    double sa, sha;
    double rsha, sb;
    double shb, rshb;
};

// Device copies of one 2-D staggered grid (all arrays FP64, lengths as in the
// reference: basic gridx_b[xnum], xstp_b[xnum-1], vy-grid gridx_vy[xnum+1],
// xstp_vy[xnum], vx-grid gridy_vx[ynum+1], ystp_vx[ynum]).
struct Grid2D {
    int64_t xnum = 0, ynum = 0;
    double xstpavg = 0, ystpavg = 0;
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
    // uniform-axis hints for O(1) cell search (exactness restored by fix-up)
    double x0 = 0, y0 = 0, inv_hx = 0, inv_hy = 0;
    int uniform_x = 0, uniform_y = 0;
    // every node within 1e-9 of a step of x0 + (i-1)*h: the tolerance-mode
    // RK4 may then use the nominal coordinates for its weights
    int nominal_x = 0, nominal_y = 0;
    double *gridx_b = nullptr, *gridy_b = nullptr;
    double *xstp_b = nullptr, *ystp_b = nullptr;
    double *gridx_vy = nullptr, *xstp_vy = nullptr;
    double *gridy_vx = nullptr, *ystp_vx = nullptr;
    AxisCell *cells_x = nullptr, *cells_y = nullptr;  // xnum-1 and ynum-1 records
    uint64_t generation = 0;  // changes on every ebg_grid2d_set
    std::vector<double> h_gridx_b, h_gridy_b, h_xstp_b, h_ystp_b;
    std::vector<double> h_gridx_vy, h_xstp_vy, h_gridy_vx, h_ystp_vx;
};

}  // namespace ebg

struct ebg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    ebg::Grid2D grid;
    bool has_grid = false;
    // solver state lives in a separate object created on demand
    void* stokes = nullptr;
    void* scatter = nullptr;
    void* mg_ref = nullptr;
    void* cnn = nullptr;  // learned preconditioner correction (ebg_cnn_create)
    void* mg3_ref = nullptr;   // reference-order 3-D multigrid (ebg_mg3d_create)
    void* stokes3d = nullptr;  // 3-D Stokes solve state (ebg_stokes3d_set_system)
    void* plastic = nullptr;   // reduction scratch of ebg_update_nodes_for_plastic_yielding
    void* heat = nullptr;      // heat-equation solve state (ebg_heat.hip)
    void* subgrid = nullptr;   // per-marker buffers of the sub-grid diffusion (ebg_subgrid.hip)
    int rk4_arithmetic = EBG_RK4_EXACT;   // ebg_rk4_set_arithmetic
    int rk4_tiles = 1;                     // ebg_rk4_set_tiling
    int stokes_graphs = 1;                 // ebg_stokes_set_graphs
};

namespace ebg {
int grid_release(Grid2D& g);
void plastic_release(ebg_ctx* ctx);
int plastic_scratch_setup(int device);   // ebg_plastic.hip
void heat_release(ebg_ctx* ctx);         // ebg_heat.hip
void subgrid_release(ebg_ctx* ctx);      // ebg_subgrid.hip
}
