// Learned correction of the multigrid preconditioner: a fully convolutional
// four-level network (3x3 convolutions, 32 channels, GELU, BF16 weights and
// activations with FP32 accumulation on the MFMA units) that maps
// (log viscosity, residual r, V-cycle output z) on the cell grid to a
// correction of z.  Training and inference kernels are in ebg_learned.hip.
// Inputs are scaled by the RMS of all of r and of all of z (s_r, s_z).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

struct ebg_ctx;

namespace ebg {

struct CnnNet;

// Network input channels: log10 viscosity, r (vx, vy, p), z (vx, vy, p).
constexpr int CNN_IN = 7;
constexpr int CNN_OUT = 3;

int cnn_create(CnnNet** out, uint64_t seed);
void cnn_destroy(CnnNet* net);
int64_t cnn_num_params(const CnnNet* net);

// Preconditioner correction on one image of H x W cells (H = xnum-1 x-columns,
// W = ynum-1 rows, pixel index = reference cell index): zc = z + s_z * N(...)
// where N sees (log10 etan, r / s_r, z / s_z) and s_r, s_z are the RMS
// values of r and z over all three components.  r, z, zc are planar FP64 vectors (3 planes of H*W).
int cnn_correct(CnnNet* net, int H, int W, const double* etan, const double* r, const double* z,
                double* zc, hipStream_t s);

// Frees the network owned by a context (ebg_ctx_destroy).
void cnn_release_ctx(ebg_ctx* ctx);

// Training sample from (r, z, e) in planar FP64: in_img [H*W][7] and
// tgt_img [H*W][3] = (e - z) / s_z (FP32, device).
int cnn_pack_sample(CnnNet* net, int H, int W, const double* etan, const double* r, const double* z,
                    const double* e, float* in_img, float* tgt_img, hipStream_t s);

}  // namespace ebg
