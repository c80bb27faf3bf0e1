// FP64 Krylov building blocks shared by the 2-D and 3-D flexible GMRES
// solvers: deterministic multi-dot products (fixed-order block partials, then
// one fixed-order reduction), vector updates, and a growable device buffer.
// The kernels are defined in ebg_stokes.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ebg_common.hpp"

namespace ebg {

constexpr int RED_BLOCKS = 2048;  // partial sums per reduction (8 wavefronts per SIMD: 2.2x the throughput of 512)
constexpr int MAXK = 256;        // longest Krylov basis a multi-dot handles (2-D restart <= MAXK - 2)
// Selective reorthogonalization (2-D and 3-D Krylov solves): a second
// Gram-Schmidt pass when the first removed more than half of the norm (a
// quarter of the squared norm left).  With 1/2 of the squared norm the second
// pass ran at almost every step of the plastic rift solves: same iterations,
// 7 % more solve time (3-D spheres at contrast 1e2-1e6: 5 %).  A looser 1/100
// saved 24 % on the rift but took 119 instead of 80 iterations on the
// 501 x 501 falling block (contrast 1e6) with the FP64 V-cycle, and 117
// instead of 71 on the 101 x 101 block with the FP32 one.
constexpr double REORTH_RATIO2 = 0.25;
constexpr int DOT_CHUNK = 8;     // basis vectors per pass of the multi-dot (register accumulators)

// partial[q*RED_BLOCKS + b] = sum over block b's grid-stride slice of w * V[q]
// for every q < k, DOT_CHUNK basis vectors per pass; a non-null gate holding
// 0 skips the launch's work (device-side condition)
__global__ void multidot_chunks_kernel(const double* __restrict__ w, const double* __restrict__ V, int64_t ldv,
                                       int k, int64_t n, double* __restrict__ partial,
                                       const int* __restrict__ gate);
// out[q] (+)= sum_b partial[q*nblocks + b], one wavefront per q, fixed order;
// gate as above
__global__ void reduce_columns_kernel(const double* __restrict__ partial, int nblocks, int k,
                                      double* __restrict__ out, int accumulate, const int* __restrict__ gate);
// w -= sum_q h[q] V[q]
__global__ void update_kernel(double* __restrict__ w, const double* __restrict__ V, int64_t ldv,
                              const double* __restrict__ h, int k, int64_t n);
// x += sum_q y[q] Z[q]
__global__ void combine_kernel(double* __restrict__ x, const double* __restrict__ Z, int64_t ldz,
                               const double* __restrict__ y, int k, int64_t n);
// out = in / nrm[0]
__global__ void scale_kernel(double* __restrict__ out, const double* __restrict__ in, const double* nrm,
                             int64_t n);

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int ensure(size_t b)
    {
        if (bytes >= b && p) return EBG_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        EBG_HIP(hipMalloc(&p, b ? b : 8));
        bytes = b;
        return EBG_OK;
    }
    template <typename T> T* as() { return (T*)p; }
};

}  // namespace ebg
