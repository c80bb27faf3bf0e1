// Learned correction network: BF16 MFMA convolutions with FP32 accumulation,
// GELU, 2x2 average pooling, nearest upsampling with skip additions, MSE loss
// and AdamW, forward and backward, all as HIP kernels.
//
// Architecture (fully convolutional, four levels, 32 channels, 3x3 kernels):
//   level 0: in(7->32) a0            -> x0      pool
//   level 1: e1a e1b                 -> x1      pool
//   level 2: e2a e2b                 -> x2      pool
//   level 3: e3a e3b                 -> x3
//   level 2: d2a(up(x3)+x2) d2b      -> y2
//   level 1: d1a(up(y2)+x1) d1b      -> y1
//   level 0: d0a(up(y1)+x0) d0b      -> y0,  out(32->3), no activation
// Every convolution but the last is followed by GELU (tanh form).  Tensors are
// channel-last [pixel][channel] with channels padded to a multiple of 16; the
// pixel index of an image with H rows of W pixels is h*W + w, batch-major.
//
// Convolution = implicit GEMM on v_mfma_f32_16x16x16bf16_1k: output channels
// are the M dimension, pixels the N dimension and (tap, input channel) the K
// dimension.  Operand layout of that instruction (wave64): lane l supplies
// A[l%16][4*(l/16) .. +3] and B[4*(l/16) .. +3][l%16] and receives
// D[4*(l/16) .. +3][l%16].
#include "ebg_common.hpp"
#include "ebg_learned.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace ebg {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- bf16
__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline v4s load4bf(const uint16_t* p)  // 8-byte aligned
{
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    v4s r;
    r[0] = (short)(u.x & 0xffff);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffff);
    r[3] = (short)(u.y >> 16);
    return r;
}

// ---------------------------------------------------------------- conv
// y[p][co] = bias[co] + sum_{tap, ci} w[tap][co][ci] * x[p + off(tap)][ci],
// off(tap) = (tap/3 - 1, tap%3 - 1) in (h, w); out-of-image pixels read 0.
// One wave computes 64 consecutive pixels x COUT channels.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) conv3x3_fwd_kernel(const uint16_t* __restrict__ x, int B, int H,
                                                          int W, const uint16_t* __restrict__ wt,
                                                          const float* __restrict__ bias,
                                                          float* __restrict__ y)
{
    constexpr int MB = COUT / 16, NB = 4, KC = CIN / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t P = (int64_t)B * H * W;
    const int64_t p0 = ((int64_t)blockIdx.x * 4 + wave) * 64;
    if (p0 >= P) return;
    const int col = lane & 15, kq = lane >> 4;
    int pb[NB], ph[NB], pw[NB];
    bool pv[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int64_t p = p0 + nb * 16 + col;
        pv[nb] = p < P;
        const int64_t q = pv[nb] ? p : 0;
        pw[nb] = (int)(q % W);
        ph[nb] = (int)((q / W) % H);
        pb[nb] = (int)(q / ((int64_t)W * H));
    }
    v4f acc[MB][NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
            v4s bop[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int hh = ph[nb] + dy, ww = pw[nb] + dx;
                if (pv[nb] && hh >= 0 && hh < H && ww >= 0 && ww < W)
                    bop[nb] = load4bf(x + (((int64_t)pb[nb] * H + hh) * W + ww) * CIN + kc * 16 + kq * 4);
                else
                    bop[nb] = v4s{0, 0, 0, 0};
            }
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const v4s aop = load4bf(wt + ((size_t)(tap * COUT + mb * 16 + col)) * CIN + kc * 16 + kq * 4);
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aop, bop[nb], acc[mb][nb], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            if (!pv[nb]) continue;
            const int co = mb * 16 + kq * 4;
            const int64_t p = p0 + nb * 16 + col;
            float4 o;
            o.x = acc[mb][nb][0] + (bias ? bias[co + 0] : 0.f);
            o.y = acc[mb][nb][1] + (bias ? bias[co + 1] : 0.f);
            o.z = acc[mb][nb][2] + (bias ? bias[co + 2] : 0.f);
            o.w = acc[mb][nb][3] + (bias ? bias[co + 3] : 0.f);
            *reinterpret_cast<float4*>(y + p * COUT + co) = o;
        }
}

// The same for CIN = 32 on gfx950's double-K instruction
// v_mfma_f32_16x16x32_bf16: one instruction per (tap, 16 output channels,
// 16 pixels) covers all 32 input channels.  Operand layout (wave64): lane l
// supplies A[l%16][8*(l/16) .. +7] (the weights w[tap][co][ci], 16 contiguous
// bytes) and B[8*(l/16) .. +7][l%16] (x[pixel][ci], 16 contiguous bytes) and
// receives D[4*(l/16) .. +3][l%16].  The block's weights (9 x COUT x 32 bf16,
// 18 KB at COUT = 32) are staged once in LDS and shared by its four waves.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <int COUT>
__global__ void __launch_bounds__(256) conv3x3_fwd32_kernel(const uint16_t* __restrict__ x, int B, int H, int W,
                                                           const uint16_t* __restrict__ wt,
                                                           const float* __restrict__ bias, float* __restrict__ y)
{
    constexpr int CIN = 32, MB = COUT / 16, NB = 4, WV = 9 * COUT * CIN / 8;   // 16-byte units
    __shared__ uint4 wsh[WV];
    const uint4* wg = reinterpret_cast<const uint4*>(wt);
    for (int i = threadIdx.x; i < WV; i += 256) wsh[i] = wg[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t P = (int64_t)B * H * W;
    const int64_t p0 = ((int64_t)blockIdx.x * 4 + wave) * 64;
    if (p0 >= P) return;
    const int col = lane & 15, kq = lane >> 4;
    int pb[NB], ph[NB], pw[NB];
    bool pv[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int64_t p = p0 + nb * 16 + col;
        pv[nb] = p < P;
        const int64_t q = pv[nb] ? p : 0;
        pw[nb] = (int)(q % W);
        ph[nb] = (int)((q / W) % H);
        pb[nb] = (int)(q / ((int64_t)W * H));
    }
    v4f acc[MB][NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        v8bf bop[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int hh = ph[nb] + dy, ww = pw[nb] + dx;
            uint4 u = make_uint4(0u, 0u, 0u, 0u);
            if (pv[nb] && hh >= 0 && hh < H && ww >= 0 && ww < W)
                u = *reinterpret_cast<const uint4*>(x + (((int64_t)pb[nb] * H + hh) * W + ww) * CIN + kq * 8);
            bop[nb] = __builtin_bit_cast(v8bf, u);
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const v8bf aop = __builtin_bit_cast(v8bf, wsh[(tap * COUT + mb * 16 + col) * (CIN / 8) + kq]);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aop, bop[nb], acc[mb][nb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            if (!pv[nb]) continue;
            const int co = mb * 16 + kq * 4;
            const int64_t p = p0 + nb * 16 + col;
            float4 o;
            o.x = acc[mb][nb][0] + (bias ? bias[co + 0] : 0.f);
            o.y = acc[mb][nb][1] + (bias ? bias[co + 1] : 0.f);
            o.z = acc[mb][nb][2] + (bias ? bias[co + 2] : 0.f);
            o.w = acc[mb][nb][3] + (bias ? bias[co + 3] : 0.f);
            *reinterpret_cast<float4*>(y + p * COUT + co) = o;
        }
}

// Weight gradient partial sums: part[chunk][tap][co][ci] =
// sum over the chunk's pixels p of dy[p][co] * x[p + off(tap)][ci].
// One wave per (chunk, tap); K = pixels in groups of 16.
template <int CIN, int COUT>
__global__ void __launch_bounds__(64) conv3x3_wgrad_kernel(const uint16_t* __restrict__ x,
                                                           const uint16_t* __restrict__ dy, int B,
                                                           int H, int W, int chunk,
                                                           float* __restrict__ part)
{
    constexpr int MB = COUT / 16, NB = CIN / 16;
    const int lane = threadIdx.x, col = lane & 15, kq = lane >> 4;
    const int c = blockIdx.x, tap = blockIdx.y;
    const int ddy = tap / 3 - 1, ddx = tap % 3 - 1;
    const int64_t P = (int64_t)B * H * W;
    v4f acc[MB][NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = v4f{0.f, 0.f, 0.f, 0.f};
    const int64_t pstart = (int64_t)c * chunk;
    for (int g = 0; g < chunk; g += 16) {
        int64_t pp[4];
        int64_t src[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t p = pstart + g + kq * 4 + t;
            pp[t] = p < P ? p : -1;
            src[t] = -1;
            if (pp[t] >= 0) {
                const int w = (int)(p % W), h = (int)((p / W) % H);
                const int hh = h + ddy, ww = w + ddx;
                if (hh >= 0 && hh < H && ww >= 0 && ww < W) src[t] = p + (int64_t)ddy * W + ddx;
            }
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            v4s aop;
#pragma unroll
            for (int t = 0; t < 4; ++t)
                aop[t] = (short)(pp[t] >= 0 ? dy[pp[t] * COUT + mb * 16 + col] : 0);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                v4s bop;
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    bop[t] = (short)(src[t] >= 0 ? x[src[t] * CIN + nb * 16 + col] : 0);
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aop, bop, acc[mb][nb], 0, 0, 0);
            }
        }
    }
    float* out = part + ((size_t)c * 9 + tap) * COUT * CIN;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int t = 0; t < 4; ++t)
                out[(size_t)(mb * 16 + kq * 4 + t) * CIN + nb * 16 + col] = acc[mb][nb][t];
}

// g[k] = sum over chunks of part[chunk][k]: a block takes 32 consecutive k;
// its eight 32-lane groups sum every eighth chunk, then a fixed-order tree
// over the groups in LDS (deterministic)
__global__ void __launch_bounds__(256) sum_chunks_kernel(const float* __restrict__ part, int nchunks, int n,
                                                         float* __restrict__ g)
{
    __shared__ float sh[256];
    const int kl = threadIdx.x & 31, grp = threadIdx.x >> 5;
    const int k = blockIdx.x * 32 + kl;
    float s = 0.f;
    if (k < n)
        for (int c = grp; c < nchunks; c += 8) s += part[(size_t)c * n + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int half = 4; half >= 1; half /= 2) {
        if (grp < half) sh[threadIdx.x] += sh[threadIdx.x + half * 32];
        __syncthreads();
    }
    if (grp == 0 && k < n) g[k] = sh[kl];
}

// bias gradient partials: part[chunk][co] = sum_p dpre[p][co] over the chunk.
// 256 threads: 256 / COUT pixel lanes per channel, each summing every
// (256 / COUT)-th pixel of the chunk, then a fixed-order tree over the lanes
// in LDS (deterministic).
template <int COUT>
__global__ void __launch_bounds__(256) bias_grad_partial_kernel(const float* __restrict__ dpre, int64_t P,
                                                                int chunk, float* __restrict__ part)
{
    constexpr int G = 256 / COUT;
    __shared__ float sh[256];
    const int co = threadIdx.x % COUT, g = threadIdx.x / COUT;
    const int64_t p0 = (int64_t)blockIdx.x * chunk;
    const int64_t pend = min(p0 + chunk, P);
    float s = 0.f;
    for (int64_t p = p0 + g; p < pend; p += G) s += dpre[p * COUT + co];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int half = G / 2; half >= 1; half /= 2) {
        if (g < half) sh[threadIdx.x] += sh[threadIdx.x + half * COUT];
        __syncthreads();
    }
    if (g == 0) part[(size_t)blockIdx.x * COUT + co] = sh[co];
}

// ---------------------------------------------------------------- pointwise
__device__ inline float gelu_f(float x)
{
    const float c = 0.7978845608028654f;
    const float u = c * (x + 0.044715f * x * x * x);
    return 0.5f * x * (1.f + tanhf(u));
}
__device__ inline float gelu_d(float x)
{
    const float c = 0.7978845608028654f;
    const float u = c * (x + 0.044715f * x * x * x);
    const float t = tanhf(u);
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * x * x);
}

__global__ void gelu_fwd_kernel(const float* __restrict__ pre, uint16_t* __restrict__ act, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) act[t] = f2bf(gelu_f(pre[t]));
}

// dpre = dact * gelu'(pre) (FP32 and BF16 copies)
__global__ void gelu_bwd_kernel(const float* __restrict__ pre, const float* __restrict__ dact,
                                float* __restrict__ dpre, uint16_t* __restrict__ dpre_bf, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float g = dact[t] * gelu_d(pre[t]);
    dpre[t] = g;
    dpre_bf[t] = f2bf(g);
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = f2bf(a[t]);
}

// 2x2 average pooling with zero padding at odd edges
__global__ void pool_fwd_kernel(const uint16_t* __restrict__ x, int B, int H, int W, int C,
                                uint16_t* __restrict__ y)
{
    const int H2 = (H + 1) / 2, W2 = (W + 1) / 2;
    const int64_t n = (int64_t)B * H2 * W2 * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = (int)(t % C);
    const int64_t q = t / C;
    const int w2 = (int)(q % W2), h2 = (int)((q / W2) % H2), b = (int)(q / ((int64_t)W2 * H2));
    float s = 0.f;
    for (int a = 0; a < 2; ++a)
        for (int e = 0; e < 2; ++e) {
            const int h = 2 * h2 + a, w = 2 * w2 + e;
            if (h < H && w < W) s += bf2f(x[(((int64_t)b * H + h) * W + w) * C + c]);
        }
    y[t] = f2bf(0.25f * s);
}

// dx[b][h][w][c] += 0.25 * dy[b][h/2][w/2][c]
__global__ void pool_bwd_add_kernel(const float* __restrict__ dy, int B, int H, int W, int C,
                                    float* __restrict__ dx)
{
    const int H2 = (H + 1) / 2, W2 = (W + 1) / 2;
    const int64_t n = (int64_t)B * H * W * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = (int)(t % C);
    const int64_t q = t / C;
    const int w = (int)(q % W), h = (int)((q / W) % H), b = (int)(q / ((int64_t)W * H));
    dx[t] += 0.25f * dy[(((int64_t)b * H2 + h / 2) * W2 + w / 2) * C + c];
}

// u = up(coarse) + skip (nearest upsampling, sum rounded to BF16)
__global__ void upadd_fwd_kernel(const uint16_t* __restrict__ coarse, const uint16_t* __restrict__ skip,
                                 int B, int H, int W, int C, uint16_t* __restrict__ u)
{
    const int H2 = (H + 1) / 2, W2 = (W + 1) / 2;
    const int64_t n = (int64_t)B * H * W * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = (int)(t % C);
    const int64_t q = t / C;
    const int w = (int)(q % W), h = (int)((q / W) % H), b = (int)(q / ((int64_t)W * H));
    u[t] = f2bf(bf2f(coarse[(((int64_t)b * H2 + h / 2) * W2 + w / 2) * C + c]) + bf2f(skip[t]));
}

// dskip = du; dcoarse[h2][w2] = sum of du over the 2x2 block
__global__ void upadd_bwd_kernel(const float* __restrict__ du, int B, int H, int W, int C,
                                 float* __restrict__ dskip, float* __restrict__ dcoarse)
{
    const int H2 = (H + 1) / 2, W2 = (W + 1) / 2;
    const int64_t n2 = (int64_t)B * H2 * W2 * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (int64_t)B * H * W * C) dskip[t] = du[t];
    if (t >= n2) return;
    const int c = (int)(t % C);
    const int64_t q = t / C;
    const int w2 = (int)(q % W2), h2 = (int)((q / W2) % H2), b = (int)(q / ((int64_t)W2 * H2));
    float s = 0.f;
    for (int a = 0; a < 2; ++a)
        for (int e = 0; e < 2; ++e) {
            const int h = 2 * h2 + a, w = 2 * w2 + e;
            if (h < H && w < W) s += du[(((int64_t)b * H + h) * W + w) * C + c];
        }
    dcoarse[t] = s;
}

// MSE over the first CNN_OUT of 16 output channels: partial sums per block and
// dout = 2 (out - tgt) / (P * CNN_OUT)
__global__ void __launch_bounds__(256) mse_kernel(const float* __restrict__ out, const float* __restrict__ tgt,
                                                  int64_t P, float* __restrict__ dout,
                                                  double* __restrict__ part)
{
    __shared__ double sh[256];
    const double inv = 1.0 / ((double)P * CNN_OUT);
    double s = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (int64_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < 16; ++c) {
            float g = 0.f;
            if (c < CNN_OUT) {
                const float d = out[p * 16 + c] - tgt[p * CNN_OUT + c];
                s += (double)d * d;
                g = (float)(2.0 * d * inv);
            }
            if (dout) dout[p * 16 + c] = g;
        }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] * inv;
}

__global__ void sum_double_kernel(const double* __restrict__ part, int n, double* __restrict__ out)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += part[k];
        out[0] = s;
    }
}

// AdamW on the FP32 master parameters
__global__ void adamw_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                             float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                             float wd, float c1, float c2)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    const float mh = mt / c1, vh = vt / c2;
    p[t] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[t]);
}

// BF16 weights: w[tap][co][ci] and the flipped transpose wT[tap][ci][co] =
// w[8-tap][co][ci] used by the input-gradient convolution.
__global__ void weights_bf16_kernel(const float* __restrict__ w, int cin, int cout,
                                    uint16_t* __restrict__ wb, uint16_t* __restrict__ wT)
{
    const int n = 9 * cin * cout;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int ci = t % cin, co = (t / cin) % cout, tap = t / (cin * cout);
    const uint16_t b = f2bf(w[t]);
    wb[t] = b;
    if (wT) wT[((size_t)(8 - tap) * cin + ci) * cout + co] = b;
}

// ------------------------------------------------- solver-vector packing
// RMS of the three components of r and z (planar, n per plane):
// part[block][6] partial sums of squares
__global__ void __launch_bounds__(256) rms6_partial_kernel(const double* __restrict__ r,
                                                           const double* __restrict__ z, int64_t n,
                                                           double* __restrict__ part)
{
    __shared__ double sh[6][256];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        for (int k = 0; k < 3; ++k) {
            const double a = r[k * n + p], b = z[k * n + p];
            s[k] += a * a;
            s[3 + k] += b * b;
        }
    for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 6) part[(size_t)blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

// One scale for all of r and one for all of z (RMS over the three
// components): the components keep their relative sizes, so a component that
// is zero (the x-momentum and continuity parts of a body-force right-hand
// side) stays zero instead of being blown up to unit size.
__global__ void rms6_finish_kernel(const double* __restrict__ part, int nb, int64_t n, double* __restrict__ sc)
{
    const int k = threadIdx.x;
    if (k >= 2) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < 3; ++c) s += part[(size_t)b * 6 + 3 * k + c];
    const double rms = std::sqrt(s / (3.0 * (double)n));
    const double v = (rms > 0.0 && rms < 1e300) ? rms : 1.0;
    for (int c = 0; c < 3; ++c) sc[3 * k + c] = v;
}

// network input [p][7] (FP32) or [p][16] (BF16, zero padded)
__device__ inline float cnn_channel(int k, int64_t p, int64_t n, const double* etan, const double* r,
                                    const double* z, const double* sc)
{
    if (k == 0) return (float)((log10(etan[p]) - 21.5) / 3.5);
    if (k <= 3) return (float)(r[(k - 1) * n + p] / sc[k - 1]);
    return (float)(z[(k - 4) * n + p] / sc[k - 1]);
}

__global__ void pack_input_f32_kernel(int64_t n, const double* __restrict__ etan, const double* __restrict__ r,
                                      const double* __restrict__ z, const double* __restrict__ sc,
                                      float* __restrict__ in)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int k = 0; k < CNN_IN; ++k) in[p * CNN_IN + k] = cnn_channel(k, p, n, etan, r, z, sc);
}

__global__ void pack_input_bf16_kernel(int64_t n, const double* __restrict__ etan, const double* __restrict__ r,
                                       const double* __restrict__ z, const double* __restrict__ sc,
                                       uint16_t* __restrict__ in)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int k = 0; k < 16; ++k) in[p * 16 + k] = k < CNN_IN ? f2bf(cnn_channel(k, p, n, etan, r, z, sc)) : 0;
}

// FP32 [P][7] -> BF16 [P][16]
__global__ void input_to_bf16_kernel(const float* __restrict__ in, int64_t P, uint16_t* __restrict__ out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    for (int k = 0; k < 16; ++k) out[p * 16 + k] = k < CNN_IN ? f2bf(in[p * CNN_IN + k]) : 0;
}

__global__ void pack_target_kernel(int64_t n, const double* __restrict__ z, const double* __restrict__ e,
                                   const double* __restrict__ sc, float* __restrict__ tgt)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int k = 0; k < 3; ++k) tgt[p * 3 + k] = (float)((e[k * n + p] - z[k * n + p]) / sc[3 + k]);
}

__global__ void apply_correction_kernel(int64_t n, const double* __restrict__ z, const float* __restrict__ out,
                                        const double* __restrict__ sc, double* __restrict__ zc)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int k = 0; k < 3; ++k) zc[k * n + p] = z[k * n + p] + sc[3 + k] * (double)out[p * 16 + k];
}

__global__ void output_unpad_kernel(const float* __restrict__ out16, int64_t P, float* __restrict__ out3)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    for (int k = 0; k < CNN_OUT; ++k) out3[p * CNN_OUT + k] = out16[p * 16 + k];
}

// ---------------------------------------------------------------- network
static unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

struct ConvLayer {
    int cin, cout;            // padded channel counts (multiples of 16)
    int cin_real, cout_real;
    int level;
    bool gelu;
    int64_t w_off, b_off;     // offsets into the flat FP32 parameter vector
};

struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevMem() { if (p) (void)hipFree(p); }
    int ensure(size_t b)
    {
        if (p && bytes >= b) return EBG_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        EBG_HIP(hipMalloc(&p, b ? b : 16));
        bytes = b;
        return EBG_OK;
    }
    template <typename T> T* as() const { return (T*)p; }
};

enum { L_IN = 0, L_A0, L_E1A, L_E1B, L_E2A, L_E2B, L_E3A, L_E3B, L_D2A, L_D2B, L_D1A, L_D1B,
       L_D0A, L_D0B, L_OUT, NLAYERS };

constexpr int WGRAD_CHUNK = 256;

struct CnnNet {
    std::vector<ConvLayer> layers;
    int64_t nparams = 0;
    DevMem params, grads, adam_m, adam_v;
    std::vector<DevMem> wbf, wT;     // per layer BF16 weights and flipped transposes
    int64_t step = 0;
    // activations for the current (B, H, W)
    int B = 0, H = 0, W = 0;
    bool train = false;
    int lh[4], lw[4];
    int64_t lp[4];
    DevMem in_bf, pre[NLAYERS], act[NLAYERS], pool1, pool2, pool3, up2, up1, up0;
    // backward
    DevMem dact[NLAYERS], dpre[NLAYERS], dpre_bf[NLAYERS], dpool1, dpool2, dpool3, dup2, dup1, dup0;
    DevMem dA1, dA3, dA5, wpart, bpart, loss_part, loss;
    // inference helpers
    DevMem scales, rms_part, in_f32;
};

static void build_layers(CnnNet& n)
{
    struct Spec { int cin, cout, cin_real, cout_real, level; bool gelu; };
    const Spec spec[NLAYERS] = {
        {16, 32, CNN_IN, 32, 0, true}, {32, 32, 32, 32, 0, true},
        {32, 32, 32, 32, 1, true},     {32, 32, 32, 32, 1, true},
        {32, 32, 32, 32, 2, true},     {32, 32, 32, 32, 2, true},
        {32, 32, 32, 32, 3, true},     {32, 32, 32, 32, 3, true},
        {32, 32, 32, 32, 2, true},     {32, 32, 32, 32, 2, true},
        {32, 32, 32, 32, 1, true},     {32, 32, 32, 32, 1, true},
        {32, 32, 32, 32, 0, true},     {32, 32, 32, 32, 0, true},
        {32, 16, 32, CNN_OUT, 0, false}};
    int64_t off = 0;
    n.layers.clear();
    for (int L = 0; L < NLAYERS; ++L) {
        ConvLayer c;
        c.cin = spec[L].cin;
        c.cout = spec[L].cout;
        c.cin_real = spec[L].cin_real;
        c.cout_real = spec[L].cout_real;
        c.level = spec[L].level;
        c.gelu = spec[L].gelu;
        c.w_off = off;
        off += 9LL * c.cin * c.cout;
        c.b_off = off;
        off += c.cout;
        n.layers.push_back(c);
    }
    n.nparams = off;
}

static void conv_fwd(const ConvLayer& c, const uint16_t* x, int B, int H, int W, const uint16_t* wb,
                     const float* bias, float* y, hipStream_t s)
{
    const int64_t P = (int64_t)B * H * W;
    const dim3 g(nblk(P, 256)), b(256);
    if (c.cin == 16 && c.cout == 32) hipLaunchKernelGGL((conv3x3_fwd_kernel<16, 32>), g, b, 0, s, x, B, H, W, wb, bias, y);
    else if (c.cin == 32 && c.cout == 32) hipLaunchKernelGGL((conv3x3_fwd32_kernel<32>), g, b, 0, s, x, B, H, W, wb, bias, y);
    else if (c.cin == 32 && c.cout == 16) hipLaunchKernelGGL((conv3x3_fwd32_kernel<16>), g, b, 0, s, x, B, H, W, wb, bias, y);
}

// input gradient: conv of dy (cout channels) with the flipped transpose -> cin channels
static void conv_bwd_data(const ConvLayer& c, const uint16_t* dy, int B, int H, int W, const uint16_t* wT,
                          float* dx, hipStream_t s)
{
    const int64_t P = (int64_t)B * H * W;
    const dim3 g(nblk(P, 256)), b(256);
    if (c.cout == 32 && c.cin == 32) hipLaunchKernelGGL((conv3x3_fwd32_kernel<32>), g, b, 0, s, dy, B, H, W, wT, (const float*)nullptr, dx);
    else if (c.cout == 16 && c.cin == 32) hipLaunchKernelGGL((conv3x3_fwd_kernel<16, 32>), g, b, 0, s, dy, B, H, W, wT, (const float*)nullptr, dx);
    else if (c.cout == 32 && c.cin == 16) hipLaunchKernelGGL((conv3x3_fwd32_kernel<16>), g, b, 0, s, dy, B, H, W, wT, (const float*)nullptr, dx);
}

static int conv_wgrad(CnnNet& n, const ConvLayer& c, const uint16_t* x, const uint16_t* dy_bf,
                      const float* dy_f32, int B, int H, int W, hipStream_t s)
{
    const int64_t P = (int64_t)B * H * W;
    const int nch = (int)((P + WGRAD_CHUNK - 1) / WGRAD_CHUNK);
    const size_t wn = 9ULL * c.cin * c.cout;
    int rc;
    if ((rc = n.wpart.ensure((size_t)nch * wn * sizeof(float)))) return rc;
    if ((rc = n.bpart.ensure((size_t)nch * c.cout * sizeof(float)))) return rc;
    const dim3 g(nch, 9);
    if (c.cin == 16 && c.cout == 32)
        hipLaunchKernelGGL((conv3x3_wgrad_kernel<16, 32>), g, dim3(64), 0, s, x, dy_bf, B, H, W, WGRAD_CHUNK, n.wpart.as<float>());
    else if (c.cin == 32 && c.cout == 32)
        hipLaunchKernelGGL((conv3x3_wgrad_kernel<32, 32>), g, dim3(64), 0, s, x, dy_bf, B, H, W, WGRAD_CHUNK, n.wpart.as<float>());
    else
        hipLaunchKernelGGL((conv3x3_wgrad_kernel<32, 16>), g, dim3(64), 0, s, x, dy_bf, B, H, W, WGRAD_CHUNK, n.wpart.as<float>());
    float* gw = n.grads.as<float>() + c.w_off;
    hipLaunchKernelGGL(sum_chunks_kernel, dim3((unsigned)((wn + 31) / 32)), dim3(256), 0, s,
                       (const float*)n.wpart.as<float>(), nch, (int)wn, gw);
    if (c.cout == 32)
        hipLaunchKernelGGL(bias_grad_partial_kernel<32>, dim3(nch), dim3(256), 0, s, dy_f32, P, WGRAD_CHUNK, n.bpart.as<float>());
    else
        hipLaunchKernelGGL(bias_grad_partial_kernel<16>, dim3(nch), dim3(256), 0, s, dy_f32, P, WGRAD_CHUNK, n.bpart.as<float>());
    hipLaunchKernelGGL(sum_chunks_kernel, dim3((unsigned)((c.cout + 31) / 32)), dim3(256), 0, s,
                       (const float*)n.bpart.as<float>(), nch, c.cout, n.grads.as<float>() + c.b_off);
    return EBG_OK;
}

static int refresh_bf16(CnnNet& n, hipStream_t s)
{
    int rc;
    for (int L = 0; L < NLAYERS; ++L) {
        const ConvLayer& c = n.layers[L];
        const size_t wn = 9ULL * c.cin * c.cout;
        if ((rc = n.wbf[L].ensure(wn * 2)) || (rc = n.wT[L].ensure(wn * 2))) return rc;
        hipLaunchKernelGGL(weights_bf16_kernel, dim3(nblk((int64_t)wn)), dim3(256), 0, s,
                           (const float*)(n.params.as<float>() + c.w_off), c.cin, c.cout, n.wbf[L].as<uint16_t>(),
                           n.wT[L].as<uint16_t>());
    }
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

static int alloc_activations(CnnNet& n, int B, int H, int W, bool train)
{
    n.B = B;
    n.H = H;
    n.W = W;
    n.train = train;
    n.lh[0] = H;
    n.lw[0] = W;
    for (int k = 1; k < 4; ++k) {
        n.lh[k] = (n.lh[k - 1] + 1) / 2;
        n.lw[k] = (n.lw[k - 1] + 1) / 2;
    }
    for (int k = 0; k < 4; ++k) n.lp[k] = (int64_t)B * n.lh[k] * n.lw[k];
    int rc;
    if ((rc = n.in_bf.ensure((size_t)n.lp[0] * 16 * 2))) return rc;
    for (int L = 0; L < NLAYERS; ++L) {
        const ConvLayer& c = n.layers[L];
        const int64_t P = n.lp[c.level];
        if ((rc = n.pre[L].ensure((size_t)P * c.cout * 4))) return rc;
        if ((rc = n.act[L].ensure((size_t)P * c.cout * 2))) return rc;
        if (train) {
            if ((rc = n.dact[L].ensure((size_t)P * c.cout * 4))) return rc;
            if ((rc = n.dpre[L].ensure((size_t)P * c.cout * 4))) return rc;
            if ((rc = n.dpre_bf[L].ensure((size_t)P * c.cout * 2))) return rc;
        }
    }
    const size_t c32 = 32;
    if ((rc = n.pool1.ensure(n.lp[1] * c32 * 2)) || (rc = n.pool2.ensure(n.lp[2] * c32 * 2)) ||
        (rc = n.pool3.ensure(n.lp[3] * c32 * 2)) || (rc = n.up2.ensure(n.lp[2] * c32 * 2)) ||
        (rc = n.up1.ensure(n.lp[1] * c32 * 2)) || (rc = n.up0.ensure(n.lp[0] * c32 * 2)))
        return rc;
    if (train) {
        if ((rc = n.dpool1.ensure(n.lp[1] * c32 * 4)) || (rc = n.dpool2.ensure(n.lp[2] * c32 * 4)) ||
            (rc = n.dpool3.ensure(n.lp[3] * c32 * 4)) || (rc = n.dup2.ensure(n.lp[2] * c32 * 4)) ||
            (rc = n.dup1.ensure(n.lp[1] * c32 * 4)) || (rc = n.dup0.ensure(n.lp[0] * c32 * 4)) ||
            (rc = n.dA1.ensure(n.lp[0] * c32 * 4)) || (rc = n.dA3.ensure(n.lp[1] * c32 * 4)) ||
            (rc = n.dA5.ensure(n.lp[2] * c32 * 4)))
            return rc;
    }
    return EBG_OK;
}

// input pointer (BF16) of layer L after the forward pass
static const uint16_t* layer_input(const CnnNet& n, int L)
{
    switch (L) {
        case L_IN: return n.in_bf.as<uint16_t>();
        case L_E1A: return n.pool1.as<uint16_t>();
        case L_E2A: return n.pool2.as<uint16_t>();
        case L_E3A: return n.pool3.as<uint16_t>();
        case L_D2A: return n.up2.as<uint16_t>();
        case L_D1A: return n.up1.as<uint16_t>();
        case L_D0A: return n.up0.as<uint16_t>();
        default: return n.act[L - 1].as<uint16_t>();
    }
}

// forward from n.in_bf; the output (FP32 [P0][16]) is n.pre[L_OUT]
static void forward(CnnNet& n, hipStream_t s)
{
    const int B = n.B;
    auto run = [&](int L) {
        const ConvLayer& c = n.layers[L];
        const int k = c.level;
        conv_fwd(c, layer_input(n, L), B, n.lh[k], n.lw[k], n.wbf[L].as<uint16_t>(),
                 n.params.as<float>() + c.b_off, n.pre[L].as<float>(), s);
        if (c.gelu) {
            const int64_t m = n.lp[k] * c.cout;
            hipLaunchKernelGGL(gelu_fwd_kernel, dim3(nblk(m)), dim3(256), 0, s, (const float*)n.pre[L].as<float>(),
                               n.act[L].as<uint16_t>(), m);
        }
    };
    auto pool = [&](const DevMem& x, int k, DevMem& y) {
        const int64_t m = n.lp[k + 1] * 32;
        hipLaunchKernelGGL(pool_fwd_kernel, dim3(nblk(m)), dim3(256), 0, s, (const uint16_t*)x.as<uint16_t>(), B,
                           n.lh[k], n.lw[k], 32, y.as<uint16_t>());
    };
    auto upadd = [&](const DevMem& coarse, const DevMem& skip, int k, DevMem& u) {
        const int64_t m = n.lp[k] * 32;
        hipLaunchKernelGGL(upadd_fwd_kernel, dim3(nblk(m)), dim3(256), 0, s, (const uint16_t*)coarse.as<uint16_t>(),
                           (const uint16_t*)skip.as<uint16_t>(), B, n.lh[k], n.lw[k], 32, u.as<uint16_t>());
    };
    run(L_IN);
    run(L_A0);
    pool(n.act[L_A0], 0, n.pool1);
    run(L_E1A);
    run(L_E1B);
    pool(n.act[L_E1B], 1, n.pool2);
    run(L_E2A);
    run(L_E2B);
    pool(n.act[L_E2B], 2, n.pool3);
    run(L_E3A);
    run(L_E3B);
    upadd(n.act[L_E3B], n.act[L_E2B], 2, n.up2);
    run(L_D2A);
    run(L_D2B);
    upadd(n.act[L_D2B], n.act[L_E1B], 1, n.up1);
    run(L_D1A);
    run(L_D1B);
    upadd(n.act[L_D1B], n.act[L_A0], 0, n.up0);
    run(L_D0A);
    run(L_D0B);
    run(L_OUT);
}

// backward from dL/dout in n.dpre[L_OUT] (FP32 and BF16); fills n.grads
static int backward(CnnNet& n, hipStream_t s)
{
    const int B = n.B;
    int rc;
    // conv L: weight grads, then dL/d(input) into dx (FP32) when wanted
    auto conv_back = [&](int L, float* dx) -> int {
        const ConvLayer& c = n.layers[L];
        const int k = c.level;
        int r = conv_wgrad(n, c, layer_input(n, L), n.dpre_bf[L].as<uint16_t>(), n.dpre[L].as<float>(), B,
                           n.lh[k], n.lw[k], s);
        if (r) return r;
        if (dx) conv_bwd_data(c, n.dpre_bf[L].as<uint16_t>(), B, n.lh[k], n.lw[k], n.wT[L].as<uint16_t>(), dx, s);
        return EBG_OK;
    };
    // dpre = dact * gelu'(pre)
    auto gelu_back = [&](int L, const float* dact) {
        const ConvLayer& c = n.layers[L];
        const int64_t m = n.lp[c.level] * c.cout;
        hipLaunchKernelGGL(gelu_bwd_kernel, dim3(nblk(m)), dim3(256), 0, s, (const float*)n.pre[L].as<float>(), dact,
                           n.dpre[L].as<float>(), n.dpre_bf[L].as<uint16_t>(), m);
    };
    auto upadd_back = [&](const DevMem& du, int k, DevMem& dskip, DevMem& dcoarse) {
        const int64_t m = n.lp[k] * 32;
        hipLaunchKernelGGL(upadd_bwd_kernel, dim3(nblk(m)), dim3(256), 0, s, (const float*)du.as<float>(), B,
                           n.lh[k], n.lw[k], 32, dskip.as<float>(), dcoarse.as<float>());
    };
    auto pool_back = [&](const DevMem& dp, int k, DevMem& dx) {
        const int64_t m = n.lp[k] * 32;
        hipLaunchKernelGGL(pool_bwd_add_kernel, dim3(nblk(m)), dim3(256), 0, s, (const float*)dp.as<float>(), B,
                           n.lh[k], n.lw[k], 32, dx.as<float>());
    };
    if ((rc = conv_back(L_OUT, n.dact[L_D0B].as<float>()))) return rc;
    gelu_back(L_D0B, n.dact[L_D0B].as<float>());
    if ((rc = conv_back(L_D0B, n.dact[L_D0A].as<float>()))) return rc;
    gelu_back(L_D0A, n.dact[L_D0A].as<float>());
    if ((rc = conv_back(L_D0A, n.dup0.as<float>()))) return rc;
    upadd_back(n.dup0, 0, n.dA1, n.dact[L_D1B]);          // skip x0, coarse y1
    gelu_back(L_D1B, n.dact[L_D1B].as<float>());
    if ((rc = conv_back(L_D1B, n.dact[L_D1A].as<float>()))) return rc;
    gelu_back(L_D1A, n.dact[L_D1A].as<float>());
    if ((rc = conv_back(L_D1A, n.dup1.as<float>()))) return rc;
    upadd_back(n.dup1, 1, n.dA3, n.dact[L_D2B]);          // skip x1, coarse y2
    gelu_back(L_D2B, n.dact[L_D2B].as<float>());
    if ((rc = conv_back(L_D2B, n.dact[L_D2A].as<float>()))) return rc;
    gelu_back(L_D2A, n.dact[L_D2A].as<float>());
    if ((rc = conv_back(L_D2A, n.dup2.as<float>()))) return rc;
    upadd_back(n.dup2, 2, n.dA5, n.dact[L_E3B]);          // skip x2, coarse x3
    gelu_back(L_E3B, n.dact[L_E3B].as<float>());
    if ((rc = conv_back(L_E3B, n.dact[L_E3A].as<float>()))) return rc;
    gelu_back(L_E3A, n.dact[L_E3A].as<float>());
    if ((rc = conv_back(L_E3A, n.dpool3.as<float>()))) return rc;
    pool_back(n.dpool3, 2, n.dA5);
    gelu_back(L_E2B, n.dA5.as<float>());
    if ((rc = conv_back(L_E2B, n.dact[L_E2A].as<float>()))) return rc;
    gelu_back(L_E2A, n.dact[L_E2A].as<float>());
    if ((rc = conv_back(L_E2A, n.dpool2.as<float>()))) return rc;
    pool_back(n.dpool2, 1, n.dA3);
    gelu_back(L_E1B, n.dA3.as<float>());
    if ((rc = conv_back(L_E1B, n.dact[L_E1A].as<float>()))) return rc;
    gelu_back(L_E1A, n.dact[L_E1A].as<float>());
    if ((rc = conv_back(L_E1A, n.dpool1.as<float>()))) return rc;
    pool_back(n.dpool1, 0, n.dA1);
    gelu_back(L_A0, n.dA1.as<float>());
    if ((rc = conv_back(L_A0, n.dact[L_IN].as<float>()))) return rc;
    gelu_back(L_IN, n.dact[L_IN].as<float>());
    if ((rc = conv_back(L_IN, nullptr))) return rc;
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

// deterministic host RNG for the initial weights (splitmix64)
static uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

int cnn_create(CnnNet** out, uint64_t seed)
{
    std::unique_ptr<CnnNet> n(new CnnNet());
    build_layers(*n);
    n->wbf.resize(NLAYERS);
    n->wT.resize(NLAYERS);
    std::vector<float> h((size_t)n->nparams, 0.f);
    uint64_t st = seed ? seed : 1;
    for (const ConvLayer& c : n->layers) {
        if (!c.gelu) continue;  // output layer starts at zero: the network starts as the plain V-cycle
        const double a = std::sqrt(6.0 / (9.0 * c.cin_real));
        for (int tap = 0; tap < 9; ++tap)
            for (int co = 0; co < c.cout; ++co)
                for (int ci = 0; ci < c.cin; ++ci) {
                    const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
                    const bool real = ci < c.cin_real && co < c.cout_real;
                    h[(size_t)c.w_off + ((size_t)tap * c.cout + co) * c.cin + ci] = real ? (float)((2.0 * u - 1.0) * a) : 0.f;
                }
    }
    int rc;
    const size_t nb = (size_t)n->nparams * sizeof(float);
    if ((rc = n->params.ensure(nb)) || (rc = n->grads.ensure(nb)) || (rc = n->adam_m.ensure(nb)) ||
        (rc = n->adam_v.ensure(nb)))
        return rc;
    EBG_HIP(hipMemcpy(n->params.p, h.data(), nb, hipMemcpyHostToDevice));
    EBG_HIP(hipMemset(n->grads.p, 0, nb));
    EBG_HIP(hipMemset(n->adam_m.p, 0, nb));
    EBG_HIP(hipMemset(n->adam_v.p, 0, nb));
    if ((rc = refresh_bf16(*n, nullptr))) return rc;
    EBG_HIP(hipDeviceSynchronize());
    *out = n.release();
    return EBG_OK;
}

void cnn_destroy(CnnNet* n) { delete n; }
int64_t cnn_num_params(const CnnNet* n) { return n->nparams; }

static int scales_of(CnnNet& n, int64_t np, const double* r, const double* z, hipStream_t s)
{
    constexpr int NB = 256;
    int rc;
    if ((rc = n.rms_part.ensure(NB * 6 * 8)) || (rc = n.scales.ensure(6 * 8))) return rc;
    hipLaunchKernelGGL(rms6_partial_kernel, dim3(NB), dim3(256), 0, s, r, z, np, n.rms_part.as<double>());
    hipLaunchKernelGGL(rms6_finish_kernel, dim3(1), dim3(64), 0, s, (const double*)n.rms_part.as<double>(), NB, np,
                       n.scales.as<double>());
    return EBG_OK;
}

int cnn_correct(CnnNet* n, int H, int W, const double* etan, const double* r, const double* z, double* zc,
                hipStream_t s)
{
    int rc;
    if (n->B != 1 || n->H != H || n->W != W || n->train)
        if ((rc = alloc_activations(*n, 1, H, W, false))) return rc;
    const int64_t np = (int64_t)H * W;
    if ((rc = scales_of(*n, np, r, z, s))) return rc;
    hipLaunchKernelGGL(pack_input_bf16_kernel, dim3(nblk(np)), dim3(256), 0, s, np, etan, r, z,
                       (const double*)n->scales.as<double>(), n->in_bf.as<uint16_t>());
    forward(*n, s);
    hipLaunchKernelGGL(apply_correction_kernel, dim3(nblk(np)), dim3(256), 0, s, np, z,
                       (const float*)n->pre[L_OUT].as<float>(), (const double*)n->scales.as<double>(), zc);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int cnn_pack_sample(CnnNet* n, int H, int W, const double* etan, const double* r, const double* z,
                    const double* e, float* in_img, float* tgt_img, hipStream_t s)
{
    const int64_t np = (int64_t)H * W;
    int rc;
    if ((rc = scales_of(*n, np, r, z, s))) return rc;
    hipLaunchKernelGGL(pack_input_f32_kernel, dim3(nblk(np)), dim3(256), 0, s, np, etan, r, z,
                       (const double*)n->scales.as<double>(), in_img);
    hipLaunchKernelGGL(pack_target_kernel, dim3(nblk(np)), dim3(256), 0, s, np, z, e,
                       (const double*)n->scales.as<double>(), tgt_img);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

// loss (and gradients when want_grad) of a batch of FP32 inputs [B][H][W][7]
// against targets [B][H][W][3]
static int loss_and_grad(CnnNet& n, int B, int H, int W, const float* in, const float* tgt, bool want_grad,
                         double* loss_host, hipStream_t s)
{
    int rc;
    if (n.B != B || n.H != H || n.W != W || n.train != want_grad)
        if ((rc = alloc_activations(n, B, H, W, want_grad))) return rc;
    const int64_t P = n.lp[0];
    hipLaunchKernelGGL(input_to_bf16_kernel, dim3(nblk(P)), dim3(256), 0, s, in, P, n.in_bf.as<uint16_t>());
    forward(n, s);
    constexpr int NB = 256;
    if ((rc = n.loss_part.ensure(NB * 8)) || (rc = n.loss.ensure(8))) return rc;
    hipLaunchKernelGGL(mse_kernel, dim3(NB), dim3(256), 0, s, (const float*)n.pre[L_OUT].as<float>(), tgt, P,
                       want_grad ? n.dpre[L_OUT].as<float>() : (float*)nullptr, n.loss_part.as<double>());
    hipLaunchKernelGGL(sum_double_kernel, dim3(1), dim3(64), 0, s, (const double*)n.loss_part.as<double>(), NB,
                       n.loss.as<double>());
    if (want_grad) {
        hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(nblk(P * 16)), dim3(256), 0, s,
                           (const float*)n.dpre[L_OUT].as<float>(), n.dpre_bf[L_OUT].as<uint16_t>(), P * 16);
        if ((rc = backward(n, s))) return rc;
    }
    if (loss_host) {
        EBG_HIP(hipMemcpyAsync(loss_host, n.loss.p, 8, hipMemcpyDeviceToHost, s));
        EBG_HIP(hipStreamSynchronize(s));
    }
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

}  // namespace ebg

using namespace ebg;

static CnnNet* net_of(ebg_ctx* ctx) { return ctx ? (CnnNet*)ctx->cnn : nullptr; }

extern "C" {

int ebg_cnn_create(ebg_ctx* ctx, uint64_t seed)
{
    EBG_CHECK(ctx, EBG_ERR_ARG, "ebg_cnn_create: ctx is NULL");
    EBG_HIP(hipSetDevice(ctx->device));
    CnnNet* n = nullptr;
    int rc = cnn_create(&n, seed);
    if (rc) return rc;
    if (ctx->cnn) cnn_destroy((CnnNet*)ctx->cnn);
    ctx->cnn = n;
    return EBG_OK;
}

int ebg_cnn_num_params(ebg_ctx* ctx, int64_t* n)
{
    CnnNet* net = net_of(ctx);
    EBG_CHECK(net && n, EBG_ERR_STATE, "ebg_cnn_num_params: no network");
    *n = net->nparams;
    return EBG_OK;
}

int ebg_cnn_get_params(ebg_ctx* ctx, float* host)
{
    CnnNet* net = net_of(ctx);
    EBG_CHECK(net && host, EBG_ERR_STATE, "ebg_cnn_get_params: no network or NULL buffer");
    EBG_HIP(hipStreamSynchronize(ctx->stream));
    EBG_HIP(hipMemcpy(host, net->params.p, net->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return EBG_OK;
}

int ebg_cnn_set_params(ebg_ctx* ctx, const float* host)
{
    CnnNet* net = net_of(ctx);
    EBG_CHECK(net && host, EBG_ERR_STATE, "ebg_cnn_set_params: no network or NULL buffer");
    EBG_HIP(hipStreamSynchronize(ctx->stream));
    EBG_HIP(hipMemcpy(net->params.p, host, net->nparams * sizeof(float), hipMemcpyHostToDevice));
    net->step = 0;
    EBG_HIP(hipMemset(net->adam_m.p, 0, net->nparams * sizeof(float)));
    EBG_HIP(hipMemset(net->adam_v.p, 0, net->nparams * sizeof(float)));
    return refresh_bf16(*net, ctx->stream);
}

int ebg_cnn_forward(ebg_ctx* ctx, int B, int H, int W, const float* in, float* out)
{
    CnnNet* net = net_of(ctx);
    EBG_CHECK(net, EBG_ERR_STATE, "ebg_cnn_forward: no network");
    EBG_CHECK(B >= 1 && H >= 1 && W >= 1 && in && out, EBG_ERR_ARG, "ebg_cnn_forward: bad arguments");
    hipStream_t s = ctx->stream;
    int rc;
    if (net->B != B || net->H != H || net->W != W || net->train)
        if ((rc = alloc_activations(*net, B, H, W, false))) return rc;
    const int64_t P = net->lp[0];
    hipLaunchKernelGGL(input_to_bf16_kernel, dim3(nblk(P)), dim3(256), 0, s, in, P, net->in_bf.as<uint16_t>());
    forward(*net, s);
    hipLaunchKernelGGL(output_unpad_kernel, dim3(nblk(P)), dim3(256), 0, s, (const float*)net->pre[L_OUT].as<float>(),
                       P, out);
    EBG_HIP(hipGetLastError());
    return EBG_OK;
}

int ebg_cnn_loss_and_grad(ebg_ctx* ctx, int B, int H, int W, const float* in, const float* target,
                          double* loss, float* grad_host)
{
    CnnNet* net = net_of(ctx);
    EBG_CHECK(net, EBG_ERR_STATE, "ebg_cnn_loss_and_grad: no network");
    EBG_CHECK(B >= 1 && H >= 1 && W >= 1 && in && target && loss, EBG_ERR_ARG,
              "ebg_cnn_loss_and_grad: bad arguments");
    int rc = loss_and_grad(*net, B, H, W, in, target, true, loss, ctx->stream);
    if (rc) return rc;
    if (grad_host) EBG_HIP(hipMemcpy(grad_host, net->grads.p, net->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return EBG_OK;
}

int ebg_cnn_loss(ebg_ctx* ctx, int B, int H, int W, const float* in, const float* target, double* loss)
{
    CnnNet* net = net_of(ctx);
    EBG_CHECK(net, EBG_ERR_STATE, "ebg_cnn_loss: no network");
    EBG_CHECK(B >= 1 && H >= 1 && W >= 1 && in && target && loss, EBG_ERR_ARG, "ebg_cnn_loss: bad arguments");
    return loss_and_grad(*net, B, H, W, in, target, false, loss, ctx->stream);
}

int ebg_cnn_adamw_step(ebg_ctx* ctx, float lr, float beta1, float beta2, float eps, float weight_decay)
{
    CnnNet* net = net_of(ctx);
    EBG_CHECK(net, EBG_ERR_STATE, "ebg_cnn_adamw_step: no network");
    hipStream_t s = ctx->stream;
    net->step += 1;
    const float c1 = 1.f - std::pow(beta1, (float)net->step);
    const float c2 = 1.f - std::pow(beta2, (float)net->step);
    hipLaunchKernelGGL(adamw_kernel, dim3(nblk(net->nparams)), dim3(256), 0, s, net->params.as<float>(),
                       (const float*)net->grads.as<float>(), net->adam_m.as<float>(), net->adam_v.as<float>(),
                       net->nparams, lr, beta1, beta2, eps, weight_decay, c1, c2);
    return refresh_bf16(*net, s);
}

}  // extern "C"

namespace ebg {
void cnn_release_ctx(ebg_ctx* ctx)
{
    if (ctx && ctx->cnn) {
        cnn_destroy((CnnNet*)ctx->cnn);
        ctx->cnn = nullptr;
    }
}
}  // namespace ebg
