// Device helpers shared by the marker kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace ebg {

// clamp(searchsortedlast(g, x), 1, n-1) as a 1-based index
// (interpolation/mapping/MarkerGridMapping.jl:78-86).
//
// For a uniform axis the cell is guessed in O(1) from (x - x0)/h and then
// corrected by exact comparisons against the stored coordinates, so the result
// equals the binary search on the reference's cumulative-sum coordinates bit
// for bit.  Non-uniform axes use the binary search directly.
__device__ __forceinline__ int left_node_index(double x, const double* __restrict__ g, int n,
                                               double x0, double inv_h, int uniform)
{
    if (x != x) return n - 1;  // NaN sorts after every node
    int c;  // candidate searchsortedlast result in [0, n]
    if (uniform) {
        const double t = (x - x0) * inv_h;
        if (t < 0.0) c = 0;
        else if (t >= (double)n) c = n;
        else c = (int)t + 1;
        while (c < n && g[c] <= x) ++c;
        while (c >= 1 && g[c - 1] > x) --c;
    } else {
        int lo = 0, hi = n + 1;
        while (lo < hi - 1) {
            const int mid = lo + ((hi - lo) >> 1);
            if (x < g[mid - 1]) hi = mid;
            else lo = mid;
        }
        c = lo;
    }
    if (c < 1) c = 1;
    if (c > n - 1) c = n - 1;
    return c;
}

// Correctly rounded a/b given rb = RN(1/b): q0 = RN(a*rb) is within 1.5 ulp
// of a/b, the first remainder step makes it faithful, and Markstein's theorem
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, "Newton-Raphson-based division with an FMA") makes
// the second step return RN(a/b): with rb within half an ulp of 1/b and q
// within one ulp of a/b, r = a - b*q is exact and RN(q + r*rb) = RN(a/b).
// The theorem needs no underflow or overflow in r and q: callers use it only
// for numerators that pass div_in_range() and divisors in [2^-100, 2^100]
// (ebg_grid2d_set rejects grid steps outside that range), and take the plain
// IEEE division otherwise, so the result always equals a / b bit for bit.
// The divisors are positive, so a / b has the sign of a; copying that sign
// onto q also turns the +0 the steps return for a = -0 into -0.
__device__ __forceinline__ double rn_div(double a, double b, double rb)
{
    double q = a * rb;
    double r = __builtin_fma(-b, q, a);
    q = __builtin_fma(r, rb, q);
    r = __builtin_fma(-b, q, a);
    q = __builtin_fma(r, rb, q);
    return __builtin_copysign(q, a);
}

// Numerators for which rn_div is exact with divisors and reciprocals between
// 2^-100 and 2^100 (grid steps): zero, or magnitude in [2^-900, 2^900].
__device__ __forceinline__ bool div_in_range(double a)
{
    const double m = fabs(a);
    return m == 0.0 || (m >= 0x1p-900 && m <= 0x1p900);
}

}  // namespace ebg
