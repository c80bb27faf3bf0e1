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

}  // namespace ebg
