// Device helpers shared by the nodal and the marker plasticity kernels:
// Julia's max / min on Float64 and YieldStress.calc_yield_stress
// (rheology/plastic_failure/utils/YieldStress.jl:28-80).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace ebg {

// Julia's max / min on Float64: NaN wins and +0.0 > -0.0
static __device__ __forceinline__ double jl_max(double a, double b)
{
    if (a != a || b != b) return a + b;
    if (a == b) return signbit(a) ? b : a;
    return a > b ? a : b;
}
static __device__ __forceinline__ double jl_min(double a, double b)
{
    if (a != a || b != b) return a + b;
    if (a == b) return signbit(a) ? a : b;
    return a < b ? a : b;
}

// YieldStress.calc_yield_stress: Drucker-Prager, optionally on the pressure
// above the hydrostatic fluid pressure (density 1000, g 9.81) with the
// tension cut-off cohesion * cos(phi) (stress_yield_tension_method2)
static __device__ __forceinline__ double yield_stress(double cohesion, double sine_fric, double pressure,
                                                      int use_fluid, double y, double y_sealevel)
{
    const double cosphi = cos(asin(sine_fric));
    double s;
    if (use_fluid == 0) {
        s = cohesion * cosphi + sine_fric * pressure;
    } else {
        const double pf = 1000.0 * 9.81 * jl_max(0.0, y - y_sealevel);
        if (pressure < pf) s = cohesion * cosphi;
        else s = cohesion * cosphi + sine_fric * (pressure - pf);
    }
    return jl_max(s, 0.0);
}

}  // namespace ebg
