// Reciprocal, division and square root for the stage kernels.
//
// IEEE-exact float64 division compiles to 11 VALU instructions on gfx950
// and sqrt to ~21.  The Riemann solvers, eigenvectors and WENO weights need
// dozens of each per interface, and most share a denominator (1/rho,
// 1/sqrt(rho), 1/c^2), so the kernels form reciprocals once with the
// hardware estimate (v_rcp_f64 / v_rsq_f64, quarter rate) refined by
// Newton-Raphson steps on the FMA pipe, and multiply.  On gfx950 the
// estimates are good to 2^-24.4 (rcp) and 2^-24.2 (rsq) relative; one Newton
// step brings them to 2^-48.7 and 2^-47.8 (20 and 37 units of 2^-53,
// measured over 2^20 values spanning 2^+-60), which is what the reciprocals
// of the stage need: they scale Roe averages, eigenvector entries and WENO
// weights, whose relative errors of 1e-14 move a step's result far less
// than the 1e-12 the device path is held to.  fdiv and fsqrt add one
// residual correction of the result and are within an ulp or two.
#pragma once
#include <hip/hip_runtime.h>

namespace sc {

// 1/x: estimate + one Newton step  r <- r + r(1 - x r)
__device__ __forceinline__ double frcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// a/b with one residual correction of the quotient
__device__ __forceinline__ double fdiv(double a, double b)
{
    double r = frcp(b);
    double q = a * r;
    double e = fma(-b, q, a);
    return fma(e, r, q);
}
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// 1/sqrt(x): estimate + one Newton step  r <- r + r(1/2 - x r^2 / 2)
__device__ __forceinline__ double frsq(double x)
{
    double r = __builtin_amdgcn_rsq(x);
    double e = fma(-(0.5 * x) * r, r, 0.5);
    return fma(r, e, r);
}
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// sqrt(x) = x / sqrt(x) with one residual correction; sqrt(0) = 0
__device__ __forceinline__ double fsqrt(double x)
{
    double r = frsq(x);
    double s = x * r;
    double e = fma(-s, s, x);
    s = fma(e, 0.5 * r, s);
    return x == 0.0 ? 0.0 : s;
}
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

}  // namespace sc
