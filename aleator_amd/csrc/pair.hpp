// pair.hpp -- FP64 pair arithmetic shared by the grid and Widom kernels.
//
// On gfx950 an FP64 FMA issues at the rate of an unpacked FP32 FMA, so the pair loops run in
// FP64 and only the two slow operations take a shortcut:
//
// - 1/sqrt(r^2) and 1/r^2 start from the hardware FP64 approximations (v_rsq_f64, v_rcp_f64,
//   relative error about 2^-23) and are refined by Newton steps: one step leaves a relative
//   error near 2e-14 (MIXED precision), two steps the last bits (FP64 precision);
// - erfc(alpha r)/r = 1/r - erf(alpha r)/r, where erf(alpha r)/r is an entire function of
//   s = r^2: (2 alpha/sqrt(pi)) G(alpha^2 s) with G(u) = int_0^1 exp(-u t^2) dt.  It is one
//   polynomial in x = s xs - 1 over s in [0, rc^2], fitted on the host (Chebyshev interpolant
//   converted to monomials in long double) and evaluated by Horner's rule.  The subtraction
//   from 1/r loses nothing absolute: the error of a term is a few ulps of kC q / r.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace aleator_gpu {

template <int NR>
__device__ inline double rsq_refined(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const double e = fma(-hx, y * y, 0.5);  // 1/2 (1 - x y^2)
        y = fma(y, e, y);
    }
    return y;
}

template <int NR>
__device__ inline double rcp_refined(double x) {
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const double e = fma(-x, y, 1.0);
        y = fma(y, e, y);
    }
    return y;
}

// erf(alpha r)/r from s = r^2: Horner's rule over the DEG + 1 coefficients c (lowest first) in
// x = s xs + xo (xo = -1 for a fit over r in [0, rmax]).
template <int DEG>
__device__ inline double erf_over_r_poly(double s, const double (&c)[DEG + 1], double xs, double xo = -1.0) {
    const double x = fma(s, xs, xo);
    double p = c[DEG];
#pragma unroll
    for (int k = DEG - 1; k >= 0; --k) p = fma(p, x, c[k]);
    return p;
}

}  // namespace aleator_gpu
