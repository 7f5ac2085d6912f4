// FP64 Bessel function J0 for the Hankel kernels.
//   x < HK_J0_XS: Chebyshev series of degree HK_J0_DEG on the unit interval holding x
//                 (coefficients fitted offline, hankel_j0_table.h);
//   x >= HK_J0_XS: Hankel asymptotic expansion
//       J0(x) = (P(x) (cos x + sin x) - Q(x) (sin x - cos x)) / sqrt(pi x),
//   which is sqrt(2/(pi x)) (P cos(x - pi/4) - Q sin(x - pi/4)) with the pi/4 shift done
//   exactly by the angle-sum identities.  The caller supplies cos x and sin x, so that kernels
//   stepping x along a uniform grid can rotate them instead of calling sincos per element.
#pragma once
#include <hip/hip_runtime.h>

#include "hankel_j0_table.h"

namespace hk {

// 0 <= x < HK_J0_XS
__device__ inline double j0_small(double x)
{
    const int n = min((int)x, HK_J0_XS - 1);
    const double t = 2.0 * (x - (double)n) - 1.0;
    const double *c = hk_j0_cheb[n];
    const double t2 = 2.0 * t;
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int k = HK_J0_DEG; k >= 1; k--) {
        const double b0 = fma(t2, b1, c[k] - b2);
        b2 = b1;
        b1 = b0;
    }
    return fma(t, b1, c[0] - b2);
}

// 1/x to full precision: hardware reciprocal estimate and two Newton steps
__device__ inline double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

__device__ inline double j0_large(double x, double cx, double sx)
{
    const double rx = recip(x);
    const double y = rx * rx;
    double P = hk_j0_p[HK_J0_NASYM - 1], Q = hk_j0_q[HK_J0_NASYM - 1];
#pragma unroll
    for (int k = HK_J0_NASYM - 2; k >= 0; k--) {
        P = fma(P, y, hk_j0_p[k]);
        Q = fma(Q, y, hk_j0_q[k]);
    }
    Q *= rx;
    const double amp = rsqrt(M_PI * x);
    return amp * (P * (cx + sx) - Q * (sx - cx));
}

// cx = cos x, sx = sin x; J0 is even (a NaN argument takes the asymptotic branch)
__device__ inline double j0(double x, double cx, double sx)
{
    if (x < 0.0) {
        x = -x;
        sx = -sx;
    }
    return x < (double)HK_J0_XS ? j0_small(x) : j0_large(x, cx, sx);
}

__device__ inline double j0(double x)
{
    x = fabs(x);
    if (x < (double)HK_J0_XS) return j0_small(x);
    double s, c;
    sincos(x, &s, &c);
    return j0_large(x, c, s);
}

} // namespace hk
