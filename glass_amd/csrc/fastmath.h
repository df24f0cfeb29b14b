// fastmath.h -- FP64 sin/cos for the UCB phase arguments, written for this
// library from textbook methods (host and device):
//
//  * sincos_red(x): Cody-Waite argument reduction x = k*(pi/2) + r with a
//    three-part pi/2 (C1 = pi/2 rounded to double, C2 = the rounding error
//    of C1, C3 = the rounding error of C1 + C2), each product k*Ci folded
//    with one fused multiply-add so the reduction error stays ~1e-16 for
//    |k| < 2^30; then Taylor polynomials of sin and cos on |r| <= pi/4
//    (terms through r^17 / r^18, truncation < 1e-18) and the quadrant
//    swap.  Domain |x| < 2^29 (the slow-series phases reach ~2e7 rad);
//    outside it the result is NaN, so a misuse poisons the likelihood
//    instead of passing silently.  No platform libm call is inlined, which
//    keeps the register footprint of the callers small.
//  * sinc_ratio(x) = sin(x)/x by its Taylor series in x^2 through x^24
//    (|x| <= 1.75, truncation < 1e-19); larger |x| use sincos_red.
//
// Accuracy is checked against glibc on the CPU (tests/test_fastmath.py):
// a few ulp, far inside the 1e-12 waveform budget.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define GH_HD __host__ __device__
#else
#define GH_HD
#endif

namespace glass {

// Taylor coefficients (-1)^n / (2n+1)!  and  (-1)^n / (2n)!
// sin r = r + r^3 Q(r^2),  Q(z) = -1/3! + z/5! - z^2/7! + ... + z^7/17!
GH_HD inline double poly_sin(double r)
{
    const double z = r * r;
    double p = 1.0 / 355687428096000.0;             // 1/17!
    p = fma(p, z, -1.0 / 1307674368000.0);          // -1/15!
    p = fma(p, z, 1.0 / 6227020800.0);              // 1/13!
    p = fma(p, z, -1.0 / 39916800.0);               // -1/11!
    p = fma(p, z, 1.0 / 362880.0);                  // 1/9!
    p = fma(p, z, -1.0 / 5040.0);                   // -1/7!
    p = fma(p, z, 1.0 / 120.0);                     // 1/5!
    p = fma(p, z, -1.0 / 6.0);                      // -1/3!
    return fma(r * z, p, r);
}

// cos r = 1 - r^2/2 + r^4 R(r^2),  R(z) = 1/4! - z/6! + ... - z^7/18!
GH_HD inline double poly_cos(double r)
{
    const double z = r * r;
    double p = -1.0 / 6402373705728000.0;           // -1/18!
    p = fma(p, z, 1.0 / 20922789888000.0);          // 1/16!
    p = fma(p, z, -1.0 / 87178291200.0);            // -1/14!
    p = fma(p, z, 1.0 / 479001600.0);               // 1/12!
    p = fma(p, z, -1.0 / 3628800.0);                // -1/10!
    p = fma(p, z, 1.0 / 40320.0);                   // 1/8!
    p = fma(p, z, -1.0 / 720.0);                    // -1/6!
    p = fma(p, z, 1.0 / 24.0);                      // 1/4!
    const double h = 0.5 * z;
    const double w = 1.0 - h;
    // (1 - w) - h recovers the rounding of 1 - r^2/2
    return w + (((1.0 - w) - h) + z * z * p);
}

GH_HD inline void sincos_red(double x, double* s, double* c)
{
    if (!(fabs(x) < 536870912.0))  // outside |x| < 2^29 (or NaN): no result
    {
        *s = *c = __builtin_nan("");
        return;
    }
    const double TWO_OVER_PI = 0.63661977236758134308;
    const double C1 = 1.5707963267948966;          // pi/2 rounded
    const double C2 = 6.123233995736766e-17;       // pi/2 - C1
    const double C3 = -1.4973849048591698e-33;     // pi/2 - C1 - C2
    const double k = rint(x * TWO_OVER_PI);
    double r = fma(-k, C1, x);
    r = fma(-k, C2, r);
    r = fma(-k, C3, r);
    const double ps = poly_sin(r), pc = poly_cos(r);
    const long q = (long)k & 3;
    // x = k pi/2 + r: quadrant rotation of (sin r, cos r)
    double ss, cc;
    if (q == 0) { ss = ps; cc = pc; }
    else if (q == 1) { ss = pc; cc = -ps; }
    else if (q == 2) { ss = -ps; cc = -pc; }
    else { ss = -pc; cc = ps; }
    *s = ss;
    *c = cc;
}

// sin(x)/x, even Taylor series
GH_HD inline double sinc_ratio(double x)
{
    const double x2 = x * x;
    if (!(x2 <= 3.0625))  // |x| > 1.75
    {
        double s, c;
        sincos_red(x, &s, &c);
        return s / x;
    }
    double p = 1.0 / 15511210043330985984000000.0;  // 1/25!
    p = fma(p, x2, -1.0 / 25852016738884976640000.0); // -1/23!
    p = fma(p, x2, 1.0 / 51090942171709440000.0);     // 1/21!
    p = fma(p, x2, -1.0 / 121645100408832000.0);      // -1/19!
    p = fma(p, x2, 1.0 / 355687428096000.0);          // 1/17!
    p = fma(p, x2, -1.0 / 1307674368000.0);           // -1/15!
    p = fma(p, x2, 1.0 / 6227020800.0);               // 1/13!
    p = fma(p, x2, -1.0 / 39916800.0);                // -1/11!
    p = fma(p, x2, 1.0 / 362880.0);                   // 1/9!
    p = fma(p, x2, -1.0 / 5040.0);                    // -1/7!
    p = fma(p, x2, 1.0 / 120.0);                      // 1/5!
    p = fma(p, x2, -1.0 / 6.0);                       // -1/3!
    return fma(p, x2, 1.0);
}

}  // namespace glass
