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

#if defined(__GNUC__) && !defined(__clang__)
#pragma GCC diagnostic ignored "-Wunknown-pragmas"
#endif

#if defined(__HIPCC__)
#define GH_HD __host__ __device__
#else
#define GH_HD
#endif

namespace glass {

// A polynomial coefficient held in a scalar register pair.  gfx9 VALU
// instructions cannot take a 64-bit literal, so a coefficient used as the
// addend of a Horner step is otherwise rebuilt in vector registers (two
// moves per step, as costly as the fma itself); from a scalar pair the step
// is one v_fma_f64.  The value is unchanged.
GH_HD inline double kc(double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(c));
#endif
    return c;
}

// Taylor coefficients (-1)^n / (2n+1)!  and  (-1)^n / (2n)!
// sin r = r + r^3 Q(r^2),  Q(z) = -1/3! + z/5! - z^2/7! + ... + z^7/17!
GH_HD inline double poly_sin(double r)
{
    const double z = r * r;
    double p = 1.0 / 355687428096000.0;             // 1/17!
    p = fma(p, z, kc(-1.0 / 1307674368000.0));          // -1/15!
    p = fma(p, z, kc(1.0 / 6227020800.0));              // 1/13!
    p = fma(p, z, kc(-1.0 / 39916800.0));               // -1/11!
    p = fma(p, z, kc(1.0 / 362880.0));                  // 1/9!
    p = fma(p, z, kc(-1.0 / 5040.0));                   // -1/7!
    p = fma(p, z, kc(1.0 / 120.0));                     // 1/5!
    p = fma(p, z, kc(-1.0 / 6.0));                      // -1/3!
    return fma(r * z, p, r);
}

// cos r = 1 - r^2/2 + r^4 R(r^2),  R(z) = 1/4! - z/6! + ... - z^7/18!
GH_HD inline double poly_cos(double r)
{
    const double z = r * r;
    double p = -1.0 / 6402373705728000.0;           // -1/18!
    p = fma(p, z, kc(1.0 / 20922789888000.0));          // 1/16!
    p = fma(p, z, kc(-1.0 / 87178291200.0));            // -1/14!
    p = fma(p, z, kc(1.0 / 479001600.0));               // 1/12!
    p = fma(p, z, kc(-1.0 / 3628800.0));                // -1/10!
    p = fma(p, z, kc(1.0 / 40320.0));                   // 1/8!
    p = fma(p, z, kc(-1.0 / 720.0));                    // -1/6!
    p = fma(p, z, kc(1.0 / 24.0));                      // 1/4!
    const double h = 0.5 * z;
    const double w = 1.0 - h;
    // (1 - w) - h recovers the rounding of 1 - r^2/2
    return w + (((1.0 - w) - h) + z * z * p);
}

// Flips the sign of x when bit 31 of m is set (an exact negation).
GH_HD inline double flip_sign(double x, unsigned m)
{
    unsigned long long b;
    __builtin_memcpy(&b, &x, sizeof b);
    b ^= (unsigned long long)(m & 0x80000000u) << 32;
    __builtin_memcpy(&x, &b, sizeof b);
    return x;
}

// Without branches: arguments outside the domain are reduced as 0 and the
// result replaced by NaN at the end; the quadrant rotation is a swap by
// selects plus sign flips of the sign bits (bit-identical to negation).
GH_HD inline void sincos_red(double x, double* s, double* c)
{
    const bool bad = !(fabs(x) < 536870912.0);  // outside |x| < 2^29 (or NaN): no result
    const double xs = bad ? 0.0 : x;
    const double TWO_OVER_PI = 0.63661977236758134308;
    const double C1 = 1.5707963267948966;          // pi/2 rounded
    const double C2 = 6.123233995736766e-17;       // pi/2 - C1
    const double C3 = -1.4973849048591698e-33;     // pi/2 - C1 - C2
    const double k = rint(xs * TWO_OVER_PI);
    double r = fma(-k, C1, xs);
    r = fma(-k, C2, r);
    r = fma(-k, C3, r);
    const double ps = poly_sin(r), pc = poly_cos(r);
    // x = k pi/2 + r, q = k mod 4: (sin, cos) = (ps, pc), (pc, -ps),
    // (-ps, -pc), (-pc, ps)
    const int q = (int)k;
    const bool odd = (q & 1) != 0;
    const double ss = odd ? pc : ps, cc = odd ? ps : pc;
    const double so = flip_sign(ss, (unsigned)(q & 2) << 30);
    const double co = flip_sign(cc, (unsigned)((q ^ (q >> 1)) & 1) << 31);
    *s = bad ? __builtin_nan("") : so;
    *c = bad ? __builtin_nan("") : co;
}

// a / b correctly rounded from y = 1/b correctly rounded: q = a y is within
// one ulp of a/b, e = a - b q is exact (fma), q + e y rounds to a/b
// (Markstein's theorem; no underflow or overflow).  Three dependent
// operations instead of a division's chain of ten.
GH_HD inline double div_by_recip(double a, double b, double y)
{
    const double q = a * y;
    return fma(fma(-q, b, a), y, q);
}

// Double-double values hi + lo (|lo| <= ulp(hi)/2), passed by value.  The
// error-free sum / product and everything built on them keep contraction
// off: fusing a product into a later sum (across inlined calls) would drop
// the rounding error they track.
struct dd_t
{
    double h, l;
};
struct sc_t
{
    double s, c;
};
GH_HD inline dd_t dd_sum(double a, double b)   // two_sum as a value
{
#pragma clang fp contract(off)
    const double t = a + b;
    const double bb = t - a;
    return dd_t{t, (a - (t - bb)) + (b - bb)};
}
GH_HD inline dd_t dd_prod(double a, double b)  // two_prod as a value
{
#pragma clang fp contract(off)
    const double t = a * b;
    return dd_t{t, fma(a, b, -t)};
}
GH_HD inline dd_t dd_add(dd_t a, dd_t b)
{
#pragma clang fp contract(off)
    dd_t s = dd_sum(a.h, b.h);
    return dd_sum(s.h, s.l + (a.l + b.l));
}
GH_HD inline dd_t dd_mul(dd_t a, dd_t b)
{
#pragma clang fp contract(off)
    dd_t p = dd_prod(a.h, b.h);
    return dd_sum(p.h, p.l + (a.h * b.l + a.l * b.h));
}
GH_HD inline dd_t dd_div_small(dd_t a, double d)
{
#pragma clang fp contract(off)
    // (a.h + a.l) / d for a small integer d (exact remainder by fma)
    const double q = a.h / d;
    return dd_t{q, (fma(-q, d, a.h) + a.l) / d};
}
GH_HD inline dd_t dd_neg(dd_t a) { return dd_t{-a.h, -a.l}; }

// sin and cos rounded to nearest except in rare cases (~1e-5 of arguments
// lie closer to a rounding boundary than the evaluation error), for
// |x| < 2^29: the argument is reduced exactly to r in [-pi/8, pi/8] by
// octants in double-double (Dekker product k*P1, Sterbenz subtraction,
// three-part pi/4), sin r and cos r are summed in double-double up to the
// r^5 / z^3 terms, the remaining Taylor terms (below 1e-6 of the result) in
// double, and odd octants are rotated by sqrt(1/2) in double-double.  Used
// where a result must round as a correctly rounded libm does (phases whose
// one-ulp differences would be amplified downstream).  Returns (sin, cos).
GH_HD inline sc_t sincos_cr2(double x)
{
#pragma clang fp contract(off)
    if (!(fabs(x) < 536870912.0))
    {
        const double nan = __builtin_nan("");
        return sc_t{nan, nan};
    }
    const double FOUR_OVER_PI = 1.2732395447351628;
    const double P1 = 0.78539816339744828, P2 = 3.061616997868383e-17, P3 = -7.486924524295849e-34;
    const double k = rint(x * FOUR_OVER_PI);
    const dd_t p1 = dd_prod(k, P1);
    dd_t r = dd_sum(x - p1.h, -p1.l);          // x - p1.h is exact (Sterbenz)
    const dd_t p2 = dd_prod(k, P2);
    {
        const dd_t t = dd_sum(r.h, -p2.h);
        r = dd_sum(t.h, t.l + ((r.l - p2.l) - k * P3));
    }
    const dd_t z = dd_mul(r, r);
    const dd_t z2 = dd_mul(z, z);
    const dd_t z3 = dd_mul(z2, z);
    // sin r = r (1 - z/6 + z^2/120 - z^3/5040 + z^4 (1/9! - z/11! + ... + z^4/17!))
    dd_t S = dd_add(dd_t{1.0, 0.0}, dd_neg(dd_div_small(z, 6.0)));
    S = dd_add(S, dd_div_small(z2, 120.0));
    {
        double q = fma(z.h, 1.0 / 355687428096000.0, -1.0 / 1307674368000.0);  // 1/17!, -1/15!
        q = fma(q, z.h, 1.0 / 6227020800.0);                                 // 1/13!
        q = fma(q, z.h, -1.0 / 39916800.0);                                  // -1/11!
        q = fma(q, z.h, 1.0 / 362880.0);                                     // 1/9!
        S.l += -z3.h / 5040.0 + z2.h * z2.h * q;
    }
    const dd_t sn = dd_mul(r, S);
    // cos r = 1 - z/2 + z^2/24 - z^3/720 + z^4 (1/8! - z/10! + ... - z^5/18!)
    dd_t C = dd_add(dd_t{1.0, 0.0}, dd_t{-0.5 * z.h, -0.5 * z.l});
    C = dd_add(C, dd_div_small(z2, 24.0));
    C = dd_add(C, dd_neg(dd_div_small(z3, 720.0)));
    {
        double q = fma(z.h, -1.0 / 6402373705728000.0, 1.0 / 20922789888000.0);  // -1/18!, 1/16!
        q = fma(q, z.h, -1.0 / 87178291200.0);                                   // -1/14!
        q = fma(q, z.h, 1.0 / 479001600.0);                                      // 1/12!
        q = fma(q, z.h, -1.0 / 3628800.0);                                       // -1/10!
        q = fma(q, z.h, 1.0 / 40320.0);                                          // 1/8!
        C.l += z2.h * z2.h * q;
    }
    const long kl = (long)k;
    const int j = (int)(((kl % 8) + 8) % 8);
    double ss, cc;
    if (j & 1)
    {
        // pi/4 + r: sqrt(1/2) (cos r + sin r), sqrt(1/2) (cos r - sin r)
        const dd_t R{0.70710678118654757, -4.8336466567264567e-17};
        const dd_t a = dd_mul(dd_add(C, sn), R);
        const dd_t b = dd_mul(dd_add(C, dd_neg(sn)), R);
        ss = a.h + a.l;
        cc = b.h + b.l;
    }
    else
    {
        ss = sn.h + sn.l;
        cc = C.h + C.l;
    }
    // rotate by the quadrant q = j/2: (s,c) -> (c,-s) -> (-s,-c) -> (-c,s)
    const int q = j >> 1;
    double so = (q & 1) ? cc : ss;
    double co = (q & 1) ? ss : cc;
    if (q >= 2) so = -so;
    if (q == 1 || q == 2) co = -co;
    return sc_t{so, co};
}

GH_HD inline void sincos_cr(double x, double* s, double* c)
{
    const sc_t v = sincos_cr2(x);
    *s = v.s;
    *c = v.c;
}

// atan(k/8), k = 0..8, and 2^(j/64), j = 0..63, as double-double (values
// computed at 200-bit precision for this library).
constexpr dd_t kAtanTab[9] = {
    {0.0, 0.0},
    {0.12435499454676144, -3.1253241424539383e-18},
    {0.24497866312686414, 1.0698755618734451e-17},
    {0.35877067027057225, -2.4623815582638635e-17},
    {0.4636476090008061, 2.2698777452961687e-17},
    {0.5585993153435624, -5.4556305485916264e-18},
    {0.6435011087932844, 1.5834785051444286e-17},
    {0.7188299996216245, -2.1478388444456983e-17},
    {0.7853981633974483, 3.061616997868383e-17},
};
constexpr dd_t kExp2Tab[64] = {
    {1.0, 0.0}, {1.0108892860517005, -1.5234778603368577e-17},
    {1.0218971486541166, 5.109225028973444e-17}, {1.0330248790212284, 7.600838874027088e-18},
    {1.0442737824274138, 8.551889705537965e-17}, {1.0556451783605572, 1.759325738772092e-18},
    {1.0671404006768237, -7.899853966841582e-17}, {1.0787607977571199, -6.656660436056593e-17},
    {1.0905077326652577, -3.046782079812471e-17}, {1.102382583307841, 5.2660368715706944e-17},
    {1.1143867425958924, 1.0410278456845571e-16}, {1.1265216186082418, 5.165856758795457e-17},
    {1.1387886347566916, 8.912812676025408e-17}, {1.1511892299529827, 3.250710218863827e-17},
    {1.1637248587775775, 3.8292048369240935e-17}, {1.1763969916502812, 5.554203254218079e-17},
    {1.189207115002721, 3.982015231465646e-17}, {1.202156731452703, 6.644981499252301e-17},
    {1.215247359980469, -7.712630692681488e-17}, {1.22848053610687, -1.89878163130253e-17},
    {1.241857812073484, 4.658027591836937e-17}, {1.255380757024691, -6.7113898212968784e-18},
    {1.2690509571917332, 2.667932131342186e-18}, {1.2828700160787783, 1.713594918243561e-17},
    {1.2968395546510096, 2.5382502794888315e-17}, {1.3109612115247644, -7.181536135519454e-17},
    {1.3252366431597413, -2.8587312100388614e-17}, {1.339667524053303, 8.927282594831732e-17},
    {1.3542555469368927, 7.70094837980299e-17}, {1.3690024229745905, 9.593797919118849e-17},
    {1.383909881963832, -6.770511658794786e-17}, {1.3989796725383112, -9.614213209051323e-17},
    {1.4142135623730951, -9.667293313452913e-17}, {1.42961333839197, -1.2031642489053655e-17},
    {1.4451808069770467, -3.0237581349939873e-17}, {1.460917794180647, -5.600377186075216e-17},
    {1.4768261459394993, -3.483994556892796e-17}, {1.4929077282912648, 1.4192920154284036e-17},
    {1.5091644275934228, -1.016455327754295e-16}, {1.5255981507445384, -1.1024941712342561e-16},
    {1.5422108254079407, 7.949834809697621e-17}, {1.559004400237837, 3.7812070533575275e-17},
    {1.5759808451078865, -1.0136916471278304e-17}, {1.593142151342267, -1.0094406542311964e-16},
    {1.6104903319492543, 2.4707192569797888e-17}, {1.6280274218573478, -6.712955084707084e-17},
    {1.645755478153965, -1.0125679913674773e-16}, {1.6636765803267364, 5.8909926967131e-17},
    {1.681792830507429, 8.199010020581497e-17}, {1.7001063537185235, -8.0237193703977e-18},
    {1.718619298122478, -1.851380418263111e-17}, {1.7373338352737062, 3.164389299292957e-17},
    {1.7562521603732995, 2.960140695448873e-17}, {1.7753764925265212, 6.429731796556572e-17},
    {1.7947090750031072, 1.8227458427912087e-17}, {1.8142521755003989, -9.969531538920349e-17},
    {1.8340080864093424, 3.283107224245627e-17}, {1.8539791250833855, 9.761887490727594e-17},
    {1.8741676341103, -6.122763413004143e-17}, {1.8945759815869656, 3.4034035352165297e-17},
    {1.9152065613971474, -1.0619946056195963e-16}, {1.9360617934922943, 1.0332385960676326e-16},
    {1.9571441241754002, 8.960767791036668e-17}, {1.978456026387951, 4.0388753109278167e-17},
};

// (a.h + a.l) / (b.h + b.l) to ~2^-100 relative (one correction step).
GH_HD inline dd_t dd_div(dd_t a, dd_t b)
{
#pragma clang fp contract(off)
    const double q1 = a.h / b.h;
    const dd_t r = dd_add(a, dd_neg(dd_mul(dd_t{q1, 0.0}, b)));
    const double q2 = r.h / b.h;
    const dd_t r2 = dd_add(r, dd_neg(dd_mul(dd_t{q2, 0.0}, b)));
    const double q3 = r2.h / b.h;
    const dd_t q = dd_sum(q1, q2);
    return dd_sum(q.h, q.l + q3);
}

// atan2(y, x) rounded to nearest except in rare cases, for the phase of a
// complex response: t = min/max of |x|, |y| in double-double, reduced by the
// nearest c = k/8 through atan t = atan c + atan((t - c)/(1 + t c)) with
// atan(k/8) tabulated in double-double (values computed at 200-bit
// precision), the remaining |u| <= 1/16 by its Taylor series (u^3/3 in
// double-double, the rest in double), then the octant and sign of (x, y).
// Non-finite or (0, 0) arguments go to the platform atan2.
GH_HD inline double atan2_cr(double y, double x)
{
#pragma clang fp contract(off)
    if (!(isfinite(x) && isfinite(y)) || (x == 0.0 && y == 0.0)) return atan2(y, x);
    const dd_t PI_2{1.5707963267948966, 6.123233995736766e-17};
    const dd_t PI{3.141592653589793, 1.2246467991473532e-16};
    const double ax = fabs(x), ay = fabs(y);
    const bool swap = ay > ax;
    const dd_t t = dd_div(dd_t{swap ? ax : ay, 0.0}, dd_t{swap ? ay : ax, 0.0});   // in [0, 1]
    const int k = (int)rint(t.h * 8.0);
    const double c = k * 0.125;
    const dd_t u = dd_div(dd_add(t, dd_t{-c, 0.0}), dd_add(dd_t{1.0, 0.0}, dd_mul(t, dd_t{c, 0.0})));
    const dd_t w = dd_mul(u, u);
    dd_t S = dd_add(dd_t{1.0, 0.0}, dd_neg(dd_div_small(w, 3.0)));
    {
        double q = fma(w.h, -1.0 / 15.0, 1.0 / 13.0);
        q = fma(q, w.h, -1.0 / 11.0);
        q = fma(q, w.h, 1.0 / 9.0);
        q = fma(q, w.h, -1.0 / 7.0);
        q = fma(q, w.h, 1.0 / 5.0);
        S.l += w.h * w.h * q;
    }
    dd_t a = dd_add(kAtanTab[k], dd_mul(u, S));
    if (swap) a = dd_add(PI_2, dd_neg(a));
    if (x < 0.0) a = dd_add(PI, dd_neg(a));
    const double r = a.h + a.l;
    return signbit(y) ? -r : r;
}

// exp(x) rounded to nearest except in rare cases: x = (64 m + j) ln2/64 + r
// with r in double-double (three-part ln2/64, Dekker product, Sterbenz
// subtraction), exp r = 1 + r + r^2/2 in double-double plus the r^3..r^6
// terms (|r| <= ln2/128), times 2^(j/64) in double-double, scaled by 2^m.
// |x| >= 700 goes to the platform exp.
GH_HD inline double exp_cr(double x)
{
#pragma clang fp contract(off)
    if (!(fabs(x) < 700.0)) return exp(x);
    const double k = rint(x * 92.33248261689366);   // 64 / ln 2
    const double L1 = 0.010830424696249145, L2 = 3.623510646634843e-19, L3 = 8.918294435025331e-36;
    const dd_t p1 = dd_prod(k, L1);
    dd_t r = dd_sum(x - p1.h, -p1.l);                // x - p1.h is exact (Sterbenz)
    const dd_t p2 = dd_prod(k, L2);
    {
        const dd_t t = dd_sum(r.h, -p2.h);
        r = dd_sum(t.h, t.l + ((r.l - p2.l) - k * L3));
    }
    dd_t E = dd_add(dd_t{1.0, 0.0}, r);
    const dd_t r2 = dd_mul(r, r);
    E = dd_add(E, dd_t{0.5 * r2.h, 0.5 * r2.l});
    {
        const double h = r.h;
        E.l += h * h * h * (1.0 / 6.0 + h * (1.0 / 24.0 + h * (1.0 / 120.0 + h * (1.0 / 720.0))));
    }
    const long ki = (long)k;
    const long m = ki >> 6;                          // floor(k / 64)
    const int j = (int)(ki & 63);
    const dd_t R = dd_mul(kExp2Tab[j], E);
    return ldexp(R.h + R.l, (int)m);
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
    p = fma(p, x2, kc(-1.0 / 25852016738884976640000.0)); // -1/23!
    p = fma(p, x2, kc(1.0 / 51090942171709440000.0));     // 1/21!
    p = fma(p, x2, kc(-1.0 / 121645100408832000.0));      // -1/19!
    p = fma(p, x2, kc(1.0 / 355687428096000.0));          // 1/17!
    p = fma(p, x2, kc(-1.0 / 1307674368000.0));           // -1/15!
    p = fma(p, x2, kc(1.0 / 6227020800.0));               // 1/13!
    p = fma(p, x2, kc(-1.0 / 39916800.0));                // -1/11!
    p = fma(p, x2, kc(1.0 / 362880.0));                   // 1/9!
    p = fma(p, x2, kc(-1.0 / 5040.0));                    // -1/7!
    p = fma(p, x2, kc(1.0 / 120.0));                      // 1/5!
    p = fma(p, x2, kc(-1.0 / 6.0));                       // -1/3!
    return fma(p, x2, 1.0);
}

}  // namespace glass
