// hm_math.hpp -- FP64 elementary functions of the per-particle pass, written
// from their series: sin/cos by Cody-Waite reduction to |r| <= pi/4 and
// Taylor polynomials (truncation error below 1e-19), log by the atanh series
// of s = (m-1)/(m+1) on the mantissa m in [sqrt(1/2), sqrt(2)).  Results are
// within a few ulp of the correctly rounded values (tests/test_math.py runs
// this same code compiled for the CPU against numpy).
//
// The includer defines HM_QUAL (the function qualifier) and an Ops policy
// with rcp(a) (an estimate of 1/a), frexp_mant(u) in [0.5, 1) and
// frexp_exp(u), so one source serves the gfx950 kernels and the host test.
// HM_FMAK(a, b, k) is a*b + k with one rounding for a constant k; the
// kernels define it to keep k in scalar registers.
#pragma once
#include <stdint.h>

#ifndef HM_QUAL
#define HM_QUAL inline
#endif
#ifndef HM_FMAK
#define HM_FMAK(a, b, k) __builtin_fma(a, b, k)
#endif

namespace hm {
namespace fm {

// pi/2 = Pio2_1 + Pio2_2 (+ 2^-109 residue); with FMA, x - k*Pio2_1 is exact
// for |x| >= 1, so the reduced argument is good to ~1e-17 for |k| < 2^19
constexpr double kPio2_1 = 0x1.921fb54442d18p+0;
constexpr double kPio2_2 = 0x1.1a62633145c07p-54;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kReduceMax = 524288.0;   // 2^19
// ln 2 = kLn2Hi + kLn2Lo
constexpr double kLn2Hi = 0x1.62e42fefa39efp-1;
constexpr double kLn2Lo = 0x1.abc9e3b39803fp-56;
constexpr double kSqrtHalf = 0.70710678118654752440;

// sin r, |r| <= pi/4: r - r^3 (1/3! - r^2/5! + ... - r^14/17!)
HM_QUAL double sin_poly(double r) {
    const double r2 = r * r;
    double p = -1.0 / 355687428096000.0;               // -1/17!
    p = HM_FMAK(p, r2,1.0 / 1307674368000.0);   // +1/15!
    p = HM_FMAK(p, r2,-1.0 / 6227020800.0);     // -1/13!
    p = HM_FMAK(p, r2,1.0 / 39916800.0);        // +1/11!
    p = HM_FMAK(p, r2,-1.0 / 362880.0);         // -1/9!
    p = HM_FMAK(p, r2,1.0 / 5040.0);            // +1/7!
    p = HM_FMAK(p, r2,-1.0 / 120.0);            // -1/5!
    p = HM_FMAK(p, r2,1.0 / 6.0);               // +1/3!
    return __builtin_fma(-(p * r2), r, r);
}

// cos r, |r| <= pi/4: 1 - r^2/2 + r^4 (1/4! - r^2/6! + ... - r^14/18!)
HM_QUAL double cos_poly(double r) {
    const double r2 = r * r;
    double p = -1.0 / 6402373705728000.0;              // -1/18!
    p = HM_FMAK(p, r2,1.0 / 20922789888000.0);  // +1/16!
    p = HM_FMAK(p, r2,-1.0 / 87178291200.0);    // -1/14!
    p = HM_FMAK(p, r2,1.0 / 479001600.0);       // +1/12!
    p = HM_FMAK(p, r2,-1.0 / 3628800.0);        // -1/10!
    p = HM_FMAK(p, r2,1.0 / 40320.0);           // +1/8!
    p = HM_FMAK(p, r2,-1.0 / 720.0);            // -1/6!
    p = HM_FMAK(p, r2,1.0 / 24.0);              // +1/4!
    // 1 - r^2/2 evaluated as w + (exact remainder) to keep the last bits
    const double hr2 = 0.5 * r2;
    const double w = 1.0 - hr2;
    return w + (((1.0 - w) - hr2) + (r2 * r2) * p);
}

// x = k pi/2 + r, |r| <= pi/4 (|x| < 2^19)
HM_QUAL double reduce_pio2(double x, int &q) {
    const double k = __builtin_rint(x * kTwoOverPi);
    double r = __builtin_fma(-k, kPio2_1, x);
    r = __builtin_fma(-k, kPio2_2, r);
    q = (int)k;
    return r;
}

// sin x for |x| < 2^19
HM_QUAL double sin_reduced(double x) {
    int q;
    const double r = reduce_pio2(x, q);
    const double v = (q & 1) ? cos_poly(r) : sin_poly(r);
    return (q & 2) ? -v : v;
}

// sin x and cos x for |x| < 2^19
HM_QUAL void sincos_reduced(double x, double &s, double &c) {
    int q;
    const double r = reduce_pio2(x, q);
    const double sr = sin_poly(r), cr = cos_poly(r);
    const double ss = (q & 1) ? cr : sr;
    const double cc = (q & 1) ? sr : cr;
    s = (q & 2) ? -ss : ss;
    c = ((q + 1) & 2) ? -cc : cc;
}

// 1/a from an estimate and one Newton-Raphson step (relative error ~1e-14
// from a 2^-23 estimate): for terms that enter a sum as a small correction
template <class Ops>
HM_QUAL double rcp1(double a) {
    const double y = Ops::rcp(a);
    const double e = __builtin_fma(-a, y, 1.0);
    return __builtin_fma(y, e, y);
}

// 1/a from an estimate and two Newton-Raphson steps
template <class Ops>
HM_QUAL double rcp(double a) {
    double y = Ops::rcp(a);
    double e = __builtin_fma(-a, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-a, y, 1.0);
    return __builtin_fma(y, e, y);
}

// log u for a positive normal double u
template <class Ops>
HM_QUAL double log_pos(double u) {
    int e = Ops::frexp_exp(u);        // u = m 2^e, m in [0.5, 1)
    double m = Ops::frexp_mant(u);
    if (m < kSqrtHalf) {              // m in [sqrt(1/2), sqrt(2))
        m = m + m;
        e -= 1;
    }
    const double s = (m - 1.0) * rcp<Ops>(m + 1.0);   // |s| <= 0.1716
    const double s2 = s * s;
    // log m = 2 atanh s = 2 s (1 + s^2/3 + s^4/5 + ... + s^22/23)
    double p = 1.0 / 23.0;
    p = HM_FMAK(p, s2,1.0 / 21.0);
    p = HM_FMAK(p, s2,1.0 / 19.0);
    p = HM_FMAK(p, s2,1.0 / 17.0);
    p = HM_FMAK(p, s2,1.0 / 15.0);
    p = HM_FMAK(p, s2,1.0 / 13.0);
    p = HM_FMAK(p, s2,1.0 / 11.0);
    p = HM_FMAK(p, s2,1.0 / 9.0);
    p = HM_FMAK(p, s2,1.0 / 7.0);
    p = HM_FMAK(p, s2,1.0 / 5.0);
    p = HM_FMAK(p, s2,1.0 / 3.0);
    const double t = s + s;
    const double lm = __builtin_fma(t * s2, p, t);
    const double de = (double)e;
    return __builtin_fma(de, kLn2Hi, __builtin_fma(de, kLn2Lo, lm));
}

}  // namespace fm
}  // namespace hm
