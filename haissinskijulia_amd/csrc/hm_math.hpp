// hm_math.hpp -- FP64 elementary functions of the per-particle pass: sin/cos
// by Cody-Waite reduction to |r| <= pi/4, log through atanh of
// s = (m-1)/(m+1) on the mantissa m in [sqrt(1/2), sqrt(2)).  The polynomial
// parts are interpolants at Chebyshev nodes (near-minimax; the method of
// Chebyshev economization, e.g. Muller, "Elementary Functions", ch. 3), fitted
// in 60-digit arithmetic to the remainder after the leading Taylor terms;
// their truncation error is below 1e-19 on the reduced interval, so each needs
// fewer terms than the Taylor series.  Results are within a few ulp of the
// correctly rounded values (tests/test_math.py runs this same code compiled for
// the CPU against numpy).
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

// sin r, |r| <= pi/4: r - r^3 p(r^2), p(t) ~ (r - sin r)/r^3 = 1/3! - t/5! + ...
// (degree 6 in t, truncation error 6e-21 at pi/4)
HM_QUAL double sin_poly(double r) {
    const double r2 = r * r;
    double p = 7.586697117706917984663e-13;
    p = HM_FMAK(p, r2, -1.605853161898614683644e-10);
    p = HM_FMAK(p, r2, 2.505210623244757732801e-8);
    p = HM_FMAK(p, r2, -2.755731921933916651755e-6);
    p = HM_FMAK(p, r2, 1.984126984126506494653e-4);
    p = HM_FMAK(p, r2, -8.333333333333331492084e-3);
    p = HM_FMAK(p, r2, 1.666666666666666666551e-1);
    return __builtin_fma(-(p * r2), r, r);
}

// cos r, |r| <= pi/4: 1 - r^2/2 + r^4 p(r^2), p(t) ~ (cos r - 1 + t/2)/t^2 =
// 1/4! - t/6! + ... (degree 5 in t, truncation error 5e-19 at pi/4)
HM_QUAL double cos_poly(double r) {
    const double r2 = r * r;
    double p = -1.138263242552171800541e-11;
    p = HM_FMAK(p, r2, 2.087614626840319789335e-9);
    p = HM_FMAK(p, r2, -2.755731727172979288199e-7);
    p = HM_FMAK(p, r2, 2.480158729876568792727e-5);
    p = HM_FMAK(p, r2, -1.388888888888739723669e-3);
    p = HM_FMAK(p, r2, 4.166666666666666538875e-2);
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
    // log m = 2 atanh s = 2 s (1 + s^2 p(s^2)), p(t) ~ (atanh(s)/s - 1)/t =
    // 1/3 + t/5 + ... (degree 7 in t, truncation error 1e-18 at |s| = 0.1716)
    double p = 0.06543313073920051069031;
    p = HM_FMAK(p, s2, 0.06634386569328442592349);
    p = HM_FMAK(p, s2, 0.0769311985140732909423);
    p = HM_FMAK(p, s2, 0.09090897820066453138014);
    p = HM_FMAK(p, s2, 0.1111111119585695872093);
    p = HM_FMAK(p, s2, 0.1428571428540180590539);
    p = HM_FMAK(p, s2, 0.2000000000000043849136);
    p = HM_FMAK(p, s2, 0.3333333333333333323247);
    const double t = s + s;
    const double lm = __builtin_fma(t * s2, p, t);
    const double de = (double)e;
    return __builtin_fma(de, kLn2Hi, __builtin_fma(de, kLn2Lo, lm));
}

}  // namespace fm
}  // namespace hm
