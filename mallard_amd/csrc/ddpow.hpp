// pow(x, y) evaluated as exp(y log x) in double-double arithmetic
// (Dekker / Knuth error-free sums and FMA products) with a 16-entry
// 2^(i/16) reduction for both the logarithm and the exponential. The result
// carries a relative error of a few 2^-70 before the final rounding, so it is
// the correctly rounded power except in rare near-halfway cases; glibc's pow
// has the same property, so the two agree on nearly every argument. Special
// operands (zero, negative or non-finite x, infinite or zero y, NaN) follow
// C99 Annex F.9.4.4 case by case, with no call to the platform pow: on the
// device its inlined body would add ~36 VGPRs to every kernel that reaches
// this function. Results below the normal range are rounded twice (once to
// 53 bits, once by ldexp), so they can differ from the correctly rounded
// subnormal by one unit (2^-1074).
//
// The same source compiles for the host (g++) and the device (hipcc).
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define MLR_HD __host__ __device__
#else
#define MLR_HD
#endif
#define MLR_CONST static constexpr

#include "ddpow_tables.hpp"

namespace mlr::ddpow {

struct DD {
    double hi, lo;
};

MLR_HD inline DD two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
MLR_HD inline DD fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
MLR_HD inline DD two_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}

// ln(x) for normal x > 0 as a double-double.
MLR_HD inline DD log_dd(double x) {
    int k;
    double m = frexp(x, &k);  // [0.5, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        k -= 1;
    }
    // nearest j with c = 2^(j/16), j in [-8, 8]
    int j = -8;
    for (int i = 0; i < 16; ++i) j += (m >= kThr[i]) ? 1 : 0;
    const int idx = j >= 0 ? j : j + 16;
    double chi = kExp2Hi[idx & 15], clo = kExp2Lo[idx & 15];
    if (j < 0) { chi *= 0.5; clo *= 0.5; }
    // s = (m - c) / (m + c) as a double-double; ln(m / c) = 2 atanh(s)
    const double n0 = m - chi;  // exact (Sterbenz)
    const DD num = two_sum(n0, -clo);
    DD den = two_sum(m, chi);
    den = fast_two_sum(den.hi, den.lo + clo);
    const double shi = num.hi / den.hi;
    const double r = (fma(-shi, den.hi, num.hi) + num.lo) - shi * den.lo;
    const double slo = r / den.hi;
    const double z = shi * shi;
    // 2 s^3 (1/3 + z/5 + z^2/7 + z^3/9 + z^4/11), |s| < 0.011
    const double tail = shi * z * (2.0 / 3 + z * (2.0 / 5 + z * (2.0 / 7 + z * (2.0 / 9 + z * (2.0 / 11)))));
    const DD lm = fast_two_sum(2.0 * shi, 2.0 * slo + tail);
    // (16 k + j) ln2/16
    const double kk = (double)(16 * k + j);
    const DD kl = two_prod(kk, kLn2_16Hi);
    const double kl_lo = kk * kLn2_16Lo;
    DD s = two_sum(kl.hi, lm.hi);
    return fast_two_sum(s.hi, s.lo + (kl.lo + kl_lo + lm.lo));
}

// exp(p) for a double-double p with |p.hi| < 708.
MLR_HD inline double exp_dd(DD p) {
    const double n = rint(p.hi * (16.0 / 0.69314718055994530942));
    const DD q = two_prod(n, kLn2_16Hi);
    const double rh0 = p.hi - q.hi;  // exact for n != 0 (Sterbenz); p.hi when n = 0
    const double rl0 = (p.lo - q.lo) - n * kLn2_16Lo;
    const DD r = fast_two_sum(rh0, rl0);
    // exp(r) - 1 = r + r^2/2 + r^3 (1/6 + r/24 + ...), |r| < 0.0218
    const DD sq = two_prod(r.hi, r.hi);
    const double h = r.hi;
    const double cub = h * sq.hi *
                       (1.0 / 6 + h * (1.0 / 24 + h * (1.0 / 120 + h * (1.0 / 720 + h * (1.0 / 5040 + h * (1.0 / 40320))))));
    DD e = two_sum(h, 0.5 * sq.hi);
    e = fast_two_sum(e.hi, e.lo + (r.lo + r.lo * h + 0.5 * sq.lo + cub));
    const int ni = (int)n;
    const int i = ni & 15;
    const int sc = (ni - i) / 16;
    const double th = kExp2Hi[i], tl = kExp2Lo[i];
    // 2^(i/16) (1 + e)
    const DD a = two_prod(th, e.hi);
    const DD s = two_sum(th, a.hi);
    const double res = s.hi + (s.lo + (a.lo + tl + th * e.lo + tl * e.hi));
    return ldexp(res, sc);
}

// finite y is an integer / an odd integer (0.5 y is exact; every double
// of magnitude >= 2^53 is even)
MLR_HD inline bool is_int(double y) { return floor(y) == y; }
MLR_HD inline bool is_odd(double y) { return is_int(y) && !is_int(0.5 * y); }

MLR_HD inline double pow_dd(double x, double y) {
    if (y == 0.0 || x == 1.0) return 1.0;   // even for a NaN operand
    if (x != x || y != y) return x + y;     // NaN
    const double ax = fabs(x);
    if (fabs(y) == INFINITY) {
        if (ax == 1.0) return 1.0;          // x = -1
        return (ax < 1.0) == (y > 0.0) ? 0.0 : INFINITY;
    }
    const bool odd = is_odd(y);
    if (x == 0.0 || ax == INFINITY) {
        // 0^y and inf^y: infinite when exactly one of (x = 0, y < 0) holds;
        // the sign of x survives for odd integer y
        const double r = (x == 0.0) == (y < 0.0) ? INFINITY : 0.0;
        return (std::signbit(x) && odd) ? -r : r;
    }
    if (x < 0.0 && !is_int(y)) return __builtin_nan("");
    // |x| finite, positive (subnormals included: frexp normalises them)
    const DD L = log_dd(ax);
    DD p = two_prod(y, L.hi);
    p = fast_two_sum(p.hi, p.lo + y * L.lo);
    double r;
    if (p.hi > 709.8) r = INFINITY;         // above ln(DBL_MAX) = 709.78...
    else if (p.hi < -745.2) r = 0.0;        // below ln(2^-1075) = -745.13...
    else r = exp_dd(p);
    return (x < 0.0 && odd) ? -r : r;
}

}  // namespace mlr::ddpow
