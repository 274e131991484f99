// Correctly rounded (to within a 2^-80-ish evaluation error) double log and
// cos for the Box-Muller transform of the white noise, so the Gaussians match
// a correctly rounded libm bit for bit.  Both evaluate in double-double
// arithmetic (error-free sums and FMA products) with table-driven argument
// reduction, then round the double-double result once.
//
//  log x: x = 2^e m, m in [1,2); bin j = top 8 bits of m, c_j ~ 1/m_j,
//         r = m c_j - 1 (exact as a double-double), and
//         log x = (e + k_j) ln2 - log(2^k_j c_j) + log1p(r),  |r| <= 2^-8,
//         with log1p by its Taylor series (r^2/2 and r^3/3 in double-double).
//  cos t: t = k pi/2 + s (pi/2 in three parts, ~150 bits), s = a + d with
//         a = j/64 from a table of sin/cos(a), |d| <= 1/128, and Taylor
//         series for sin d, cos d; quadrant k selects +-cos s / +-sin s.
// Every function disables FMA contraction: a fused p = m*c; p - 1 (for one)
// would break the error-free transforms.
#pragma once

#include "crmath_tables.h"
#include "mitos_common.h"

namespace mitos {

struct dd {
    double hi, lo;
};

__device__ __forceinline__ dd two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return dd{s, e};
}
__device__ __forceinline__ dd fast_two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b;
    const double e = b - (s - a);
    return dd{s, e};
}
__device__ __forceinline__ dd two_prod(double a, double b) {
#pragma clang fp contract(off)
    const double p = a * b;
    return dd{p, fma(a, b, -p)};
}
__device__ __forceinline__ dd dd_add(dd x, dd y) {
#pragma clang fp contract(off)
    dd s = two_sum(x.hi, y.hi);
    dd t = two_sum(x.lo, y.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ dd dd_add_d(dd x, double y) {
#pragma clang fp contract(off)
    dd s = two_sum(x.hi, y);
    s.lo += x.lo;
    return fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ dd dd_mul(dd x, dd y) {
#pragma clang fp contract(off)
    dd p = two_prod(x.hi, y.hi);
    p.lo = fma(x.hi, y.lo, fma(x.lo, y.hi, p.lo));
    return fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ dd dd_mul_d(dd x, double b) {
#pragma clang fp contract(off)
    dd p = two_prod(x.hi, b);
    p.lo = fma(x.lo, b, p.lo);
    return fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ dd dd_neg(dd x) { return dd{-x.hi, -x.lo}; }

// natural log of a positive normal double
__device__ __forceinline__ double cr_log(double x) {
#pragma clang fp contract(off)
    const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
    const int e = (int)((bits >> 52) & 0x7ff) - 1023;
    const double m = __longlong_as_double((long long)((bits & 0xfffffffffffffULL) |
                                                      0x3ff0000000000000ULL));
    const int j = (int)((bits >> 44) & 0xff);
    const int k = (j >= CR_LOG_KSTART) ? 1 : 0;
    const double c = CR_LOG_C[j];
    // r = m c - 1 exactly
    const double p = m * c;
    const double pe = fma(m, c, -p);
    const dd r = two_sum(p - 1.0, pe);
    // log1p(r) = r - r^2/2 + r^3/3 - r^4/4 + ...
    dd r2 = two_prod(r.hi, r.hi);
    r2.lo = fma(2.0 * r.hi, r.lo, r2.lo);
    r2 = fast_two_sum(r2.hi, r2.lo);
    const dd r3 = dd_mul(r2, r);
    const dd t3 = dd_mul(r3, dd{CR_THIRD_HI, CR_THIRD_LO});
    const double rh = r.hi;
    double poly = -1.0 / 12;
    poly = fma(poly, rh, 1.0 / 11);
    poly = fma(poly, rh, -1.0 / 10);
    poly = fma(poly, rh, 1.0 / 9);
    poly = fma(poly, rh, -1.0 / 8);
    poly = fma(poly, rh, 1.0 / 7);
    poly = fma(poly, rh, -1.0 / 6);
    poly = fma(poly, rh, 1.0 / 5);
    poly = fma(poly, rh, -0.25);
    const double tail = (r2.hi * r2.hi) * poly;
    dd s = dd_add(r, dd{-0.5 * r2.hi, -0.5 * r2.lo});
    s = dd_add(s, t3);
    s = dd_add_d(s, tail);
    // (e + k) ln2 - log(2^k c) + log1p(r)
    const double ek = (double)(e + k);
    dd L = two_prod(ek, CR_LN2_MID);
    L = dd_add_d(L, ek * CR_LN2_LO);
    L = dd_add_d(L, ek * CR_LN2_HI);  // exact product (42-bit head)
    dd res = dd_add(L, dd{CR_LOG_T[j][0], CR_LOG_T[j][1]});
    res = dd_add(res, s);
    return res.hi + res.lo;
}

// sin and cos of a double-double s, |s| <= pi/4 + 2^-20
__device__ __forceinline__ void dd_sincos(dd s, dd& sn, dd& cs) {
#pragma clang fp contract(off)
    const int j = (int)rint(s.hi * CR_NTRIG);
    const int aj = j < 0 ? -j : j;
    const double a = (double)j / CR_NTRIG;
    const dd d = fast_two_sum(s.hi - a, s.lo);  // s.hi - a is exact
    // d^2 as double-double
    dd d2 = two_prod(d.hi, d.hi);
    d2.lo = fma(2.0 * d.hi, d.lo, d2.lo);
    d2 = fast_two_sum(d2.hi, d2.lo);
    const double z = d2.hi;
    // sin d = d - d^3/6 + d^5 (1/120 - d^2/5040 + d^4/362880 - d^6/39916800)
    const dd d3 = dd_mul(d2, d);
    const dd d3_6 = dd_mul(d3, dd{1.0 / 6, 9.2518585385429707e-18});
    double ps = -1.0 / 39916800;
    ps = fma(ps, z, 1.0 / 362880);
    ps = fma(ps, z, -1.0 / 5040);
    ps = fma(ps, z, 1.0 / 120);
    const double s5 = (d3.hi * z) * ps;
    dd sind = dd_add(d, dd_neg(d3_6));
    sind = dd_add_d(sind, s5);
    // cos d = 1 - d^2/2 + d^4 (1/24 - d^2/720 + d^4/40320 - d^6/3628800)
    double pc = -1.0 / 3628800;
    pc = fma(pc, z, 1.0 / 40320);
    pc = fma(pc, z, -1.0 / 720);
    pc = fma(pc, z, 1.0 / 24);
    const double c4 = (z * z) * pc;
    dd cosd = dd_add(dd{1.0, 0.0}, dd{-0.5 * d2.hi, -0.5 * d2.lo});
    cosd = dd_add_d(cosd, c4);
    if (aj == 0) {
        sn = sind;
        cs = cosd;
        return;
    }
    dd sa{CR_TRIG[aj][0], CR_TRIG[aj][1]};
    const dd ca{CR_TRIG[aj][2], CR_TRIG[aj][3]};
    if (j < 0) sa = dd_neg(sa);
    // sin(a + d) = sin a cos d + cos a sin d; cos(a + d) = cos a cos d - sin a sin d
    sn = dd_add(dd_mul(sa, cosd), dd_mul(ca, sind));
    cs = dd_add(dd_mul(ca, cosd), dd_neg(dd_mul(sa, sind)));
}

// cos of a double t with |t| < 8 (the Box-Muller angle lies in (0, 2 pi))
__device__ __forceinline__ double cr_cos(double t) {
#pragma clang fp contract(off)
    const double kf = rint(t * 0.63661977236758134308);  // 2/pi
    const int q = ((int)kf) & 3;
    dd s{fma(-kf, CR_PIO2_1, t), 0.0};  // exact: k*P1 has <= 51 bits
    s = dd_add(s, two_prod(-kf, CR_PIO2_2));
    s = dd_add_d(s, -kf * CR_PIO2_3);
    dd sn, cs;
    dd_sincos(s, sn, cs);
    dd r = (q == 0) ? cs : (q == 1) ? dd_neg(sn) : (q == 2) ? dd_neg(cs) : sn;
    return r.hi + r.lo;
}

}  // namespace mitos
