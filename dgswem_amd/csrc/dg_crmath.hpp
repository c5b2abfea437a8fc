// dg_crmath.hpp -- correctly rounded cos(x) and x**(1/3) for the step's two transcendental terms.
//
// The reference evaluates the tidal boundary elevation with COS (src/ocean_edge_hydro.F:459-488,
// src/flow_edge_hydro.F:143-173) and Manning's friction with DEPTH**(1.D0/3.D0)
// (src/rhs_dg_hydro.F:302), i.e. the C library's cos and pow with the exponent 1/3 rounded to a
// double.  The slope limiter responds discontinuously to last-bit differences in its inputs, so
// on large meshes a 1-ulp difference between two libraries grows into differences far above
// round-off.  These functions return the correctly rounded value (the double nearest to the exact
// result), which is what an accurate C library returns, so that the GPU step reproduces the CPU
// step bit for bit.
//
// Both work in double-double arithmetic (a value as an unevaluated sum hi + lo) built from exact
// error-free transformations (fma-based products, two-sums) and round once at the end; the
// result is correctly rounded unless the exact value lies within about 2^-100 (relative) of a
// midpoint between two doubles.  Everything here must be compiled without floating-point
// contraction.
//
// Usable from host and device code (the host build is what the CPU tests check against the C
// library).
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define DG_CRM_FN __host__ __device__ __forceinline__
#else
#define DG_CRM_FN inline
#endif

namespace dg {
namespace crm {

struct dd {
    double hi, lo;
};

// a + b = s + e exactly
DG_CRM_FN dd two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
// the same for |a| >= |b|
DG_CRM_FN dd fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
// a * b = p + e exactly
DG_CRM_FN dd two_prod(double a, double b) {
    const double p = a * b;
    return {p, std::fma(a, b, -p)};
}
DG_CRM_FN dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return fast_two_sum(s.hi, s.lo);
}
DG_CRM_FN dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo = p.lo + (a.hi * b.lo + a.lo * b.hi);
    return fast_two_sum(p.hi, p.lo);
}

// 1/n! for n = 0..27 as double-double (hi, lo): the Taylor coefficients of cos and sin.
// (Exact rationals rounded to two doubles.)
constexpr double INV_FACT[28][2] = {
    {1.0, 0.0},
    {1.0, 0.0},
    {0x1.0000000000000p-1, 0.0},
    {0x1.5555555555555p-3, 0x1.5555555555555p-57},
    {0x1.5555555555555p-5, 0x1.5555555555555p-59},
    {0x1.1111111111111p-7, 0x1.1111111111111p-63},
    {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65},
    {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73},
    {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76},
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},
    {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76},
    {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80},
    {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83},
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87},
    {0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92},
    {0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97},
    {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101},
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103},
    {0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107},
    {0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112},
    {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120},
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120},
    {0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124},
    {0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130},
    {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135},
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139},
    {0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143},
    {0x1.d1ab1c2dccea3p-94, 0x1.054d0c78aea14p-149},
};

// pi/2 = P1 + P2 + P3 + P4 + O(2^-176): P1 and P2 carry 33 significant bits, so k*P1 and k*P2 are
// exact for |k| < 2^20; k*P3 is formed exactly as a two-product and k*P4 (about k*2^-123) needs no
// more than its rounding.  2/pi rounded to a double picks the quadrant.
constexpr double PIO2_1 = 0x1.921fb54400000p+0;
constexpr double PIO2_2 = 0x1.0b4611a600000p-34;
constexpr double PIO2_3 = 0x1.3198a2e037073p-69;
constexpr double PIO2_4 = 0x1.129024e088a68p-123;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;

// sum_{j=0..13} (-1)^j z^j / (2j+s)!  (s = 0: the cosine series, s = 1: the sine series over r),
// by Horner's rule p <- 1/(2j+s)! - z*p from the highest term down.  Kept as a loop: the kernels
// call it only on tidal and flux boundary edges, and unrolled it would bloat their code.
DG_CRM_FN dd even_series(dd z, int s) {
    dd p = {INV_FACT[26 + s][0], INV_FACT[26 + s][1]};
#pragma unroll 1
    for (int j = 12; j >= 0; j--) {
        const dd zp = dd_mul(z, p);
        p = dd_add({INV_FACT[2 * j + s][0], INV_FACT[2 * j + s][1]}, {-zp.hi, -zp.lo});
    }
    return p;
}

// Correctly rounded cos(x) for |x| < 2^16 (the step's arguments -- tidal and flux boundary phases,
// tidal potential arguments -- are below 30); other arguments, NaN and infinities go to the
// library's cos.  The reduced argument r = x - k*pi/2 is a double-double whose absolute error is
// below k*2^-170 (x - k*P1 and k*P2 exact, k*P3 an exact two-product, k*P4 rounded once), so even
// for the doubles closest to a multiple of pi/2 (|r| above 2^-62 for |x| < 2^16) its relative
// error stays below 2^-90, inside the series' accuracy.
DG_CRM_FN double cr_cos(double x) {
    if (!(std::fabs(x) < 65536.0)) return std::cos(x);
    const double k = std::rint(x * TWO_OVER_PI);
    const double a = x - k * PIO2_1;                 // exact (Sterbenz)
    const dd s = two_sum(a, -(k * PIO2_2));          // k*P2 exact
    const dd p3 = two_prod(k, PIO2_3);
    const dd u = dd_add(s, {-p3.hi, -p3.lo});
    const dd r = fast_two_sum(u.hi, u.lo - k * PIO2_4);
    const dd z = dd_mul(r, r);
    const int q = (int)((long long)k & 3);
    // even quadrants: cos r; odd: sin r = r * (the sine series)
    const int odd = q & 1;
    dd v = even_series(z, odd);
    if (odd) v = dd_mul(r, v);
    const double res = v.hi + v.lo;
    return (q == 1 || q == 2) ? -res : res;
}

// Correctly rounded x**y for y = 1/3 rounded to a double (0x1.5555555555555p-2), for positive x
// in [2^-1000, 2^1000]: x**y = x**(1/3) * x**(-d) with d = 1/3 - y = 2^-54/3.  The cube root is
// refined to double-double by one Newton step from the library's cbrt, and the second factor is
// 1 - d*ln(x) to 2^-110.  The correction terms need only modest accuracy (the Newton quotient
// 2^-45 relative, ln(x) 2^-20 absolute: their errors enter at 2^-97 and 2^-75 of the result), so
// they are formed in single precision: on the device from the hardware's approximate reciprocal
// and base-2 logarithm (about 1 ulp each), ln(x) as ln(2) * (exponent + log2(significand)).
// Outside that range as pow: +0 for +-0, NaN for negative x; the library's cube root otherwise.
DG_CRM_FN double cr_pow_third(double x) {
    const double c0 = std::cbrt(x);
    if (!(x >= 0x1p-1000 && x <= 0x1p+1000)) return x < 0.0 ? (x - x) / (x - x) : (x == 0.0 ? 0.0 : c0);
    const dd s = two_prod(c0, c0);
    const dd t = two_prod(c0, s.hi);
    const double e = (t.hi - x) + (t.lo + c0 * s.lo);     // c0^3 - x (t.hi - x is exact)
    const double a = 3.0 * s.hi;
#if defined(__HIP_DEVICE_COMPILE__)
    const float ra = __builtin_amdgcn_rcpf((float)a);
    const float l2 = (float)__builtin_amdgcn_frexp_exp(x) + __builtin_amdgcn_logf((float)__builtin_amdgcn_frexp_mant(x));
    const double eps = (double)l2 * (0.6931471805599453 * 0x1.5555555555555p-56);   // ln(x) * 2^-54 / 3
#else
    const float ra = 1.0f / (float)a;
    const double eps = (double)std::log((float)x) * 0x1.5555555555555p-56;         // ln(x) * 2^-54 / 3
#endif
    double r = (double)ra;
    r = r * (2.0 - a * r);                                 // 1/a to 2^-46
    const double d = e * r;                                // Newton: cbrt(x) = c0 - d
    return c0 + (-d - c0 * eps);
}

}  // namespace crm
}  // namespace dg
