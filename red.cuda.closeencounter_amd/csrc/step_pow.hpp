// step_pow.hpp -- x^y rounded to nearest, the step controller's factor (tol/err)^(1/(p+1)).
//
// The reference updates the trial step with its C runtime's pow (dt *= 0.9 * pow(tol / err,
// 1/(order+1)), int_rungekutta8.cu:777, int_rungekutta4.cu:253), whose last bit depends on
// the platform's libm.  Here the factor is the correctly rounded x^y, computed by the same
// operations on the host (the single-disk controller) and on the device (the ensemble's
// per-disk controller), so both give the same bits: every operation below is an IEEE
// addition, multiplication, fused multiply-add or division, or an exact bit manipulation;
// no library function is called.
//
// Method (|x - 1| > 2^-30): x = 2^e m with m in (sqrt(1/2), sqrt(2)]; log m = 2 atanh(s),
// s = (m-1)/(m+1), by its odd series in double-double arithmetic (|s| <= 0.172, 24 terms);
// z = y log x;
// exp z = 2^k exp(r), |r| <= ln2/2, exp(r) = (exp(r/16))^16 with a 15-term Taylor sum.
// The double-double result carries a relative error below about 2^-98, so the final
// rounding is the correct one unless x^y lies within that distance of a midpoint between
// two doubles (for inputs spread over the controller's range, about one call in 2^44).
// Within 2^-30 of x = 1 a binomial series and an exact rounding decision take over.
//
// Domain: 0 < y <= 1/2 (the controllers use 1/4, 1/5 and 1/8), any x.  x^y is then a normal
// double or +inf for every positive x.  x = +inf gives +inf, x = 0 gives 0, x < 0 or NaN
// gives NaN, as pow does for these y.
#pragma once

#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define STEP_POW_FN __host__ __device__ inline
#define STEP_POW_UNROLL _Pragma("unroll")
#else
#define STEP_POW_FN inline
#define STEP_POW_UNROLL
#endif

namespace red {
namespace steppow {

struct dd {
    double hi, lo;
};

STEP_POW_FN double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

STEP_POW_FN uint64_t bits_of(double x)
{
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    return u;
}

STEP_POW_FN double of_bits(uint64_t u)
{
    double x;
    __builtin_memcpy(&x, &u, 8);
    return x;
}

// 2^k for -1022 <= k <= 1023
STEP_POW_FN double pow2i(int k) { return of_bits((uint64_t)(k + 1023) << 52); }

STEP_POW_FN dd two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return dd{s, e};
}

// |a| >= |b| or a == 0
STEP_POW_FN dd fast_two_sum(double a, double b)
{
    const double s = a + b;
    const double e = b - (s - a);
    return dd{s, e};
}

STEP_POW_FN dd two_prod(double a, double b)
{
    const double p = a * b;
    return dd{p, fma_(a, b, -p)};
}

STEP_POW_FN dd add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return fast_two_sum(s.hi, s.lo);
}

STEP_POW_FN dd neg(dd a) { return dd{-a.hi, -a.lo}; }

STEP_POW_FN dd mul(dd a, dd b)
{
    dd p = two_prod(a.hi, b.hi);
    p.lo = p.lo + ((a.hi * b.lo) + (a.lo * b.hi));
    return fast_two_sum(p.hi, p.lo);
}

STEP_POW_FN dd mul_d(dd a, double b)
{
    dd p = two_prod(a.hi, b);
    p.lo = p.lo + a.lo * b;
    return fast_two_sum(p.hi, p.lo);
}

// a / b: three quotient digits, each from the remainder of the previous ones
STEP_POW_FN dd div(dd a, dd b)
{
    const double q1 = a.hi / b.hi;
    dd r = add(a, neg(mul_d(b, q1)));
    const double q2 = r.hi / b.hi;
    r = add(r, neg(mul_d(b, q2)));
    const double q3 = r.hi / b.hi;
    return add(fast_two_sum(q1, q2), dd{q3, 0.0});
}

// 1/(2k+1) (k = 0..24) and 1/i! (i = 0..15) as double-doubles: hi is the double nearest the
// value, lo the double nearest the rest (gen_step_pow_constants.py prints them).
constexpr double INV_ODD[25][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.5555555555555p-2, 0x1.5555555555555p-56},
    {0x1.999999999999ap-3, -0x1.999999999999ap-57},
    {0x1.2492492492492p-3, 0x1.2492492492492p-57},
    {0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58},
    {0x1.745d1745d1746p-4, -0x1.745d1745d1746p-59},
    {0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58},
    {0x1.1111111111111p-4, 0x1.1111111111111p-60},
    {0x1.e1e1e1e1e1e1ep-5, 0x1.e1e1e1e1e1e1ep-61},
    {0x1.af286bca1af28p-5, 0x1.af286bca1af28p-59},
    {0x1.8618618618618p-5, 0x1.8618618618618p-59},
    {0x1.642c8590b2164p-5, 0x1.642c8590b2164p-60},
    {0x1.47ae147ae147bp-5, -0x1.eb851eb851eb8p-61},
    {0x1.2f684bda12f68p-5, 0x1.2f684bda12f68p-59},
    {0x1.1a7b9611a7b96p-5, 0x1.1a7b9611a7b96p-61},
    {0x1.0842108421084p-5, 0x1.0842108421084p-60},
    {0x1.f07c1f07c1f08p-6, -0x1.f07c1f07c1f08p-61},
    {0x1.d41d41d41d41dp-6, 0x1.0750750750750p-60},
    {0x1.bacf914c1bad0p-6, -0x1.bacf914c1bad0p-60},
    {0x1.a41a41a41a41ap-6, 0x1.0690690690690p-60},
    {0x1.8f9c18f9c18fap-6, -0x1.f3831f3831f38p-61},
    {0x1.7d05f417d05f4p-6, 0x1.7d05f417d05f4p-62},
    {0x1.6c16c16c16c17p-6, -0x1.f49f49f49f49fp-61},
    {0x1.5c9882b931057p-6, 0x1.310572620ae4cp-61},
    {0x1.4e5e0a72f0539p-6, 0x1.e0a72f0539783p-60},
};
constexpr double INV_FACT[16][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.0000000000000p-1, 0x0.0p+0},
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
};

constexpr double LN2_HI = 0x1.62e42fefa39efp-1;    // ln 2 = LN2_HI + LN2_LO to 2^-108
constexpr double LN2_LO = 0x1.abc9e3b39803fp-56;
constexpr double INV_LN2 = 0x1.71547652b82fep+0;
constexpr double SQRT2 = 0x1.6a09e667f3bcdp+0;

// k ln 2 for an integer |k| <= 2200
STEP_POW_FN dd k_ln2(double k) { return add(two_prod(k, LN2_HI), dd{k * LN2_LO, 0.0}); }

// ln x for a positive finite x
STEP_POW_FN dd log_dd(double x)
{
    int e = 0;
    if (x < 0x1p-1022) {   // subnormal: scale into the normal range first
        x = x * 0x1p54;
        e = -54;
    }
    const uint64_t u = bits_of(x);
    e += (int)((u >> 52) & 0x7ff) - 1023;
    double m = of_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);   // [1, 2)
    if (m > SQRT2) {
        m = m * 0.5;
        e += 1;
    }
    // s = (m - 1) / (m + 1); m - 1 is exact for m in (sqrt(1/2), sqrt(2)]
    const dd s = div(dd{m - 1.0, 0.0}, two_sum(m, 1.0));
    const dd s2 = mul(s, s);
    // atanh(s) / s = sum_k s^(2k) / (2k + 1): s2 <= 0.0295, so 24 terms reach 2^-120
    constexpr int K = 24;
    dd p = dd{INV_ODD[K][0], INV_ODD[K][1]};
    STEP_POW_UNROLL
    for (int k = K - 1; k >= 0; k--) p = add(mul(p, s2), dd{INV_ODD[k][0], INV_ODD[k][1]});
    const dd logm = mul_d(mul(s, p), 2.0);
    return add(k_ln2((double)e), logm);
}

// exp z for |z.hi| <= 709 (the caller handles the rest)
STEP_POW_FN double exp_dd_rounded(dd z)
{
    const double k = __builtin_floor(z.hi * INV_LN2 + 0.5);
    dd r = add(z, neg(k_ln2(k)));                 // |r| <= ln2/2 (plus rounding)
    r.hi = r.hi * 0.0625;                         // r / 16, exact
    r.lo = r.lo * 0.0625;
    // exp(r/16) = sum_{i<=15} (r/16)^i / i!, |r/16| <= 0.0217
    constexpr int N = 15;
    dd p = dd{INV_FACT[N][0], INV_FACT[N][1]};
    STEP_POW_UNROLL
    for (int i = N - 1; i >= 0; i--) p = add(mul(p, r), dd{INV_FACT[i][0], INV_FACT[i][1]});
    for (int q = 0; q < 4; q++) p = mul(p, p);      // (exp(r/16))^16
    // p is normalised (fast_two_sum), so p.hi is p rounded to nearest; 2^k scales exactly
    // while the result stays normal, which y <= 1/2 guarantees
    const int ki = (int)k;
    if (ki > 1023) return p.hi * pow2i(1023) * pow2i(ki - 1023);
    if (ki < -1022) return p.hi * pow2i(-1022) * pow2i(ki + 1022);
    return p.hi * pow2i(ki);
}

// x^y for |x - 1| <= 2^-30 by the binomial series 1 + sum_k C(y, k) d^k, d = x - 1 (exact),
// rounded with an exact decision: near 1 the result can lie within ~2^-110 of a midpoint
// (x = 1 - 2^-53, y = 1/2, say), closer than the general path resolves.
STEP_POW_FN double pow_near_one(double x, double y)
{
    const double d = x - 1.0;
    dd c = dd{y, 0.0};                 // C(y, k)
    dd dk = dd{d, 0.0};                // d^k
    dd t = two_prod(y, d);
    for (int k = 2; k <= 5; k++) {
        c = div(mul(c, two_sum(y, -(double)(k - 1))), dd{(double)k, 0.0});
        dk = mul_d(dk, d);
        t = add(t, mul(c, dk));
    }
    // 1 + t = s + e + t.lo exactly; round it to nearest (ties to even)
    const dd se = two_sum(1.0, t.hi);
    const double s = se.hi;
    const dd w = two_sum(se.lo, t.lo);
    if (w.hi == 0.0) return s;
    const uint64_t bs = bits_of(s);
    const double nb = of_bits(w.hi > 0.0 ? bs + 1 : bs - 1);   // s > 0: the neighbour towards w
    const double half = (nb - s) * 0.5;                        // exact: a power of two
    bool away;
    if (w.hi != half) away = (w.hi > 0.0) ? (w.hi > half) : (w.hi < half);
    else if (w.lo != 0.0) away = (w.lo > 0.0) == (w.hi > 0.0);
    else away = (bits_of(nb) & 1) == 0;                        // exact tie: to even
    return away ? nb : s;
}

}  // namespace steppow

// x^y rounded to nearest, for 0 < y <= 1/2.
STEP_POW_FN double step_pow(double x, double y)
{
    if (x != x || y != y || x < 0.0) return __builtin_nan("");
    if (x == 0.0) return 0.0;
    if (x == __builtin_inf()) return __builtin_inf();
    if (x == 1.0) return 1.0;
    if (x >= 1.0 - 0x1p-30 && x <= 1.0 + 0x1p-30) return steppow::pow_near_one(x, y);
    const steppow::dd z = steppow::mul_d(steppow::log_dd(x), y);
    if (z.hi > 709.0) return __builtin_inf();     // x^y > 2^1023 needs y log x > 709.1
    return steppow::exp_dd_rounded(z);
}

}  // namespace red
