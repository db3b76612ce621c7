// Portable binary64 pow(x, y) built only from IEEE-754 operations that give
// the same bits on the CPU and on CDNA4 (+, -, *, /, fma, rint, integer bit
// manipulation), so the HIP kernels and the CPU oracle can share one
// definition and agree bit for bit.
//
// The reference computes the HLLC wave-speed estimate with std::pow
// (Source/Fluxes/HLLC.H:270-273): pow(p_K, -z) and pow(., 1/z) with
// z = (gamma-1)/(2 gamma).  glibc's pow and OCML's pow differ from each
// other in the last bit for a fraction of inputs, so neither can be used on
// both sides.  This one follows the textbook table-driven method
// (Tang, "Table-driven implementation of the logarithm/exponential function",
// ACM TOMS 1989/1990; Muller, "Handbook of Floating-Point Arithmetic", ch. 11):
//
//   log: x = 2^k m, m in [1,2) (m/2 and k+1 when m >= 1.5); j = top 7 bits
//        of m; c_j the centre of m's interval (c = 1 on the two intervals
//        next to 1); m*invc_j = 1 + r as an exact double-double, |r| < 2^-7;
//        log x = k ln2 + (-log invc_j) + log1p(r), kept as a double-double.
//   exp: t = y log x (double-double); t = (64 k + j) ln2/64 + r, |r| <= ln2/128;
//        e^t = 2^k * 2^(j/64) * e^r, with e^r - 1 a degree-7 polynomial.
//
// Internal error is below 2^-62 relative before the final rounding, so the
// result is within 0.51 ulp of the exact value (tested against 60-digit
// decimal arithmetic and against glibc in the test suite).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define BX_POW_TABLE static __device__ __constant__ const
#define BX_POW_FN __device__ __forceinline__
#else
#define BX_POW_TABLE static const
#define BX_POW_FN static inline
#endif

#include "bx_pow_tables.h"

namespace bxpow {

BX_POW_FN uint64_t d2u(double x) { return __builtin_bit_cast(uint64_t, x); }
BX_POW_FN double u2d(uint64_t u) { return __builtin_bit_cast(double, u); }

// Exact sum s + e = a + b (Knuth TwoSum).
BX_POW_FN void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Exact sum s + e = a + b, valid when |a| >= |b| or a == 0.
BX_POW_FN void fast_two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    e = b - (s - a);
}

// log(x) as a double-double, x positive, finite, normal.
BX_POW_FN void log_dd(double x, double &lhi, double &llo)
{
    const uint64_t ix = d2u(x);
    const int      j  = (int)((ix >> 45) & 127);
    // m in [1,2); for j >= 64 use m/2 in [0.75,1) and k+1
    const int hi = j >> 6;
    const int k  = (int)((ix >> 52) & 0x7ff) - 1023 + hi;
    const double m
        = u2d((ix & 0x000fffffffffffffULL) | ((uint64_t)(1023 - hi) << 52));

    const double invc = bx_log_invc[j];
    // m * invc = p_hi + p_lo exactly; r = p - 1
    const double p_hi = m * invc;
    const double p_lo = __builtin_fma(m, invc, -p_hi);
    const double r0   = p_hi - 1.0; // exact (Sterbenz)
    // r0 is 0 or at least one ulp of p_hi, which bounds |p_lo|
    double r_hi, r_lo;
    fast_two_sum(r0, p_lo, r_hi, r_lo);

    // log1p(r) = r - r^2/2 + r^3 (1/3 - r/4 + r^2/5 - r^3/6 + r^4/7 - r^5/8)
    const double sq_hi = r_hi * r_hi;
    const double sq_lo = __builtin_fma(r_hi, r_hi, -sq_hi);
    const double r3    = sq_hi * r_hi;
    double       poly  = -0.125;
    poly = __builtin_fma(poly, r_hi, 0x1.2492492492492p-3);  // 1/7
    poly = __builtin_fma(poly, r_hi, -0x1.5555555555555p-3); // -1/6
    poly = __builtin_fma(poly, r_hi, 0x1.999999999999ap-3);  // 1/5
    poly = __builtin_fma(poly, r_hi, -0.25);
    poly = __builtin_fma(poly, r_hi, 0x1.5555555555555p-2);  // 1/3

    // k ln2 + logc: both multiples of 2^-42, the sum is exact; it is 0 or
    // larger in magnitude than r (table design), and b dominates r^2/2
    const double kd = (double)k;
    const double a  = kd * BX_LN2_HI + bx_log_logc_hi[j];
    double       b_hi, b_lo;
    fast_two_sum(a, r_hi, b_hi, b_lo);
    double c_hi, c_lo;
    fast_two_sum(b_hi, -0.5 * sq_hi, c_hi, c_lo);

    const double small = kd * BX_LN2_LO + bx_log_logc_lo[j] + r_lo
                         - 0.5 * sq_lo - r_hi * r_lo + r3 * poly;
    fast_two_sum(c_hi, c_lo + b_lo + small, lhi, llo);
}

// e^(t_hi + t_lo) for |t_hi| well inside the finite range.
BX_POW_FN double exp_dd(double t_hi, double t_lo)
{
    const double kk   = __builtin_rint(t_hi * BX_INV_LN2_64);
    const int    ki   = (int)kk;
    const int    j    = ki & 63;
    const int    k    = (ki - j) >> 6;
    const double r_hi = __builtin_fma(-kk, BX_LN2_64_HI, t_hi); // exact
    const double r_lo = t_lo - kk * BX_LN2_64_LO;
    double       r, rl;
    fast_two_sum(r_hi, r_lo, r, rl);

    // e^r - 1 = r + [rl + r^2/2 + r^3 (1/6 + r/24 + r^2/120 + r^3/720 + r^4/5040)]
    double poly = 0x1.a01a01a01a01ap-13;                 // 1/5040
    poly        = __builtin_fma(poly, r, 0x1.6c16c16c16c17p-10); // 1/720
    poly        = __builtin_fma(poly, r, 0x1.1111111111111p-7);  // 1/120
    poly        = __builtin_fma(poly, r, 0x1.5555555555555p-5);  // 1/24
    poly        = __builtin_fma(poly, r, 0x1.5555555555555p-3);  // 1/6
    const double r2  = r * r;
    const double elo = rl + 0.5 * r2 + r * rl + (r2 * r) * poly;

    const double th  = bx_exp_t_hi[j];
    const double tl  = bx_exp_t_lo[j];
    const double phi = th * r;
    const double plo = __builtin_fma(th, r, -phi);
    const double sml = plo + th * elo + tl + tl * (r + elo);
    double       s_hi, s_lo;
    fast_two_sum(th, phi, s_hi, s_lo); // th >= 1 > |phi|
    double res = s_hi + (s_lo + sml);

    // scale by 2^k without overflow in the scale factor itself
    if (k > -1022 && k < 1024)
        return res * u2d((uint64_t)(k + 1023) << 52);
    const int k1 = k / 2;
    const int k2 = k - k1;
    res *= u2d((uint64_t)(k1 + 1023) << 52);
    return res * u2d((uint64_t)(k2 + 1023) << 52);
}

BX_POW_FN bool is_integer(double y)
{
    return __builtin_rint(y) == y;
}

BX_POW_FN bool is_odd_integer(double y)
{
    if (!is_integer(y) || __builtin_fabs(y) >= 0x1p53)
        return false;
    const double h = 0.5 * y;
    return __builtin_rint(h) != h;
}

} // namespace bxpow

// Special cases of pow (C99 Annex F) and the rare inputs the fast path
// below does not take: x <= 0, NaN or inf operands, subnormal x, results
// near the overflow / underflow thresholds.
#if defined(__HIPCC__)
__device__ __attribute__((noinline))
#else
static
#endif
double bx_pow_slow(double x, double y)
{
    using namespace bxpow;
    if (y == 0.0)
        return 1.0;
    if (x != x || y != y)
        return x == 1.0 ? 1.0 : x + y;
    double sign = 1.0;
    if (!(x > 0.0))
    {
        if (x == 0.0)
        {
            const bool odd = is_odd_integer(y);
            if (y > 0)
                return odd ? x : 0.0;
            return odd ? 1.0 / x : __builtin_inf();
        }
        // x < 0
        if (!is_integer(y))
            return __builtin_nan("");
        if (is_odd_integer(y))
            sign = -1.0;
        x = -x;
    }
    if (x == __builtin_inf())
        return sign * (y > 0 ? __builtin_inf() : 0.0);
    if (__builtin_fabs(y) == __builtin_inf())
    {
        if (x == 1.0)
            return 1.0;
        return ((x > 1.0) == (y > 0)) ? __builtin_inf() : 0.0;
    }
    // subnormal x: rescale into the normal range
    int    kadj = 0;
    double xs   = x;
    if (xs < 0x1p-1022)
    {
        xs *= 0x1p54;
        kadj = 54;
    }
    double lhi, llo;
    log_dd(xs, lhi, llo);
    if (kadj)
    {
        // log x = log xs - 54 ln2
        double s, e;
        two_sum(lhi, -54.0 * BX_LN2_HI, s, e);
        llo = llo + e - 54.0 * BX_LN2_LO;
        fast_two_sum(s, llo, lhi, llo);
    }
    const double t_hi = y * lhi;
    const double t_lo = __builtin_fma(y, lhi, -t_hi) + y * llo;
    double       th, tl;
    fast_two_sum(t_hi, t_lo, th, tl);
    if (th > 709.8)
        return sign * __builtin_inf();
    if (th < -745.2)
        return sign * 0.0;
    return sign * exp_dd(th, tl);
}

// pow(x, y).  Fast path: x positive and normal, y finite and nonzero, and
// |y log x| < 708 (result well inside the normal range); it performs exactly
// the same operations as the general path for those inputs.
BX_POW_FN double bx_pow(double x, double y)
{
    using namespace bxpow;
    if (x >= 0x1p-1022 && x < 0x1p1023 && __builtin_fabs(y) < 0x1p1023
        && y != 0.0)
    {
        double lhi, llo;
        log_dd(x, lhi, llo);
        const double t_hi = y * lhi;
        const double t_lo = __builtin_fma(y, lhi, -t_hi) + y * llo;
        double       th, tl;
        fast_two_sum(t_hi, t_lo, th, tl);
        if (__builtin_fabs(th) < 708.0)
            return exp_dd(th, tl);
    }
    return bx_pow_slow(x, y);
}

// ---------------------------------------------------------------------------
// The two exponents of the HLLC pressure estimate, -z and 1/z with
// z = (gamma-1)/(2 gamma) rounded to double, are within a few ulps of -1/n
// and n for the integer n = round(1/z) whenever 2 gamma/(gamma-1) is an
// integer (gamma = 1.4: n = 7; gamma = 5/3: n = 5).  Then
//   x^(n + d)     = x^n (1 + d log x + ...),        d = 1/z - n   (exact),
//   x^(-1/n + w)  = x^(-1/n) (1 + w log x + ...),   w = 1/n - z,
// with |d|, |w| <= 2^-48, so the dropped terms stay below 4e-24 relative
// for every double x.  x^n is formed in double-double by binary powering
// and x^(-1/n) by one double-double Newton step on x r^n = 1 from a 1e-12
// estimate, so both values are known to ~1e-23 relative before the final
// rounding: the result is the correctly rounded pow(x, y) except when the
// exact value lies within ~1e-23 of a rounding boundary, the same answer the
// general routine above gives, for a fraction of its instruction count.
// The double-double work runs on operands rescaled by exact powers of two
// (x r^n is scale-free), so no input range needs the general routine: the
// *_ni forms below have no branch into it and keep a kernel that calls them
// free of its registers.
namespace bxpow {

// Double-double products left unnormalised: the pair (h, l) stands for
// h + l with |l| at most a few ulps of h.  Each step forms the leading
// product exactly (p + e = h x, by fma) and folds the cross terms into the
// new low part; the dropped l*l terms and the roundings of the low parts
// stay below ~2^-100 relative over the few steps of x^n (n <= 32), so no
// renormalising TwoSum is needed between steps.
BX_POW_FN void ddu_sqr(double &h, double &l)
{
    const double p = h * h;
    const double e = __builtin_fma(h, h, -p); // h^2 - p, exact
    l              = __builtin_fma(h + h, l, e);
    h              = p;
}

BX_POW_FN void ddu_mul_d(double &h, double &l, double x)
{
    const double p = h * x;
    const double e = __builtin_fma(h, x, -p); // h x - p, exact
    l              = __builtin_fma(l, x, e);
    h              = p;
}

// x^n (n >= 2) as an unnormalised double-double, left-to-right binary
// powering; the first square is of the double x itself (exact: x^2 = h + l).
// NF > 0 fixes n = NF at compile time (the loop unrolls); the operations,
// and so the result, are the same as for a run-time n = NF.
template <int NF>
BX_POW_FN void ddu_powi(double x, int n, double &h, double &l)
{
    if (NF > 0)
        n = NF;
    const int top = 31 - __builtin_clz((unsigned)n);
    h             = x * x;
    l             = __builtin_fma(x, x, -h);
    if ((n >> (top - 1)) & 1)
        ddu_mul_d(h, l, x);
    for (int b = top - 2; b >= 0; --b)
    {
        ddu_sqr(h, l);
        if ((n >> b) & 1)
            ddu_mul_d(h, l, x);
    }
}

// log(x) to ~1e-11 absolute, x positive and normal (same table split as
// log_dd; log1p(r) to degree 4 in r, |r| < 2^-7)
BX_POW_FN double log_approx(double x)
{
    const uint64_t ix = d2u(x);
    const int      j  = (int)((ix >> 45) & 127);
    const int      hi = j >> 6;
    const int      k  = (int)((ix >> 52) & 0x7ff) - 1023 + hi;
    const double   m
        = u2d((ix & 0x000fffffffffffffULL) | ((uint64_t)(1023 - hi) << 52));
    const double r = __builtin_fma(m, bx_log_invc[j], -1.0);
    double       p = -0.25;
    p              = __builtin_fma(p, r, 0x1.5555555555555p-2); // 1/3
    p              = __builtin_fma(p, r, -0.5);
    p              = __builtin_fma(p, r, 1.0);
    return ((double)k * BX_LN2_HI + bx_log_logc_hi[j]) + r * p;
}

// e^t to ~1e-13 relative for |t| <= 700
BX_POW_FN double exp_approx(double t)
{
    const double kk = __builtin_rint(t * BX_INV_LN2_64);
    const int    ki = (int)kk;
    const int    j  = ki & 63;
    const int    k  = (ki - j) >> 6;
    double       r  = __builtin_fma(-kk, BX_LN2_64_HI, t);
    r               = __builtin_fma(-kk, BX_LN2_64_LO, r);
    double p        = 0x1.5555555555555p-5; // 1/24
    p               = __builtin_fma(p, r, 0x1.5555555555555p-3); // 1/6
    p               = __builtin_fma(p, r, 0.5);
    p               = __builtin_fma(p, r, 1.0);
    const double th = bx_exp_t_hi[j];
    return __builtin_fma(th, r * p, th) * u2d((uint64_t)(k + 1023) << 52);
}

} // namespace bxpow

// Exponent data of the HLLC estimate for a given gamma.  n = 0 selects the
// general routine (2 gamma/(gamma-1) not an integer in [2, 32]).
struct BxPowHllc
{
    int    n;
    int    qmul;  // round(2^16 / n): (k * qmul) >> 16 is within 1 of k / n
    double inv_n; // RN(1/n)
    double d;     // 1/z - n, exact
    double w;     // 1/n - z to ~1e-16 relative
};

static inline BxPowHllc bx_pow_hllc_setup(double gamma)
{
    BxPowHllc    s  = { 0, 0, 0.0, 0.0, 0.0 };
    const double z  = (gamma - 1) / (2 * gamma);
    const double yi = 1 / z;
    if (!(yi >= 1.5 && yi <= 32.5))
        return s;
    const double nd = __builtin_rint(yi);
    const double d  = yi - nd;
    // w = (1 - n z) / n with n z as an exact double-double
    const double nz_hi = nd * z;
    const double nz_lo = __builtin_fma(nd, z, -nz_hi);
    const double w     = ((1.0 - nz_hi) - nz_lo) / nd;
    if (!(__builtin_fabs(d) <= 0x1p-48 && __builtin_fabs(w) <= 0x1p-48))
        return s;
    s.n     = (int)nd;
    s.qmul  = (65536 + s.n / 2) / s.n;
    s.inv_n = 1.0 / nd;
    s.d     = d;
    s.w     = w;
    return s;
}

namespace bxpow {
// unbiased exponent field of x (sign ignored; -1023 for zero and subnormals)
BX_POW_FN int exponent_field(double x)
{
    return (int)((d2u(x) >> 52) & 0x7ff) - 1023;
}
} // namespace bxpow

// pow(x, -z) for s.n != 0 and x >= 2^-1022, +inf or NaN (the estimate's
// pressures, floored at 1e-14).  The Newton residual x r^n - 1 is formed on
// x 2^(-n q) and r 2^q with q ~ k/n (k the exponent of x): exact rescalings
// that keep the double-double intermediates normal for every such x.
// NF > 0 fixes n = NF = s.n at compile time; for n = 7 the estimate comes
// from tables (negz_est7), which changes the estimate but not its accuracy,
// so NF = 7 is what any caller with s.n == 7 uses (bx_pow_negz dispatches).
namespace bxpow {
// x^(-1/7) 2^q to ~1e-13 relative (rs), x 2^(-7q) (xs), and log x to
// ~1e-8 absolute (lx), for x positive and normal.  x = 2^k m with the
// log_approx split (m in [0.75, 1.5)), k = 7q + i with 0 <= i < 7:
//   x^(-1/7) = 2^-q 2^(-i/7) invc_j^(1/7) (1 + t)^(-1/7),  1 + t = m invc_j,
// |t| < 2^-8, (1 + t)^(-1/7) by its binomial series to t^4.
BX_POW_FN void negz_est7(double x, double &rs, double &xs, double &lx, int &q)
{
    const uint64_t ix = d2u(x);
    const int      j  = (int)((ix >> 45) & 127);
    const int      hi = j >> 6;
    const int      k  = (int)((ix >> 52) & 0x7ff) - 1023 + hi;
    const double   m
        = u2d((ix & 0x000fffffffffffffULL) | ((uint64_t)(1023 - hi) << 52));
    const double t = __builtin_fma(m, bx_log_invc[j], -1.0);
    double       p = 110.0 / 2401.0;                // C(-1/7, 4)
    p              = __builtin_fma(p, t, -20.0 / 343.0); // C(-1/7, 3)
    p              = __builtin_fma(p, t, 4.0 / 49.0);    // C(-1/7, 2)
    p              = __builtin_fma(p, t, -1.0 / 7.0);    // C(-1/7, 1)
    p              = __builtin_fma(p, t, 1.0);
    // q = floor(k / 7): k + 1050 in [28, 2075], where (v 9363) >> 16 == v / 7
    q           = (((k + 1050) * 9363) >> 16) - 150;
    const int i = k - 7 * q;
    rs          = (bx_r7_invc[j] * bx_r7_two[i]) * p;
    xs          = __builtin_ldexp(x, -7 * q);
    lx          = ((double)k * BX_LN2_HI + bx_log_logc_hi[j]) + t * __builtin_fma(-0.5, t, 1.0);
}
} // namespace bxpow

template <int NF = 0>
BX_POW_FN double bx_pow_negz_ni(double x, const BxPowHllc &s)
{
    using namespace bxpow;
    const int n = NF > 0 ? NF : s.n;
    double    r, xs, lx;
    int       q;
    if constexpr (NF == 7)
    {
        negz_est7(x, r, xs, lx, q); // r = x^(-1/7) 2^q
    }
    else
    {
        lx = log_approx(x);
        q  = (exponent_field(x) * s.qmul) >> 16;
        r  = __builtin_ldexp(exp_approx(-lx * s.inv_n), q); // x^(-1/n) 2^q to ~1e-12
        xs = __builtin_ldexp(x, -n * q);
    }
    double h, l;
    ddu_powi<NF>(r, n, h, l);
    ddu_mul_d(h, l, xs);                        // x r^n = h + l, close to 1
    const double delta = (1.0 - h) - l;         // 1 - h exact (Sterbenz)
    const double c     = r * (delta * s.inv_n); // Newton correction
    const double res   = __builtin_ldexp(r + (c + r * (s.w * lx)), -q);
    return x == __builtin_inf() ? 0.0 : res;
}

// pow(x, -z); y = -z is the exact exponent
BX_POW_FN double bx_pow_negz(double x, double y, const BxPowHllc &s)
{
    if (!(s.n != 0 && x >= 0x1p-1022))
        return bx_pow(x, y);
    if (s.n == 7)
        return bx_pow_negz_ni<7>(x, s);
    return bx_pow_negz_ni<0>(x, s);
}

// pow(x, 1/z) for s.n != 0 and any x.  x^n is formed on the mantissa
// m = x 2^-k in [1, 2) and scaled back by 2^(n k) after the final rounding:
// exact unless the result is below 2^-1022, where that second scaling
// rounds again (results under 2^-1022 may be off by one subnormal ulp).
// NF > 0 fixes n = NF = s.n at compile time, with the same result.
template <int NF = 0>
BX_POW_FN double bx_pow_invz_ni(double x, const BxPowHllc &s)
{
    using namespace bxpow;
    const int n = NF > 0 ? NF : s.n;
    const int k = exponent_field(x);
    double    h, l;
    ddu_powi<NF>(__builtin_ldexp(x, -k), n, h, l); // signed: (-m)^n
    const double t   = s.d * log_approx(x);        // log |x|
    double       res = __builtin_ldexp(h + __builtin_fma(h, t, l), n * k);
    // 1/z is an integer only when d == 0; otherwise x < 0 gives NaN
    const bool odd_int = s.d == 0.0 && (n & 1);
    if (x == 0.0)
        res = odd_int ? x : 0.0;
    else if (__builtin_fabs(x) == __builtin_inf())
        res = (odd_int && x < 0) ? -__builtin_inf() : __builtin_inf();
    else if (x < 0 && s.d != 0.0)
        res = __builtin_nan("");
    return res;
}

// pow(x, 1/z); y = 1/z is the exact exponent
BX_POW_FN double bx_pow_invz(double x, double y, const BxPowHllc &s)
{
    if (s.n == 0)
        return bx_pow(x, y);
    return bx_pow_invz_ni<0>(x, s);
}
