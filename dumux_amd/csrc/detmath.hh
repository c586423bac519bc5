// Deterministic FP64 pow/log for host and gfx950 device code.
//
// DuMux evaluates the van Genuchten law with std::pow
// (dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:117, 224, 268). The numeric
// Jacobian divides residual differences by eps = 1e-10*(|x|+1)
// (dumux/common/numericdifferentiation.hh:40), so the device residual must reproduce the host
// residual bit for bit. The device math library's pow and the host libm differ in the last
// bits, so this header supplies one pow that uses only IEEE-754 add, multiply and divide
// (all correctly rounded on gfx950 and x86-64) in a fixed order:
//   log x  = e*ln2 + 2*atanh(s), s = f/(2+f), f = m-1, m in [sqrt(1/2), sqrt(2)), s as double-double
//   pow    = exp(y*log x), product as double-double (Dekker two-product, no FMA)
//   exp a  = 2^k*exp(r), r = a - k*ln2 (Cody-Waite), Taylor polynomial of degree 14
// Build all users with -ffp-contract=off. Error < 0.6 ulp on the argument ranges of this path.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DMX_HD __host__ __device__ __forceinline__
#else
#define DMX_HD inline
#endif

namespace dmx {

DMX_HD uint64_t dbits(double x) { return __builtin_bit_cast(uint64_t, x); }
DMX_HD double dfrombits(uint64_t u) { return __builtin_bit_cast(double, u); }

DMX_HD void two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

DMX_HD void fast_two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    e = b - (s - a);
}

DMX_HD void veltkamp(double a, double& hi, double& lo)
{
    const double c = 134217729.0 * a;
    hi = c - (c - a);
    lo = a - hi;
}

DMX_HD void two_prod(double a, double b, double& p, double& e)
{
    p = a * b;
    double ah, al, bh, bl;
    veltkamp(a, ah, al);
    veltkamp(b, bh, bl);
    e = (((ah * bh - p) + ah * bl) + al * bh) + al * bl;
}

constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kInvLn2 = 1.44269504088896338700e+00;
constexpr double kSqrt2 = 1.41421356237309514547e+00;

DMX_HD void log_dd(double x, double& hi, double& lo)
{
    int eadj = 0;
    uint64_t u = dbits(x);
    if ((u >> 52) == 0) {
        x = x * 18014398509481984.0;
        u = dbits(x);
        eadj = -54;
    }
    int e = int((u >> 52) & 0x7ff) - 1023 + eadj;
    double mant = dfrombits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (mant > kSqrt2) {
        mant = mant * 0.5;
        e += 1;
    }
    const double f = mant - 1.0;
    double dh, dl;
    fast_two_sum(2.0, f, dh, dl);
    const double sh = f / dh;
    double p, pe;
    two_prod(sh, dh, p, pe);
    const double sl = (((f - p) - pe) - sh * dl) / dh;
    const double z = sh * sh;
    double poly = 2.0 / 23.0;
    poly = 2.0 / 21.0 + z * poly;
    poly = 2.0 / 19.0 + z * poly;
    poly = 2.0 / 17.0 + z * poly;
    poly = 2.0 / 15.0 + z * poly;
    poly = 2.0 / 13.0 + z * poly;
    poly = 2.0 / 11.0 + z * poly;
    poly = 2.0 / 9.0 + z * poly;
    poly = 2.0 / 7.0 + z * poly;
    poly = 2.0 / 5.0 + z * poly;
    poly = 2.0 / 3.0 + z * poly;
    const double tail = (sh * z) * poly;
    const double de = double(e);
    double h, l;
    two_sum(de * kLn2Hi, 2.0 * sh, h, l);
    l = l + ((de * kLn2Lo + 2.0 * sl) + tail);
    fast_two_sum(h, l, hi, lo);
}

DMX_HD double pow2i(int k) { return dfrombits(uint64_t(k + 1023) << 52); }

DMX_HD double exp_dd(double ah, double al)
{
    if (ah != ah) return ah;
    if (ah > 709.8) return __builtin_huge_val();
    if (ah < -745.2) return 0.0;
    const double kd = __builtin_rint(ah * kInvLn2);
    const int k = int(kd);
    const double r0 = ah - kd * kLn2Hi;
    double rh, rl;
    two_sum(r0, -(kd * kLn2Lo), rh, rl);
    rl = rl + al;
    double t;
    fast_two_sum(rh, rl, t, rl);
    rh = t;
    double q = 1.0 / 87178291200.0;
    q = 1.0 / 6227020800.0 + rh * q;
    q = 1.0 / 479001600.0 + rh * q;
    q = 1.0 / 39916800.0 + rh * q;
    q = 1.0 / 3628800.0 + rh * q;
    q = 1.0 / 362880.0 + rh * q;
    q = 1.0 / 40320.0 + rh * q;
    q = 1.0 / 5040.0 + rh * q;
    q = 1.0 / 720.0 + rh * q;
    q = 1.0 / 120.0 + rh * q;
    q = 1.0 / 24.0 + rh * q;
    q = 1.0 / 6.0 + rh * q;
    q = 0.5 + rh * q;
    const double em1 = rh + (rh * rh) * q;
    double s, se;
    fast_two_sum(1.0, rh, s, se);
    const double mant = s + (se + ((rh * rh) * q + rl * (1.0 + em1)));
    if (k >= -1021 && k <= 1023) return mant * pow2i(k);
    if (k > 1023) return (mant * pow2i(1023)) * pow2i(k - 1023);
    return (mant * pow2i(k + 600)) * pow2i(-600);
}

// pow(x, y) for x >= 0 (NaN for negative x with y != 0)
DMX_HD double pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return y > 0.0 ? 0.0 : __builtin_huge_val();
    if (x == __builtin_huge_val()) return y > 0.0 ? x : 0.0;
    double lh, ll;
    log_dd(x, lh, ll);
    double ph, pe;
    two_prod(y, lh, ph, pe);
    pe = pe + y * ll;
    double qh, ql;
    fast_two_sum(ph, pe, qh, ql);
    return exp_dd(qh, ql);
}

DMX_HD double log(double x)
{
    if (x != x || x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_huge_val();
    if (x == __builtin_huge_val()) return x;
    double h, l;
    log_dd(x, h, l);
    return h + l;
}

} // namespace dmx
