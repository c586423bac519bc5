// Oracle (test infrastructure only): deterministic FP64 pow for the CPU restatement.
//
// The reference evaluates van Genuchten laws with std::pow (glibc), e.g.
// dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:117 (pc), :224 (krw), :268 (krn).
// A finite-difference Jacobian with eps = 1e-10*(|x|+1) (dumux/common/numericdifferentiation.hh:40)
// amplifies a 1-ulp difference in the residual by ~1e6, so the oracle and the GPU kernels must
// evaluate pow with the *same* sequence of IEEE-754 add/mul/div operations. This file is the
// oracle's own implementation of that method (no FMA, no tables):
//
//   log(x) = e*ln2 + 2*atanh(s),  s = f/(2+f),  f = m-1,  m in [sqrt(1/2), sqrt(2))
//            (s carried as a double-double; atanh series truncated after s^23)
//   pow(x,y) = exp(y*log(x)) with the product kept as a double-double (Dekker's two-product)
//   exp(a)  = 2^k * exp(r),  r = a - k*ln2 (Cody-Waite split of ln2), degree-14 Taylor polynomial
//
// Accuracy (checked against mpmath in tests/test_detmath.py): < 1 ulp for the argument ranges
// used on this path. Compile with -ffp-contract=off so no multiply-add is fused.
#pragma once
#include <cstdint>
#include <cstring>

namespace orc {

static inline uint64_t dm_bits(double x) { uint64_t u; std::memcpy(&u, &x, 8); return u; }
static inline double dm_from_bits(uint64_t u) { double x; std::memcpy(&x, &u, 8); return x; }

// Knuth two-sum: s + e == a + b exactly.
static inline void dm_two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Dekker fast two-sum, requires |a| >= |b| (or a == 0).
static inline void dm_fast_two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    e = b - (s - a);
}

// Veltkamp split: a == hi + lo with hi holding the upper 26 bits.
static inline void dm_split(double a, double& hi, double& lo)
{
    const double c = 134217729.0 * a; // 2^27 + 1
    hi = c - (c - a);
    lo = a - hi;
}

// Dekker two-product: p + e == a * b exactly (no FMA).
static inline void dm_two_prod(double a, double b, double& p, double& e)
{
    p = a * b;
    double ah, al, bh, bl;
    dm_split(a, ah, al);
    dm_split(b, bh, bl);
    e = (((ah * bh - p) + ah * bl) + al * bh) + al * bl;
}

// ln2 split so that k*LN2_HI is exact for |k| < 2^21 (31 significant bits).
static const double DM_LN2_HI = 6.93147180369123816490e-01; // 0x3FE62E42FEE00000
static const double DM_LN2_LO = 1.90821492927058770002e-10; // 0x3DEA39EF35793C76
static const double DM_INV_LN2 = 1.44269504088896338700e+00; // 0x3FF71547652B82FE
static const double DM_SQRT2 = 1.41421356237309514547e+00;

// Natural logarithm of a positive, finite, normal-or-subnormal x as a double-double (hi, lo).
static inline void dm_log_dd(double x, double& hi, double& lo)
{
    int eadj = 0;
    uint64_t u = dm_bits(x);
    if ((u >> 52) == 0) { // subnormal
        x = x * 18014398509481984.0; // 2^54
        u = dm_bits(x);
        eadj = -54;
    }
    int e = int((u >> 52) & 0x7ff) - 1023 + eadj;
    double m = dm_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); // [1,2)
    if (m > DM_SQRT2) { m = m * 0.5; e += 1; }
    const double f = m - 1.0;             // exact (Sterbenz)
    double dh, dl;
    dm_fast_two_sum(2.0, f, dh, dl);      // 2+f as double-double
    const double sh = f / dh;
    double p, pe;
    dm_two_prod(sh, dh, p, pe);
    const double sl = (((f - p) - pe) - sh * dl) / dh;
    const double z = sh * sh;
    // 2/(2k+1) for k = 1..11
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
    dm_two_sum(de * DM_LN2_HI, 2.0 * sh, h, l);
    l = l + ((de * DM_LN2_LO + 2.0 * sl) + tail);
    dm_fast_two_sum(h, l, hi, lo);
}

// 2^k as a double for -1022 <= k <= 1023.
static inline double dm_pow2i(int k) { return dm_from_bits(uint64_t(k + 1023) << 52); }

// exp(ah + al) for a double-double argument (al small w.r.t. ah).
static inline double dm_exp_dd(double ah, double al)
{
    if (ah != ah) return ah;
    if (ah > 709.8) return 1.0 / 0.0;
    if (ah < -745.2) return 0.0;
    const double kd = __builtin_rint(ah * DM_INV_LN2);
    const int k = int(kd);
    const double r0 = ah - kd * DM_LN2_HI;  // exact (Cody-Waite)
    double rh, rl;
    dm_two_sum(r0, -(kd * DM_LN2_LO), rh, rl);
    rl = rl + al;
    double t;
    dm_fast_two_sum(rh, rl, t, rl);
    rh = t;
    // exp(rh) - 1 - rh = rh^2 * Q(rh), Taylor to degree 14
    double q = 1.0 / 87178291200.0;           // 1/14!
    q = 1.0 / 6227020800.0 + rh * q;          // 1/13!
    q = 1.0 / 479001600.0 + rh * q;           // 1/12!
    q = 1.0 / 39916800.0 + rh * q;            // 1/11!
    q = 1.0 / 3628800.0 + rh * q;             // 1/10!
    q = 1.0 / 362880.0 + rh * q;              // 1/9!
    q = 1.0 / 40320.0 + rh * q;               // 1/8!
    q = 1.0 / 5040.0 + rh * q;                // 1/7!
    q = 1.0 / 720.0 + rh * q;                 // 1/6!
    q = 1.0 / 120.0 + rh * q;                 // 1/5!
    q = 1.0 / 24.0 + rh * q;                  // 1/4!
    q = 1.0 / 6.0 + rh * q;                   // 1/3!
    q = 0.5 + rh * q;                         // 1/2!
    const double em1 = rh + (rh * rh) * q;    // exp(rh) - 1 (rounded)
    double s, se;
    dm_fast_two_sum(1.0, rh, s, se);          // 1 + rh exactly
    // exp(rh + rl) = (1 + rh) + rh^2 Q + rl*(1 + em1)
    const double mant = s + (se + ((rh * rh) * q + rl * (1.0 + em1)));
    if (k >= -1021 && k <= 1023)
        return mant * dm_pow2i(k);
    if (k > 1023)
        return (mant * dm_pow2i(1023)) * dm_pow2i(k - 1023);
    return (mant * dm_pow2i(k + 600)) * dm_pow2i(-600);
}

// pow(x, y) for x >= 0 (the only domain used on this path; NaN for x < 0 with non-trivial y).
static inline double dpow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    if (x < 0.0) return 0.0 / 0.0;
    if (x == 0.0) return y > 0.0 ? 0.0 : 1.0 / 0.0;
    if (x == 1.0 / 0.0) return y > 0.0 ? x : 0.0;
    double lh, ll;
    dm_log_dd(x, lh, ll);
    double ph, pe;
    dm_two_prod(y, lh, ph, pe);
    pe = pe + y * ll;
    double qh, ql;
    dm_fast_two_sum(ph, pe, qh, ql);
    return dm_exp_dd(qh, ql);
}

// log(x) rounded to double, for the seeded log-normal generator.
static inline double dlog(double x)
{
    if (x != x || x < 0.0) return 0.0 / 0.0;
    if (x == 0.0) return -1.0 / 0.0;
    if (x == 1.0 / 0.0) return x;
    double h, l;
    dm_log_dd(x, h, l);
    return h + l;
}

} // namespace orc
