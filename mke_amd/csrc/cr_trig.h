/*
 * cr_trig.h - correctly rounded cos and sin of a binary64 argument,
 * round-to-nearest-even, for |x| <= 3.5 (the element integrals evaluate cos
 * at chart latitudes, |x| <= pi/2).
 *
 * The reference evaluates the cos weight of its element integrals with the
 * C library's cos (src/util.cpp:194-215, f_cos / f_1_cos).  The mass
 * entries cancel about 1e7-fold, so a one-ulp difference in cos moves them
 * by ~1e-10 relative; the only definition of cos that no library build can
 * disagree with is the correctly rounded one, and that is what this header
 * computes, on the host (gcc or hipcc) and on gfx950 alike.  One header
 * serves the device kernels and the CPU restatement.
 *
 * Method: Ziv's strategy (A. Ziv, "Fast evaluation of elementary
 * mathematical functions with correctly rounded last bit", ACM TOMS 17(3),
 * 1991).
 *   1. Fast path in double-double arithmetic (Dekker's two-product via
 *      fma, Knuth's two-sum):
 *        - Cody-Waite reduction r = x - k pi/2, k = rint(2x/pi) in
 *          {-2..2}, with pi/2 as a sum of four doubles; x - k*P1 is exact
 *          (both are multiples of 2^-53 and the difference is below 1), so
 *          r carries a relative error below 2^-105;
 *        - Taylor series of sin(r) and cos(r) on |r| <= 0.8 with
 *          double-double coefficients 1/n! (terms to degree 27 / 26: the
 *          first omitted term is below 2^-107 relative), the small terms
 *          of degree >= 19 in plain double;
 *      relative error of the double-double result below 2^-100 (measured
 *      maximum on random and assembly arguments about 2^-104).
 *   2. Rounding test: with the result h + l (|l| <= ulp(h)/2) and the bound
 *      e = 2^-95 |h|, RN(h + (l - e)) == RN(h + (l + e)) proves that every
 *      value within the bound, the exact cos among them, rounds to the same
 *      double (round-to-nearest is monotonic).  It fails for about one
 *      argument in 2^42.
 *   3. Slow path (only then): 256-bit fixed-point arithmetic (value =
 *      N 2^-254), exact reduction |x - k pi/2| against a 256-bit pi/2, and
 *      Taylor series to degree 56 (cos) or 57 (sin), terms built by the recurrence
 *      t_{n+2} = t_n r^2 / ((n+1)(n+2)); absolute error below 2^-240; the
 *      result is rounded to nearest from its bits.  cos and sin of a
 *      non-zero double are transcendental (Lindemann), never a midpoint, so
 *      a finite precision decides every case; 2^-240 is far beyond it.
 *
 * Outside |x| <= 3.5 the same code runs with a reduction that is no longer
 * exact (k up to 2^20; faithful, not guaranteed correctly rounded).
 *
 * Plain C99 (with unsigned __int128), compiled with -ffp-contract=off; under
 * hipcc the functions are __host__ __device__.
 */
#ifndef PHG_CR_TRIG_H
#define PHG_CR_TRIG_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define CRT_HD static inline __host__ __device__
#define CRT_NOINLINE static __host__ __device__ __attribute__((noinline))
#else
#define CRT_HD static inline
#define CRT_NOINLINE static __attribute__((noinline))
#endif

/* pi/2 = P1 + P2 + P3 + P4 (each the double nearest to what is left) */
#define CRT_P1 0x1.921fb54442d18p+0
#define CRT_P2 0x1.1a62633145c07p-54
#define CRT_P3 (-0x1.f1976b7ed8fbcp-110)
#define CRT_P4 0x1.4cf98e804177dp-164
#define CRT_TWO_OVER_PI 0x1.45f306dc9c883p-1
/* relative error bound of the double-double result used by the rounding test */
#define CRT_FAST_ERR 0x1p-95

/* ---------------------------------------------------- double-double ops -- */
typedef struct {
    double hi, lo;
} crt_dd;

CRT_HD crt_dd crt_two_sum(double a, double b)
{
    crt_dd r;
    r.hi = a + b;
    const double bb = r.hi - a;
    r.lo = (a - (r.hi - bb)) + (b - bb);
    return r;
}

/* |a| >= |b| (or a == 0) */
CRT_HD crt_dd crt_fast_two_sum(double a, double b)
{
    crt_dd r;
    r.hi = a + b;
    r.lo = b - (r.hi - a);
    return r;
}

CRT_HD crt_dd crt_dd_mul(crt_dd a, crt_dd b)
{
    const double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return crt_fast_two_sum(p, e);
}

CRT_HD crt_dd crt_dd_add(crt_dd a, crt_dd b)
{
    crt_dd s = crt_two_sum(a.hi, b.hi);
    const crt_dd t = crt_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = crt_fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return crt_fast_two_sum(s.hi, s.lo);
}

/* (-1)^(n/2) / n! for n = 2..27 as double-double pairs (hi, lo), computed
 * with 600-bit arithmetic; the CPU tests recompute every pair */
#define CRT_C2 (-0x1.0000000000000p-1), 0.0
#define CRT_C3 (-0x1.5555555555555p-3), (-0x1.5555555555555p-57)
#define CRT_C4 0x1.5555555555555p-5, 0x1.5555555555555p-59
#define CRT_C5 0x1.1111111111111p-7, 0x1.1111111111111p-63
#define CRT_C6 (-0x1.6c16c16c16c17p-10), 0x1.f49f49f49f49fp-65
#define CRT_C7 (-0x1.a01a01a01a01ap-13), (-0x1.a01a01a01a01ap-73)
#define CRT_C8 0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76
#define CRT_C9 0x1.71de3a556c734p-19, (-0x1.c154f8ddc6c00p-73)
#define CRT_C10 (-0x1.27e4fb7789f5cp-22), (-0x1.cbbc05b4fa99ap-76)
#define CRT_C11 (-0x1.ae64567f544e4p-26), 0x1.c062e06d1f209p-80
#define CRT_C12 0x1.1eed8eff8d898p-29, (-0x1.2aec959e14c06p-83)
#define CRT_C13 0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87
#define CRT_C14 (-0x1.93974a8c07c9dp-37), (-0x1.05d6f8a2efd1fp-92)
#define CRT_C15 (-0x1.ae7f3e733b81fp-41), (-0x1.1d8656b0ee8cbp-97)
#define CRT_C16 0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101
#define CRT_C17 0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103
/* degree >= 18: plain doubles */
#define CRT_D18 (-0x1.6827863b97d97p-53)
#define CRT_D19 (-0x1.2f49b46814157p-57)
#define CRT_D20 0x1.e542ba4020225p-62
#define CRT_D21 0x1.71b8ef6dcf572p-66
#define CRT_D22 (-0x1.0ce396db7f853p-70)
#define CRT_D23 (-0x1.761b41316381ap-75)
#define CRT_D24 0x1.f2cf01972f578p-80
#define CRT_D25 0x1.3f3ccdd165fa9p-84
#define CRT_D26 (-0x1.88e85fc6a4e5ap-89)
#define CRT_D27 (-0x1.d1ab1c2dccea3p-94)

CRT_HD crt_dd crt_mk(double hi, double lo)
{
    crt_dd c;
    c.hi = hi;
    c.lo = lo;
    return c;
}

/* s = c + z s in double-double */
#define CRT_HORNER(C) s = crt_dd_add(crt_mk(C), crt_dd_mul(z, s))

CRT_HD crt_dd crt_square(crt_dd r)
{
    crt_dd z;
    z.hi = r.hi * r.hi;
    z.lo = fma(r.hi, r.hi, -z.hi) + 2.0 * r.hi * r.lo;
    return crt_fast_two_sum(z.hi, z.lo);
}

/* sin(r) = r + r z S(z), z = r^2, |r| <= 0.8: S(z) to the degree-27 term
 * (the first omitted term is below 2^-111 relative); the terms of degree
 * >= 19 (below 2^-62 relative) in plain double */
CRT_HD crt_dd crt_poly_sin(crt_dd r)
{
    const crt_dd z = crt_square(r);
    double q = CRT_D27;
    q = fma(q, z.hi, CRT_D25);
    q = fma(q, z.hi, CRT_D23);
    q = fma(q, z.hi, CRT_D21);
    q = fma(q, z.hi, CRT_D19);
    crt_dd s = crt_mk(q, 0.0);
    CRT_HORNER(CRT_C17);
    CRT_HORNER(CRT_C15);
    CRT_HORNER(CRT_C13);
    CRT_HORNER(CRT_C11);
    CRT_HORNER(CRT_C9);
    CRT_HORNER(CRT_C7);
    CRT_HORNER(CRT_C5);
    CRT_HORNER(CRT_C3);
    const crt_dd t = crt_dd_mul(r, crt_dd_mul(z, s));
    return crt_dd_add(r, t);
}

/* cos(r) = 1 + z C(z): to the degree-26 term (first omitted below 2^-106
 * relative), degree >= 18 in plain double */
CRT_HD crt_dd crt_poly_cos(crt_dd r)
{
    const crt_dd z = crt_square(r);
    double q = CRT_D26;
    q = fma(q, z.hi, CRT_D24);
    q = fma(q, z.hi, CRT_D22);
    q = fma(q, z.hi, CRT_D20);
    q = fma(q, z.hi, CRT_D18);
    crt_dd s = crt_mk(q, 0.0);
    CRT_HORNER(CRT_C16);
    CRT_HORNER(CRT_C14);
    CRT_HORNER(CRT_C12);
    CRT_HORNER(CRT_C10);
    CRT_HORNER(CRT_C8);
    CRT_HORNER(CRT_C6);
    CRT_HORNER(CRT_C4);
    CRT_HORNER(CRT_C2);
    return crt_dd_add(crt_mk(1.0, 0.0), crt_dd_mul(z, s));
}
#undef CRT_HORNER

/* r = x - k pi/2 as a double-double and the quadrant k mod 4 */
CRT_HD crt_dd crt_reduce(double x, int *quadrant)
{
    crt_dd r;
    const double kd = rint(x * CRT_TWO_OVER_PI);
    if (kd == 0.0) {
        *quadrant = 0;
        r.hi = x;
        r.lo = 0.0;
        return r;
    }
    const long long k = (long long)kd;
    *quadrant = (int)(k & 3);
    if (fabs(kd) <= 2.0) {
        /* k*P1, k*P2, k*P3, k*P4 are exact; t = x - k*P1 is exact */
        const double t = x - kd * CRT_P1;
        crt_dd s = crt_two_sum(t, -kd * CRT_P2);
        double e = s.lo - kd * CRT_P3;
        e -= kd * CRT_P4;
        return crt_fast_two_sum(s.hi, e);
    }
    /* larger k: Cody-Waite with fma-exact products, no exactness claim */
    const double p1 = kd * CRT_P1, p1e = fma(kd, CRT_P1, -p1);
    const double p2 = kd * CRT_P2, p2e = fma(kd, CRT_P2, -p2);
    crt_dd s = crt_two_sum(x, -p1);
    crt_dd c = crt_two_sum(p1e, p2);
    c.lo += p2e + kd * CRT_P3;
    return crt_dd_add(s, crt_mk(-c.hi, -c.lo));
}

/* ------------------------------------------------- 256-bit fixed point -- */
/* unsigned value N 2^-254, N = w[0] + w[1] 2^64 + w[2] 2^128 + w[3] 2^192 */
typedef struct {
    uint64_t w[4];
} crt_fx;

CRT_HD crt_fx crt_fx_zero(void)
{
    crt_fx a;
    a.w[0] = a.w[1] = a.w[2] = a.w[3] = 0;
    return a;
}

CRT_HD crt_fx crt_fx_add(crt_fx a, crt_fx b)
{
    crt_fx r;
    unsigned __int128 c = 0;
    for (int i = 0; i < 4; ++i) {
        c += (unsigned __int128)a.w[i] + b.w[i];
        r.w[i] = (uint64_t)c;
        c >>= 64;
    }
    return r;
}

/* a - b, a >= b */
CRT_HD crt_fx crt_fx_sub(crt_fx a, crt_fx b)
{
    crt_fx r;
    uint64_t borrow = 0;
    for (int i = 0; i < 4; ++i) {
        const uint64_t d = a.w[i] - b.w[i];
        const uint64_t d2 = d - borrow;
        borrow = (a.w[i] < b.w[i]) || (d < borrow);
        r.w[i] = d2;
    }
    return r;
}

CRT_HD int crt_fx_less(crt_fx a, crt_fx b)
{
    for (int i = 3; i >= 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i];
    return 0;
}

/* (a b) 2^-254, truncated: bits 254.. of the 512-bit product */
CRT_HD crt_fx crt_fx_mul(crt_fx a, crt_fx b)
{
    uint64_t p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        unsigned __int128 c = 0;
        for (int j = 0; j < 4; ++j) {
            c += (unsigned __int128)a.w[i] * b.w[j] + p[i + j];
            p[i + j] = (uint64_t)c;
            c >>= 64;
        }
        p[i + 4] = (uint64_t)c;
    }
    /* shift right by 254 = 3 * 64 + 62 */
    crt_fx r;
    for (int i = 0; i < 4; ++i) {
        const uint64_t lo = p[i + 3], hi = (i + 4 < 8) ? p[i + 4] : 0;
        r.w[i] = (lo >> 62) | (hi << 2);
    }
    return r;
}

/* a / d for a small integer d, truncated (long division by 32-bit halves) */
CRT_HD crt_fx crt_fx_div_small(crt_fx a, uint32_t d)
{
    crt_fx r;
    uint64_t rem = 0;
    for (int i = 3; i >= 0; --i) {
        const uint64_t hi = (rem << 32) | (a.w[i] >> 32);
        const uint64_t qh = hi / d;
        rem = hi % d;
        const uint64_t lo = (rem << 32) | (a.w[i] & 0xffffffffu);
        const uint64_t ql = lo / d;
        rem = lo % d;
        r.w[i] = (qh << 32) | ql;
    }
    return r;
}

/* |x| (|x| < 4, |x| >= 2^-190 or 0) as N 2^-254, exactly */
CRT_HD crt_fx crt_fx_from_double(double x)
{
    crt_fx r = crt_fx_zero();
    x = fabs(x);
    if (x == 0.0) return r;
    int e;
    const double m = frexp(x, &e);                       /* x = m 2^e, m in [0.5, 1) */
    const uint64_t mant = (uint64_t)ldexp(m, 53);         /* 53-bit integer */
    const int sh = e - 53 + 254;                          /* x = mant 2^(sh - 254) */
    if (sh < 0) {
        r.w[0] = mant >> (-sh);
        return r;
    }
    const int limb = sh >> 6, bit = sh & 63;
    r.w[limb] = mant << bit;
    if (bit && limb + 1 < 4) r.w[limb + 1] = mant >> (64 - bit);
    return r;
}

/* N 2^-254 rounded to the nearest double (ties to even) */
CRT_HD double crt_fx_to_double(crt_fx a)
{
    int top = -1;
    for (int i = 3; i >= 0 && top < 0; --i)
        if (a.w[i]) {
            int b = 63;
            while (!((a.w[i] >> b) & 1)) --b;
            top = 64 * i + b;
        }
    if (top < 0) return 0.0;
    /* bit index of the lowest mantissa bit */
    const int low = top - 52;
    if (low <= 0) {
        const double v = (double)a.w[0];
        return ldexp(v, -254);
    }
#define CRT_BIT(idx) ((a.w[(idx) >> 6] >> ((idx) & 63)) & 1)
    uint64_t mant = 0;
    for (int b = top; b >= low; --b) mant = (mant << 1) | CRT_BIT(b);
    const int round_bit = (int)CRT_BIT(low - 1);
    int sticky = 0;
    for (int b = low - 2; b >= 0 && !sticky; --b) sticky = (int)CRT_BIT(b);
#undef CRT_BIT
    if (round_bit && (sticky || (mant & 1))) mant += 1;
    return ldexp((double)mant, low - 254);
}

/* sin(x) (odd = 1) or cos(x) (odd = 0) by the fixed-point series */
CRT_NOINLINE double crt_slow(double x, int odd)
{
    /* |x| < 2^-100: cos x = 1 - x^2/2 and sin x = x (1 - x^2/6) round to 1
     * and x (the deviation is below 2^-199 relative) */
    if (fabs(x) < 0x1p-100) return odd ? x : 1.0;
    /* pi/2 truncated to 254 fractional bits */
    crt_fx pio2;
    pio2.w[0] = 0x105df531d89cd91ull;
    pio2.w[1] = 0x948127044533e63aull;
    pio2.w[2] = 0x62633145c06e0e68ull;
    pio2.w[3] = 0x6487ed5110b4611aull;
    const double kd = rint(x * CRT_TWO_OVER_PI);
    const long long k = (long long)kd;
    const crt_fx ax = crt_fx_from_double(x);
    /* y = x - k pi/2 = sign(x) (|x| - |k| pi/2); R = |y|, ys = sign of y */
    crt_fx kp = crt_fx_zero();
    for (long long i = 0; i < (k < 0 ? -k : k); ++i) kp = crt_fx_add(kp, pio2);
    crt_fx R;
    int ys = x < 0.0 ? -1 : 1;
    if (crt_fx_less(ax, kp)) {
        R = crt_fx_sub(kp, ax);
        ys = -ys;
    } else {
        R = crt_fx_sub(ax, kp);
    }
    const int q = (int)(k & 3);
    /* f(x) = +-cos(y) or +-sin(y) by the quadrant */
    const int use_sin = odd ? !(q & 1) : (q & 1);
    int sign = 1;
    if (odd) {
        if (q == 2 || q == 3) sign = -1;
    } else {
        if (q == 1 || q == 2) sign = -1;
    }
    if (use_sin) sign *= ys; /* sin is odd in y; cos is even */
    /* series in R: terms t_n = R^n / n!, n = 0/1 .. 55, alternating */
    const crt_fx R2 = crt_fx_mul(R, R);
    crt_fx one = crt_fx_zero();
    one.w[3] = 1ull << 62; /* 2^254 2^-254 = 1 */
    crt_fx term = use_sin ? R : one;
    crt_fx pos = term, neg = crt_fx_zero();
    int n = use_sin ? 1 : 0;
    for (int j = 1; n + 2 <= 57; ++j) {
        term = crt_fx_mul(term, R2);
        term = crt_fx_div_small(term, (uint32_t)((n + 1) * (n + 2)));
        n += 2;
        if (j & 1) neg = crt_fx_add(neg, term);
        else pos = crt_fx_add(pos, term);
    }
    const crt_fx v = crt_fx_sub(pos, neg);
    const double d = crt_fx_to_double(v);
    return sign < 0 ? -d : d;
}

/* ----------------------------------------------------------- entries ---- */
/* The fast path with the rounding test at relative bound err; returns 1 and
 * the rounded result in *out when the test passes. */
CRT_HD int crt_fast(double x, int odd, double err, double *out)
{
    int q;
    const crt_dd r = crt_reduce(x, &q);
    const int use_sin = odd ? !(q & 1) : (q & 1);
    const crt_dd v = use_sin ? crt_poly_sin(r) : crt_poly_cos(r);
    const double e = 2.0 * err * fabs(v.hi);
    const double a = v.hi + (v.lo - e);
    const double b = v.hi + (v.lo + e);
    int neg;
    if (odd) neg = (q == 2 || q == 3);
    else neg = (q == 1 || q == 2);
    *out = neg ? -a : a;
    return a == b;
}

/* beyond |x| <= 3.5 the fixed-point path does not apply: the fast result is
 * returned as it is */
CRT_HD double cr_cos_bound(double x, double err)
{
    double f;
    if (crt_fast(x, 0, err, &f) || !(fabs(x) <= 3.5)) return f;
    return crt_slow(x, 0);
}

CRT_HD double cr_sin_bound(double x, double err)
{
    if (x == 0.0) return x; /* keeps the sign of zero */
    double f;
    if (crt_fast(x, 1, err, &f) || !(fabs(x) <= 3.5)) return f;
    return crt_slow(x, 1);
}

CRT_HD double cr_cos(double x) { return cr_cos_bound(x, CRT_FAST_ERR); }
CRT_HD double cr_sin(double x) { return cr_sin_bound(x, CRT_FAST_ERR); }

#endif /* PHG_CR_TRIG_H */
