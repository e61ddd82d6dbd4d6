/*
 * orc_math.h -- TEST INFRASTRUCTURE (oracle). Not part of the product.
 *
 * Deterministic FP64 elementary functions and the Philox4x32-10 counter-based
 * generator used by the CPU oracle.  Every routine uses only IEEE-754 +, -, *,
 * /, sqrt and explicit fma, evaluated in a fixed order, so that a GPU kernel
 * that evaluates the same recipe with FP contraction disabled produces the
 * same bits.  The recipes are textbook ones, derived from the mathematics:
 *   exp   : x = k ln2 + r (two-part ln2, Cody-Waite), Taylor series of e^r to
 *           degree 14, scaling by 2^k through the exponent field.
 *   log   : x = m 2^e with m in [1/sqrt2, sqrt2), log m = 2 atanh(s),
 *           s = (m-1)/(m+1), odd series to s^23.
 *   sinpi/cospi : exact reduction x = k/2 + r, |r| <= 1/4, Taylor series of
 *           sin and cos of pi*r to degree 17 / 18.
 *   sin/cos : sinpi/cospi of x/pi with x/pi formed from a two-part 1/pi.
 *   tanh  : expm1(2x) / (expm1(2x) + 2), expm1 by Taylor for |y| <= 1/2.
 * The reference (BVPs/EDPs/Monegros_Poisson_CPP.hpp:12-59,
 * Integrators/FKACIntegrator.hpp:15-19) calls libm and boost normals; the
 * oracle replaces both by these reproducible recipes.  Accuracy against libm
 * is checked in tests/test_oracle_math.py.
 *
 * Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11 (multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds).  Pinned by the published
 * known-answer vectors in tests/test_oracle_math.py.
 */
#ifndef ORC_MATH_H
#define ORC_MATH_H
#include <stdint.h>
#include <string.h>
#include <math.h>

static inline uint64_t orc_bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double orc_from_bits(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

#define ORC_LN2_HI 0.6931467056274414
#define ORC_LN2_LO 4.7493250390316726e-07
#define ORC_LOG2E 1.4426950408889634
#define ORC_LN2 0.6931471805599453
#define ORC_PI 3.141592653589793
#define ORC_PI_LO 1.2246467991473532e-16
#define ORC_INVPI_HI 0.3183098861837907
#define ORC_INVPI_LO (-1.9678676675182486e-17)

/* 2^k for integer k in [-1022, 1023] */
static inline double orc_pow2i(int k) { return orc_from_bits((uint64_t)(k + 1023) << 52); }

static inline double orc_exp_poly(double r) {
    /* sum_{n=0}^{14} r^n / n!  (Horner with fma) */
    double p = 1.1470745597729725e-11;
    p = fma(p, r, 1.6059043836821613e-10);
    p = fma(p, r, 2.08767569878681e-09);
    p = fma(p, r, 2.505210838544172e-08);
    p = fma(p, r, 2.755731922398589e-07);
    p = fma(p, r, 2.7557319223985893e-06);
    p = fma(p, r, 2.48015873015873e-05);
    p = fma(p, r, 0.0001984126984126984);
    p = fma(p, r, 0.001388888888888889);
    p = fma(p, r, 0.008333333333333333);
    p = fma(p, r, 0.041666666666666664);
    p = fma(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return p;
}

static inline double orc_exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.2) return 0.0;
    double k = floor(x * ORC_LOG2E + 0.5);
    double r = fma(-k, ORC_LN2_HI, x);
    r = fma(-k, ORC_LN2_LO, r);
    double p = orc_exp_poly(r);
    int ki = (int)k;
    if (ki > 1023) return (p * 2.0) * orc_pow2i(ki - 1);
    if (ki >= -1022) return p * orc_pow2i(ki);
    return (p * orc_pow2i(ki + 600)) * orc_pow2i(-600);
}

/* exp(x) for x <= 0 by Tang's table-driven method (ACM TOMS 15(2), 1989): 2^m 2^(j/64) e^r with
   |r| <= ln2/128, e^r - 1 by its degree-5 Taylor polynomial and 2^(j/64) correctly rounded from a
   64-entry table; the exponential of the engine's RBF reduction, same operations. */
static const double orc_exp2_tab[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};
static inline double orc_exp_tab(double x) {
    /* results below the normal range are flushed to zero, so 2^m stays normal */
    double xc = x < -708.0 ? -708.0 : x;
    double k = rint(xc * 0x1.71547652b82fep+6); /* round half to even */
    double r = fma(-k, 0x1.62e42fefa3800p-7, xc);
    r = fma(-k, 0x1.ef35793c76730p-51, r);
    double p = fma(r, 0.008333333333333333, 0.041666666666666664);
    p = fma(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = p * r;
    int ki = (int)k;
    double t = orc_exp2_tab[ki & 63];
    double v = fma(t, p, t);
    double y = v * orc_pow2i(ki >> 6);
    return x < -708.0 ? 0.0 : y;
}

static inline double orc_expm1(double x) {
    if (fabs(x) <= 0.5) {
        /* sum_{n=1}^{17} x^n / n! */
        double p = 2.8114572543455206e-15; /* 1/17! */
        p = fma(p, x, 4.779477332387385e-14);   /* 1/16! */
        p = fma(p, x, 7.647163731819816e-13);   /* 1/15! */
        p = fma(p, x, 1.1470745597729725e-11);  /* 1/14! */
        p = fma(p, x, 1.6059043836821613e-10);
        p = fma(p, x, 2.08767569878681e-09);
        p = fma(p, x, 2.505210838544172e-08);
        p = fma(p, x, 2.755731922398589e-07);
        p = fma(p, x, 2.7557319223985893e-06);
        p = fma(p, x, 2.48015873015873e-05);
        p = fma(p, x, 0.0001984126984126984);
        p = fma(p, x, 0.001388888888888889);
        p = fma(p, x, 0.008333333333333333);
        p = fma(p, x, 0.041666666666666664);
        p = fma(p, x, 0.16666666666666666);
        p = fma(p, x, 0.5);
        p = fma(p, x, 1.0);
        return p * x;
    }
    return orc_exp(x) - 1.0;
}

static inline double orc_log(double x) {
    if (!(x > 0.0)) return (x == 0.0) ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    uint64_t b = orc_bits(x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) { /* subnormal */
        x *= 18014398509481984.0; /* 2^54 */
        b = orc_bits(x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = orc_from_bits((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    /* s^2 <= 0.0295: the atanh series terms past s^23 are below 1e-2 ulp */
    double p = 0.08695652173913043;          /* 2/23 */
    p = fma(p, s2, 0.09523809523809523);     /* 2/21 */
    p = fma(p, s2, 0.10526315789473684);     /* 2/19 */
    p = fma(p, s2, 0.11764705882352941);     /* 2/17 */
    p = fma(p, s2, 0.13333333333333333);     /* 2/15 */
    p = fma(p, s2, 0.15384615384615385);     /* 2/13 */
    p = fma(p, s2, 0.18181818181818182);     /* 2/11 */
    p = fma(p, s2, 0.2222222222222222);      /* 2/9 */
    p = fma(p, s2, 0.2857142857142857);      /* 2/7 */
    p = fma(p, s2, 0.4);                     /* 2/5 */
    p = fma(p, s2, 0.6666666666666666);      /* 2/3 */
    p = fma(p, s2, 2.0);
    double lm = s * p;
    double de = (double)e;
    return fma(de, ORC_LN2_HI, fma(de, ORC_LN2_LO, lm));
}

/* sin(pi r), cos(pi r) for |r| <= 1/4 */
static inline void orc_sincospi_kernel(double r, double *s, double *c) {
    double y = fma(r, ORC_PI, r * ORC_PI_LO);
    double y2 = y * y;
    /* |y| <= pi/4: the Taylor terms past y^17 (sin) and y^18 (cos) are below 1e-3 ulp */
    double ps = 2.8114572543455206e-15;      /* 1/17! */
    ps = fma(ps, y2, -7.647163731819816e-13);
    ps = fma(ps, y2, 1.6059043836821613e-10);
    ps = fma(ps, y2, -2.505210838544172e-08);
    ps = fma(ps, y2, 2.7557319223985893e-06);
    ps = fma(ps, y2, -0.0001984126984126984);
    ps = fma(ps, y2, 0.008333333333333333);
    ps = fma(ps, y2, -0.16666666666666666);
    *s = fma(ps * y2, y, y);
    double pc = -1.5619206968586225e-16;     /* -1/18! */
    pc = fma(pc, y2, 4.779477332387385e-14);
    pc = fma(pc, y2, -1.1470745597729725e-11);
    pc = fma(pc, y2, 2.08767569878681e-09);
    pc = fma(pc, y2, -2.755731922398589e-07);
    pc = fma(pc, y2, 2.48015873015873e-05);
    pc = fma(pc, y2, -0.001388888888888889);
    pc = fma(pc, y2, 0.041666666666666664);
    pc = fma(pc, y2, -0.5);
    *c = fma(pc, y2, 1.0);
}

/* sin(pi x) and cos(pi x) together; exact argument reduction */
static inline void orc_sincospi(double x, double *sp, double *cp) {
    if (!(fabs(x) < 4503599627370496.0)) { /* |x| >= 2^52: x is an integer (or inf/nan) */
        if (x != x || fabs(x) == INFINITY) { *sp = NAN; *cp = NAN; return; }
        *sp = 0.0 * x;
        *cp = (fmod(x, 2.0) == 0.0) ? 1.0 : -1.0;
        return;
    }
    double k = floor(2.0 * x + 0.5);
    double r = x - 0.5 * k;
    double s, c;
    orc_sincospi_kernel(r, &s, &c);
    int64_t q = ((int64_t)k) & 3;
    switch (q) {
        case 0: *sp = s; *cp = c; break;
        case 1: *sp = c; *cp = -s; break;
        case 2: *sp = -s; *cp = -c; break;
        default: *sp = -c; *cp = s; break;
    }
}
static inline double orc_sinpi(double x) { double s, c; orc_sincospi(x, &s, &c); return s; }
static inline double orc_cospi(double x) { double s, c; orc_sincospi(x, &s, &c); return c; }
static inline double orc_x_over_pi(double x) { return fma(x, ORC_INVPI_HI, x * ORC_INVPI_LO); }
static inline double orc_sin(double x) { return orc_sinpi(orc_x_over_pi(x)); }
static inline double orc_cos(double x) { return orc_cospi(orc_x_over_pi(x)); }
static inline double orc_tanh(double x) {
    if (x > 20.0) return 1.0;
    if (x < -20.0) return -1.0;
    double e = orc_expm1(2.0 * x);
    return e / (e + 2.0);
}

/* ---------------- Philox4x32-10 ---------------- */
static inline void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in (0,1) from two 32-bit words: ((a<<21 | b>>11) + 0.5) * 2^-53 */
static inline double orc_u01_53(uint32_t a, uint32_t b) {
    uint64_t m = ((uint64_t)a << 21) | (uint64_t)(b >> 11);
    return ((double)m + 0.5) * 1.1102230246251565e-16;
}
/* 24-bit uniform in (0,1) from one word: ((a>>8) + 0.5) * 2^-24 (exact in float) */
static inline double orc_u01_24(uint32_t a) {
    return ((double)(a >> 8) + 0.5) * 5.9604644775390625e-08;
}

/* Box-Muller on (u1,u2): n1 = r cos(2 pi u2), n2 = r sin(2 pi u2), r = sqrt(-2 log u1) */
static inline void orc_box_muller(double u1, double u2, double *n1, double *n2) {
    double r = sqrt(-2.0 * orc_log(u1));
    double s, c;
    orc_sincospi(2.0 * u2, &s, &c);
    *n1 = r * c;
    *n2 = r * s;
}

#endif
