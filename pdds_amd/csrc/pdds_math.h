// pdds_math.h -- device/host arithmetic for the path engine.
//
// FP64: reproducible elementary functions built only from IEEE +,-,*,/,sqrt
// and explicit fma in a fixed order (TUs that include this for FP64 are
// compiled with -ffp-contract=off), so every FP64 path is a fixed function of
// its Philox stream.  Recipes (textbook): exp by two-part ln2 reduction and
// the Taylor series of e^r to degree 14; log by 2 atanh((m-1)/(m+1)) to
// s^23; sinpi/cospi by exact reduction to |r| <= 1/4 and Taylor series to
// degree 17/18; sin/cos through a two-part 1/pi; tanh through expm1.
// FP32: the CDNA hardware transcendentals (v_exp_f32, v_log_f32, v_sin_f32,
// v_cos_f32 take base-2 / revolution arguments) and v_rcp_f32.
//
// Philox4x32-10 after Salmon et al., SC'11 (public multipliers and Weyl
// constants), 64-bit products through v_mad_u64_u32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PDDS_HD __host__ __device__ __forceinline__

namespace pdds {

// Stopping-domain geometry of a mesh's stencils: rectangles, or annular sectors of a disc.
enum Geom { GeomRect = 0, GeomDisc = 1 };

// ------------------------------------------------------------------ FP64 --
PDDS_HD uint64_t f64_bits(double x) { return (uint64_t)__builtin_bit_cast(uint64_t, x); }
PDDS_HD double f64_from(uint64_t u) { return __builtin_bit_cast(double, u); }
PDDS_HD double pow2i(int k) { return f64_from((uint64_t)(k + 1023) << 52); }

constexpr double kLn2Hi = 0.6931467056274414;
constexpr double kLn2Lo = 4.7493250390316726e-07;
constexpr double kLog2e = 1.4426950408889634;
constexpr double kPi = 3.141592653589793;
constexpr double kPiLo = 1.2246467991473532e-16;
constexpr double kInvPiHi = 0.3183098861837907;
constexpr double kInvPiLo = -1.9678676675182486e-17;

PDDS_HD double exp_poly(double r) {
    double p = 1.1470745597729725e-11;
    p = __builtin_fma(p, r, 1.6059043836821613e-10);
    p = __builtin_fma(p, r, 2.08767569878681e-09);
    p = __builtin_fma(p, r, 2.505210838544172e-08);
    p = __builtin_fma(p, r, 2.755731922398589e-07);
    p = __builtin_fma(p, r, 2.7557319223985893e-06);
    p = __builtin_fma(p, r, 2.48015873015873e-05);
    p = __builtin_fma(p, r, 0.0001984126984126984);
    p = __builtin_fma(p, r, 0.001388888888888889);
    p = __builtin_fma(p, r, 0.008333333333333333);
    p = __builtin_fma(p, r, 0.041666666666666664);
    p = __builtin_fma(p, r, 0.16666666666666666);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return p;
}

PDDS_HD double dexp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return __builtin_huge_val();
    if (x < -745.2) return 0.0;
    double k = __builtin_floor(x * kLog2e + 0.5);
    double r = __builtin_fma(-k, kLn2Hi, x);
    r = __builtin_fma(-k, kLn2Lo, r);
    double p = exp_poly(r);
    int ki = (int)k;
    if (ki > 1023) return (p * 2.0) * pow2i(ki - 1);
    if (ki >= -1022) return p * pow2i(ki);
    return (p * pow2i(ki + 600)) * pow2i(-600);
}

// fma(p, r, c) with the constant c taken from a scalar register pair: one v_fma_f64 per Horner
// step.  (Left to itself the compiler keeps the coefficients of a loop-resident polynomial in
// vector registers and evaluates each step as a copy plus a two-address v_fmac_f64.)
PDDS_HD double fma_k(double p, double r, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    double d;
    __asm__("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(p), "v"(r), "s"(c));
    return d;
#else
    return __builtin_fma(p, r, c);
#endif
}

// exp_poly with one instruction per Horner step (same coefficients, same fused operations)
PDDS_HD double exp_poly_k(double r) {
    double p = 1.1470745597729725e-11;
    p = fma_k(p, r, 1.6059043836821613e-10);
    p = fma_k(p, r, 2.08767569878681e-09);
    p = fma_k(p, r, 2.505210838544172e-08);
    p = fma_k(p, r, 2.755731922398589e-07);
    p = fma_k(p, r, 2.7557319223985893e-06);
    p = fma_k(p, r, 2.48015873015873e-05);
    p = fma_k(p, r, 0.0001984126984126984);
    p = fma_k(p, r, 0.001388888888888889);
    p = fma_k(p, r, 0.008333333333333333);
    p = fma_k(p, r, 0.041666666666666664);
    p = fma_k(p, r, 0.16666666666666666);
    p = fma_k(p, r, 0.5);
    p = fma_k(p, r, 1.0);
    p = fma_k(p, r, 1.0);
    return p;
}

// dexp(x) for |x| <= 708, where its result is normal (k in [-1021, 1022]): the same steps without
// the special-case tests and scalings.  Its polynomial keeps the two-instruction Horner steps:
// with exp_poly_k as well, the LDS-tiled FP64 path kernel holds more coefficients in scalar
// registers than it has, spills them and runs 3 % slower.
PDDS_HD double dexp_core(double x) {
    const double k = __builtin_floor(x * kLog2e + 0.5);
    double r = __builtin_fma(-k, kLn2Hi, x);
    r = __builtin_fma(-k, kLn2Lo, r);
    return exp_poly(r) * pow2i((int)k);
}

// dexp(x) bit for bit: the short path where the result is normal, a branch to dexp elsewhere
PDDS_HD double dexp_fast(double x) {
    if (__builtin_expect(__builtin_fabs(x) <= 708.0, 1)) return dexp_core(x);
    return dexp(x);
}

// dexp(x) for x <= 0, bit for bit, without branches: the value of dexp's own steps, with the
// underflow (x < -745.2 -> 0) and subnormal scalings as selects.  The RBF moments of the
// reduction evaluate one of these per (exit, stencil entry).
PDDS_HD double dexp_nonpos(double x) {
    const bool under = x < -745.2;
    const double xc = under ? 0.0 : x;
    const double k = __builtin_floor(xc * kLog2e + 0.5);
    double r = __builtin_fma(-k, kLn2Hi, xc);
    r = __builtin_fma(-k, kLn2Lo, r);
    const double p = exp_poly_k(r);
    const int ki = (int)k;                       // in [-1075, 0]
    const bool sub = ki < -1022;
    const double v = p * pow2i(sub ? ki + 600 : ki);
    const double w = sub ? v * pow2i(-600) : v;
    return under ? 0.0 : w;
}

// exp(x) for x <= 0 by Tang's table-driven method (P. T. P. Tang, ACM TOMS 15(2), 1989):
// x = (64 m + j) ln2/64 + r with |r| <= ln2/128, exp(x) = 2^m 2^(j/64) e^r, e^r - 1 by its
// degree-5 Taylor polynomial (truncation below 0.35 ulp), 2^(j/64) correctly rounded from the
// table T (kExp2Tab, staged in LDS by the reduction).  Within 1.5 ulp of exp; about half the
// operations of dexp_nonpos.  Underflow and subnormal results are handled as in dexp_nonpos.
#define PDDS_EXP2_TAB { \
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0, \
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0, \
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0, \
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0, \
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0, \
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0, \
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0, \
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0, \
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0, \
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0, \
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0, \
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0, \
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0, \
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0, \
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0, \
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0, \
}
constexpr double kInvLn2x64 = 0x1.71547652b82fep+6;     // 64 / ln 2
constexpr double kLn2d64Hi = 0x1.62e42fefa3800p-7;      // ln 2 / 64, 42 significant bits
constexpr double kLn2d64Lo = 0x1.ef35793c76730p-51;     // ln 2 / 64 - kLn2d64Hi
// Results below the normal range (x < -708) are flushed to zero: an RBF moment that small lies
// far below the rounding of any sum it joins, and the flush keeps 2^m normal (m >= -1022), so
// the scale is a single multiply by a power of two built from its exponent bits.
PDDS_HD double dexp_tab_nonpos(double x, const double *T) {
    const double xc = __builtin_fmax(x, -708.0);
    const double k = __builtin_rint(xc * kInvLn2x64);   // round half to even
    double r = __builtin_fma(-k, kLn2d64Hi, xc);
    r = __builtin_fma(-k, kLn2d64Lo, r);
    double p = fma_k(r, 0.008333333333333333, 0.041666666666666664);
    p = fma_k(p, r, 0.16666666666666666);
    p = fma_k(p, r, 0.5);
    p = fma_k(p, r, 1.0);
    p = p * r;                                   // e^r - 1
    const int ki = (int)k;                       // in [-65371, 0]
    const double t = T[ki & 63];
    const double v = __builtin_fma(t, p, t);
    const double y = v * pow2i(ki >> 6);         // ki >> 6 >= -1022
    return x < -708.0 ? 0.0 : y;
}

PDDS_HD double dexpm1(double x) {
    if (__builtin_fabs(x) <= 0.5) {
        double p = 2.8114572543455206e-15;
        p = __builtin_fma(p, x, 4.779477332387385e-14);
        p = __builtin_fma(p, x, 7.647163731819816e-13);
        p = __builtin_fma(p, x, 1.1470745597729725e-11);
        p = __builtin_fma(p, x, 1.6059043836821613e-10);
        p = __builtin_fma(p, x, 2.08767569878681e-09);
        p = __builtin_fma(p, x, 2.505210838544172e-08);
        p = __builtin_fma(p, x, 2.755731922398589e-07);
        p = __builtin_fma(p, x, 2.7557319223985893e-06);
        p = __builtin_fma(p, x, 2.48015873015873e-05);
        p = __builtin_fma(p, x, 0.0001984126984126984);
        p = __builtin_fma(p, x, 0.001388888888888889);
        p = __builtin_fma(p, x, 0.008333333333333333);
        p = __builtin_fma(p, x, 0.041666666666666664);
        p = __builtin_fma(p, x, 0.16666666666666666);
        p = __builtin_fma(p, x, 0.5);
        p = __builtin_fma(p, x, 1.0);
        return p * x;
    }
    return dexp(x) - 1.0;
}

// (e^x - 1) / x, 1 at x = 0 (the local-time weight of a Robin side)
PDDS_HD double dexprel(double x) { return x == 0.0 ? 1.0 : dexpm1(x) / x; }

// dlog's steps for a positive, normal, finite x whose exponent is to be taken `sub` lower (0, or
// -54 for a subnormal scaled by 2^54).  The Box-Muller uniforms of u01_53 lie in [2^-54, 1), so
// their logs take this path directly, without dlog's special-case tests.
PDDS_HD double dlog_core(double x, int sub) {
    const uint64_t b = f64_bits(x);
    int e = (int)((b >> 52) & 0x7ff) + sub - 1023;
    double m = f64_from((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    // s^2 <= 0.0295: the atanh series terms past s^23 are below 1e-2 ulp
    // (Horner steps through fma_k: one instruction each, the coefficient from scalar registers)
    double p = 0.08695652173913043;
    p = fma_k(p, s2, 0.09523809523809523);
    p = fma_k(p, s2, 0.10526315789473684);
    p = fma_k(p, s2, 0.11764705882352941);
    p = fma_k(p, s2, 0.13333333333333333);
    p = fma_k(p, s2, 0.15384615384615385);
    p = fma_k(p, s2, 0.18181818181818182);
    p = fma_k(p, s2, 0.2222222222222222);
    p = fma_k(p, s2, 0.2857142857142857);
    p = fma_k(p, s2, 0.4);
    p = fma_k(p, s2, 0.6666666666666666);
    p = fma_k(p, s2, 2.0);
    double lm = s * p;
    double de = (double)e;
    return __builtin_fma(de, kLn2Hi, __builtin_fma(de, kLn2Lo, lm));
}

PDDS_HD double dlog(double x) {
    if (!(x > 0.0)) return (x == 0.0) ? -__builtin_huge_val() : __builtin_nan("");
    if (x == __builtin_huge_val()) return x;
    if (((f64_bits(x) >> 52) & 0x7ff) == 0) return dlog_core(x * 18014398509481984.0, -54);
    return dlog_core(x, 0);
}

PDDS_HD void sincospi_kernel(double r, double &s, double &c) {
    double y = __builtin_fma(r, kPi, r * kPiLo);
    double y2 = y * y;
    // |y| <= pi/4: the Taylor terms past y^17 (sin) and y^18 (cos) are below 1e-3 ulp
    double ps = 2.8114572543455206e-15;
    ps = fma_k(ps, y2, -7.647163731819816e-13);
    ps = fma_k(ps, y2, 1.6059043836821613e-10);
    ps = fma_k(ps, y2, -2.505210838544172e-08);
    ps = fma_k(ps, y2, 2.7557319223985893e-06);
    ps = fma_k(ps, y2, -0.0001984126984126984);
    ps = fma_k(ps, y2, 0.008333333333333333);
    ps = fma_k(ps, y2, -0.16666666666666666);
    s = __builtin_fma(ps * y2, y, y);
    double pc = -1.5619206968586225e-16;
    pc = fma_k(pc, y2, 4.779477332387385e-14);
    pc = fma_k(pc, y2, -1.1470745597729725e-11);
    pc = fma_k(pc, y2, 2.08767569878681e-09);
    pc = fma_k(pc, y2, -2.755731922398589e-07);
    pc = fma_k(pc, y2, 2.48015873015873e-05);
    pc = fma_k(pc, y2, -0.001388888888888889);
    pc = fma_k(pc, y2, 0.041666666666666664);
    pc = fma_k(pc, y2, -0.5);
    c = fma_k(pc, y2, 1.0);
}

// sin(pi x), cos(pi x); |x| < 2^52 assumed (path coordinates and uniforms)
PDDS_HD void dsincospi(double x, double &sp, double &cp) {
    double k = __builtin_floor(2.0 * x + 0.5);
    double r = x - 0.5 * k;
    double s, c;
    sincospi_kernel(r, s, c);
    int q = ((int)(int64_t)k) & 3;
    double a = (q & 1) ? c : s;
    double b = (q & 1) ? -s : c;
    sp = (q & 2) ? -a : a;
    cp = (q & 2) ? -b : b;
}
PDDS_HD double x_over_pi(double x) { return __builtin_fma(x, kInvPiHi, x * kInvPiLo); }
PDDS_HD void dsincos(double x, double &s, double &c) { dsincospi(x_over_pi(x), s, c); }
PDDS_HD double dsin(double x) { double s, c; dsincos(x, s, c); return s; }
PDDS_HD double dcos(double x) { double s, c; dsincos(x, s, c); return c; }
PDDS_HD double dtanh(double x) {
    if (x > 20.0) return 1.0;
    if (x < -20.0) return -1.0;
    double e = dexpm1(2.0 * x);
    return e / (e + 2.0);
}

// --------------------------------------------------------------- traits --
template <class R> struct RM;
template <> struct RM<double> {
    static PDDS_HD double exp(double x) { return dexp_fast(x); }
    static PDDS_HD double log(double x) { return dlog(x); }
    static PDDS_HD double sqrt(double x) { return __builtin_sqrt(x); }
    static PDDS_HD void sincospi(double x, double &s, double &c) { dsincospi(x, s, c); }
    static PDDS_HD void sincos(double x, double &s, double &c) { dsincos(x, s, c); }
    static PDDS_HD double sin(double x) { return dsin(x); }
    static PDDS_HD double cos(double x) { return dcos(x); }
    static PDDS_HD double tanh(double x) { return dtanh(x); }
    static PDDS_HD double exprel(double x) { return dexprel(x); }
    static PDDS_HD double div(double a, double b) { return a / b; }
    static PDDS_HD double floor(double x) { return __builtin_floor(x); }
    static PDDS_HD double fmin(double a, double b) { return __builtin_fmin(a, b); }
    static PDDS_HD double fmax(double a, double b) { return __builtin_fmax(a, b); }
    static PDDS_HD double fabs(double a) { return __builtin_fabs(a); }
};
template <> struct RM<float> {
    // v_exp_f32 computes 2^x, v_log_f32 log2, v_sin/v_cos_f32 take revolutions; v_sqrt_f32 is
    // the hardware square root (1 ulp, no IEEE rounding fix-up: FP32 paths are not bit-exact)
    static __device__ __forceinline__ float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
    static __device__ __forceinline__ float log(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ void sincospi(float x, float &s, float &c) {
        float t = x * 0.5f;
        t = t - __builtin_floorf(t);
        s = __builtin_amdgcn_sinf(t);
        c = __builtin_amdgcn_cosf(t);
    }
    static __device__ __forceinline__ void sincos(float x, float &s, float &c) {
        float t = x * 0.15915494309189535f;
        t = t - __builtin_floorf(t);
        s = __builtin_amdgcn_sinf(t);
        c = __builtin_amdgcn_cosf(t);
    }
    static __device__ __forceinline__ float sin(float x) { float s, c; sincos(x, s, c); return s; }
    static __device__ __forceinline__ float cos(float x) { float s, c; sincos(x, s, c); return c; }
    // (e^x - 1) / x: Taylor terms to x^4 for |x| < 1/16 (truncation below 2^-24), else from exp
    static __device__ __forceinline__ float exprel(float x) {
        if (__builtin_fabsf(x) < 0.0625f)
            return __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(__builtin_fmaf(x, 1.0f / 120.0f, 1.0f / 24.0f), x,
                                                                1.0f / 6.0f), x, 0.5f), x, 1.0f);
        return (exp(x) - 1.0f) * __builtin_amdgcn_rcpf(x);
    }
    static __device__ __forceinline__ float tanh(float x) {
        float ax = __builtin_fabsf(x);
        float e = exp(-2.0f * ax);
        float t = (1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e);
        return x < 0.0f ? -t : t;
    }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float floor(float x) { return __builtin_floorf(x); }
    static __device__ __forceinline__ float fmin(float a, float b) { return __builtin_fminf(a, b); }
    static __device__ __forceinline__ float fmax(float a, float b) { return __builtin_fmaxf(a, b); }
    static __device__ __forceinline__ float fabs(float a) { return __builtin_fabsf(a); }
};

// ---------------------------------------------------------------- Philox --
struct Philox4 { uint32_t v[4]; };
PDDS_HD Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    Philox4 o;
    o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

PDDS_HD double u01_53(uint32_t a, uint32_t b) {
    uint64_t m = ((uint64_t)a << 21) | (uint64_t)(b >> 11);
    return ((double)m + 0.5) * 1.1102230246251565e-16;
}

// FP64 normal pair for one step: one Philox draw, Box-Muller on 53-bit uniforms.
PDDS_HD void normals_f64(uint32_t seed, uint32_t stream, uint32_t knot, uint32_t path, int64_t step,
                         double &n1, double &n2) {
    Philox4 w = philox4x32_10((uint32_t)step, path, knot, (uint32_t)(step >> 32), seed, stream);
    double u1 = u01_53(w.v[0], w.v[1]);
    double u2 = u01_53(w.v[2], w.v[3]);
    double r = __builtin_sqrt(-2.0 * dlog_core(u1, 0));
    double s, c;
    dsincospi(2.0 * u2, s, c);
    n1 = r * c;
    n2 = r * s;
}

// ------------------------------------------------------------ Rectangle --
// Signed distance to p = [x0,y0,x1,y1] with nearest-side projection and outward
// normal; centre/half-side form with the quadrant split and tie rule of the
// reference (BVPs/Domains/rectangle.hpp:7-229).  `kind` reports the projection
// as an integer: 0 N, 1 S, 2 E, 3 W side; 4+ corner (4 NE, 5 NW, 6 SE, 7 SW).
template <class R>
PDDS_HD R rect_dist(const R p[4], R x, R y, R &ex, R &ey, R &nx, R &ny, int &kind) {
    const R cx = (R)0.5 * (p[0] + p[2]), cy = (R)0.5 * (p[1] + p[3]);
    const R hx = p[2] - cx, hy = p[3] - cy;
    const R dx = RM<R>::fabs(x - cx), dy = RM<R>::fabs(y - cy);
    const R diag = (R)0.70710678118f;
    const bool right = x > cx, top = y > cy;
    const R sxs = right ? (R)1 : (R)-1;
    const R sys = top ? (R)1 : (R)-1;
    ex = x; ey = y;
    if (dx <= hx && dy <= hy) {
        if (dx - hx > dy - hy) {
            ex = sxs * hx + cx; nx = sxs; ny = 0;
            kind = right ? 2 : 3;
            return dx - hx;
        }
        ey = sys * hy + cy; nx = 0; ny = sys;
        kind = top ? 0 : 1;
        return dy - hy;
    }
    if (dx > hx) {
        if (dy > hy) {
            R r = RM<R>::sqrt((dx - hx) * (dx - hx) + (dy - hy) * (dy - hy));
            ex = sxs * hx + cx; ey = sys * hy + cy;
            nx = sxs * diag; ny = sys * diag;
            kind = top ? (right ? 4 : 5) : (right ? 6 : 7);
            return r;
        }
        ex = sxs * hx + cx; nx = sxs; ny = 0;
        kind = right ? 2 : 3;
        return dx - hx;
    }
    ey = sys * hy + cy; nx = 0; ny = sys;
    kind = top ? 0 : 1;
    return dy - hy;
}

// What the stopping test needs of rect_dist, without branches: inside the rectangle its signed
// distance and outward normal, bit for bit (same operations, same tie rule); outside a positive
// distance (the largest side excess, which is not rect_dist's corner distance) -- a point there
// stops whatever its value, and the exit projection comes from rect_dist itself.
template <class R>
PDDS_HD R rect_dist_inside(const R p[4], R x, R y, R &nx, R &ny) {
    const R cx = (R)0.5 * (p[0] + p[2]), cy = (R)0.5 * (p[1] + p[3]);
    const R hx = p[2] - cx, hy = p[3] - cy;
    const R ax = RM<R>::fabs(x - cx) - hx, ay = RM<R>::fabs(y - cy) - hy;
    const bool xside = ax > ay;
    const R sxs = x > cx ? (R)1 : (R)-1;
    const R sys = y > cy ? (R)1 : (R)-1;
    nx = xside ? sxs : (R)0;
    ny = xside ? (R)0 : sys;
    return xside ? ax : ay;
}

// Signed distance to the annular sector g = {r0, r1, cos a, sin a, cos b, sin b}: radii
// r0 <= r <= r1 between the rays a and b (b counterclockwise from a, less than pi apart;
// r0 = 0 for a sector at the centre, which then has no inner side; r0 < 0 for the whole disc
// r <= r1, whose only side is its circle).  Inside, the distance to
// the nearest side is the largest of the four half-plane / annulus distances (exact for such
// a sector); outside, that largest violation stands for the distance (positive).  The side it
// names gives the projection (clamped to the side) and the outward normal; `kind` reports it
// as 0 outer arc, 1 inner arc, 2 ray b, 3 ray a.
template <class R>
PDDS_HD R sector_dist(const R g[6], R x, R y, R &ex, R &ey, R &nx, R &ny, int &kind) {
    const R r0 = g[0], r1 = g[1], ca = g[2], sa = g[3], cb = g[4], sb = g[5];
    const R r = RM<R>::sqrt(x * x + y * y);
    const R sA = ca * y - sa * x;          // > 0: counterclockwise of ray a (inside)
    const R sB = x * sb - y * cb;          // > 0: clockwise of ray b (inside)
    const bool whole = r0 < (R)0;
    R d = r - r1;
    int k = 0;
    const R din = r0 > (R)0 ? r0 - r : -(r + r1);
    if (din > d) { d = din; k = 1; }
    if (!whole && -sB > d) { d = -sB; k = 2; }
    if (!whole && -sA > d) { d = -sA; k = 3; }
    if (k <= 1) {
        // radial direction of X, or the nearer ray's when X lies outside the angle
        R ux = x / r, uy = y / r;
        if (!whole && sA < (R)0) { ux = ca; uy = sa; }
        else if (!whole && sB < (R)0) { ux = cb; uy = sb; }
        const R rr = k == 0 ? r1 : r0;
        ex = rr * ux; ey = rr * uy;
        nx = k == 0 ? ux : -ux; ny = k == 0 ? uy : -uy;
    } else {
        const R dx = k == 3 ? ca : cb, dy = k == 3 ? sa : sb;
        R t = x * dx + y * dy;
        t = t < r0 ? r0 : (t > r1 ? r1 : t);
        ex = t * dx; ey = t * dy;
        nx = k == 3 ? sa : -sb;
        ny = k == 3 ? -ca : cb;
    }
    kind = k;
    return d;
}

// --------------------------------------------------------------- LUT -----
// Cubic B-spline coefficient table of nx x ny texels, t = (x - ox) * sx.  Like the
// reference's FP32 texture lookups (Integrators/cubicTex2D.cu:64-87), the 16-tap weighted
// sum runs in FP32 -- here with explicit fma -- whatever the path precision; only the
// coordinate is formed in R.
//
// Device layout: the table is stored with kLutPad texels of replicated border on the
// low side and kLutPad + 1 on the high side (padded row of px = nx + 2*kLutPad + 1
// texels), so every tap of a clamped coordinate t in [-2, nx + 1] lies inside the
// padded table and no per-tap clamp is needed; the replicated border gives exactly the
// clamp-to-edge values.  Fields 0..3 live in one float4 plane (a).  For nf == 5, plane b
// holds field 4 alone, one float per texel, and the four taps of a row are one 16-byte load
// at a 4-byte aligned offset (a row's taps never run past the padded row: the last tap
// column is at most nx + 6 = px - 1); a patch's field-4 texels then take a quarter of the
// L2 that 16-byte x-quads would.  For nf > 5 plane b holds fields 4..7 per texel.  Every
// load is at a 32-bit byte offset from a uniform base.  A gather of scattered lanes costs
// one cache line per lane and instruction, so a step costs 16 + 4 line accesses per lane
// for nf == 5 (16 + 16 for nf > 5).
constexpr int kLutPad = 3;

// four consecutive floats of a 4-byte aligned address
typedef float float4_a4 __attribute__((ext_vector_type(4), aligned(4)));

struct LutView {
    const float4 *a;       // fields 0..3
    const void *b;         // nf == 5: field 4 (float per texel); nf > 5: fields 4..7 (float4); else null
    int nx, ny, px, nf;    // logical texels per row / column, padded row length, fields
    double ox, oy, sx, sy;
};

// Cubic B-spline weights at fraction f in FP32 (explicit fma)
PDDS_HD void bspline_wf(float f, float w[4]) {
    const float s6 = 1.0f / 6.0f;
    const float f2 = f * f, om = 1.0f - f;
    w[0] = (om * om * om) * s6;
    w[1] = __builtin_fmaf(__builtin_fmaf(3.0f, f, -6.0f), f2, 4.0f) * s6;
    w[2] = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(-3.0f, f, 3.0f), f, 3.0f), f, 1.0f) * s6;
    w[3] = (f2 * f) * s6;
}

template <class T>
PDDS_HD const T &lut_at(const void *base, uint32_t byte_off) {
    return *(const T *)((const char *)base + byte_off);
}

// A rectangle of plane a held in LDS: padded columns [c0, c0 + tw), rows [r0, r0 + th),
// row-major with `pitch` texels per row (pitch >= tw; the spare texels change which LDS banks
// the rows start on).  th == 0 means no tile.  The pointer is typed in the LDS address space,
// so tile reads are ds_read_b128 and never merge with the global-memory branch into generic
// (flat) loads.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) const float4 lds_float4;
#else
typedef const float4 lds_float4;   // the host pass never dereferences a tile
#endif
struct LutTile {
    lds_float4 *lds;
    int c0, r0, tw, th;
    int pitch;
};

// The 16-tap accumulation of lut_eval_t.  SRC 0: plane a from global memory; 1: from the LDS
// tile; 2: per lane, the tile when `in`, else global memory.  The values and the order of the
// fused multiply-adds are the same for every source.
template <int NF, int SRC>
PDDS_HD void lut_accum(const LutView &L, const LutTile &T, lds_float4 *tp, bool in, uint32_t t0, const float wx[4],
                       const float wy[4], float acc[NF]) {
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = 0.0f;
    // the fifth field's four taps of each row first, so that their global loads are all in
    // flight while the rows' plane-a taps are read
    float4_a4 Q[4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
        Q[b] = NF == 5 ? lut_at<float4_a4>(L.b, (t0 + (uint32_t)b * (uint32_t)L.px) * 4u) : float4_a4{0.f, 0.f, 0.f, 0.f};
    // fields 0..3 (and 4..7 for eight fields) row by row; the fifth field's quads last, so that
    // their global loads have the whole plane-a accumulation to arrive (each field's fused
    // multiply-adds keep their order, so the values do not depend on this)
    constexpr int NA = NF == 5 ? 4 : NF;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint32_t r0 = t0 + (uint32_t)b * (uint32_t)L.px;
        float4 A[4];
        if (SRC == 1 || (SRC == 2 && in)) {
#pragma unroll
            for (int a = 0; a < 4; ++a) A[a] = tp[b * T.pitch + a];
        } else {
#pragma unroll
            for (int a = 0; a < 4; ++a) A[a] = lut_at<float4>(L.a, (r0 + (uint32_t)a) * 16u);
        }
        float row[NA];
#pragma unroll
        for (int f = 0; f < NA; ++f) row[f] = 0.0f;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            float c[8];
            const float4 q0 = A[a];
            c[0] = q0.x; c[1] = q0.y; c[2] = q0.z; c[3] = q0.w;
            if constexpr (NF > 5) {
                const float4 q1 = lut_at<float4>(L.b, (r0 + (uint32_t)a) * 16u);
                c[4] = q1.x; c[5] = q1.y; c[6] = q1.z; c[7] = q1.w;
            }
#pragma unroll
            for (int f = 0; f < NA; ++f) row[f] = __builtin_fmaf(wx[a], c[f], row[f]);
        }
#pragma unroll
        for (int f = 0; f < NA; ++f) acc[f] = __builtin_fmaf(wy[b], row[f], acc[f]);
    }
    if constexpr (NF == 5) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const float4_a4 quad = Q[b];
            float row = __builtin_fmaf(wx[0], quad.x, 0.0f);
            row = __builtin_fmaf(wx[1], quad.y, row);
            row = __builtin_fmaf(wx[2], quad.z, row);
            row = __builtin_fmaf(wx[3], quad.w, row);
            acc[4] = __builtin_fmaf(wy[b], row, acc[4]);
        }
    }
}

// Evaluate fields [0, NF) (NF <= 8) at (x,y):
//   sum_b wy_b * sum_a wx_a * C[clamp(j-1+b)][clamp(i-1+a)][f]
// With TILED, a lane whose 4x4 taps lie inside the tile reads fields 0..3 from LDS (one
// 16-byte LDS read per tap) and the others read them from plane a; the values, and the
// order of the fused multiply-adds, are the same either way.  A wave whose lanes are all
// inside the tile takes a path with LDS reads only, so no LDS read waits for the wave's
// global loads.
template <class R, int NF, bool TILED>
PDDS_HD void lut_eval_t(const LutView &L, const LutTile &T, R x, R y, R out[NF]) {
    static_assert(NF <= 8, "at most 8 fields per texel");
    R t = (x - (R)L.ox) * (R)L.sx;
    R v = (y - (R)L.oy) * (R)L.sy;
    t = RM<R>::fmin(RM<R>::fmax(t, (R)-2), (R)L.nx + (R)1);
    v = RM<R>::fmin(RM<R>::fmax(v, (R)-2), (R)L.ny + (R)1);
    const R fi = RM<R>::floor(t), fj = RM<R>::floor(v);
    float wx[4], wy[4];
    bspline_wf((float)(t - fi), wx);
    bspline_wf((float)(v - fj), wy);
    // padded column / row of tap (a=0, b=0): i - 1 + kLutPad >= 0, j - 1 + kLutPad >= 0
    const int pc = (int)fi + (kLutPad - 1), pr = (int)fj + (kLutPad - 1);
    const uint32_t t0 = (uint32_t)pr * (uint32_t)L.px + (uint32_t)pc;
    float acc[NF];
    if constexpr (TILED) {
        const int lc = pc - T.c0, lr = pr - T.r0;
        const bool in = lc >= 0 && lr >= 0 && lc + 3 < T.tw && lr + 3 < T.th;
        lds_float4 *tp = T.lds + (in ? lr * T.pitch + lc : 0);
#if defined(__HIP_DEVICE_COMPILE__)
        if (__all(in)) lut_accum<NF, 1>(L, T, tp, in, t0, wx, wy, acc);
        else lut_accum<NF, 2>(L, T, tp, in, t0, wx, wy, acc);
#else
        lut_accum<NF, 2>(L, T, tp, in, t0, wx, wy, acc);
#endif
    } else {
        lut_accum<NF, 0>(L, T, T.lds, false, t0, wx, wy, acc);
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) out[f] = (R)acc[f];
}

template <class R, int NF>
PDDS_HD void lut_eval(const LutView &L, R x, R y, R out[NF]) {
    const LutTile none{nullptr, 0, 0, 0, 0};
    lut_eval_t<R, NF, false>(L, none, x, y, out);
}

}  // namespace pdds
