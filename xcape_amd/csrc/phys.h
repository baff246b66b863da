// Physical constants and scalar thermodynamics shared by the gfx950 kernels.
//
// Constants follow getcape (reference src/xcape/fortran/CAPE_CODE_model_lev.f90
// :188-211).  For T = float every derived constant is folded in single
// precision, as the reference's REAL parameters are; T = double uses the same
// decimal values in double precision (the FP64 yardstick run).
//
// Math modes:
//   MATH_ACCURATE - IEEE-rounded + - * / and exp, log, pow within about an
//                   ulp (the device library's expf and logf, the pow kernel
//                   below); translation units using it are compiled with
//                   -ffp-contract=off so every + - * / rounds exactly as the
//                   reference's does.
//   MATH_FAST     - hardware exp2/log2 and reciprocal-based division (a few
//                   ulp); FMA contraction allowed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Functions usable from host code too (the accurate path only needs the C
// math library), so the column routine can be exercised on the CPU.
#define XC_HD __host__ __device__

namespace xcape {

// MATH_LIB: the accurate mode's arithmetic done by the device library
// (IEEE division, expf/logf/powf); the fallback of its faster kernels below.
enum { MATH_ACCURATE = 0, MATH_FAST = 1, MATH_LIB = 3 };

template <typename T> struct Phys;

template <> struct Phys<float> {
    static constexpr float g = 9.81f, p00 = 100000.0f, cp = 1005.7f, rd = 287.04f,
                           rv = 461.5f, xlv = 2501000.0f, xls = 2836017.0f, t0 = 273.15f,
                           cpv = 1875.0f, cpl = 4190.0f, cpi = 2118.636f;
    static constexpr float lv1 = xlv + (cpl - cpv) * t0;
    static constexpr float lv2 = cpl - cpv;
    static constexpr float ls1 = xls + (cpi - cpv) * t0;
    static constexpr float ls2 = cpi - cpv;
    static constexpr float rp00 = 1.0f / p00;
    static constexpr float eps = rd / rv;
    static constexpr float reps = rv / rd;
    static constexpr float rddcp = rd / cp;
    static constexpr float cpdg = cp / g;
    static constexpr float converge = 0.0002f;
};

template <> struct Phys<double> {
    static constexpr double g = 9.81, p00 = 100000.0, cp = 1005.7, rd = 287.04, rv = 461.5,
                            xlv = 2501000.0, xls = 2836017.0, t0 = 273.15, cpv = 1875.0,
                            cpl = 4190.0, cpi = 2118.636;
    static constexpr double lv1 = xlv + (cpl - cpv) * t0;
    static constexpr double lv2 = cpl - cpv;
    static constexpr double ls1 = xls + (cpi - cpv) * t0;
    static constexpr double ls2 = cpi - cpv;
    static constexpr double rp00 = 1.0 / p00;
    static constexpr double eps = rd / rv;
    static constexpr double reps = rv / rd;
    static constexpr double rddcp = rd / cp;
    static constexpr double cpdg = cp / g;
    static constexpr double converge = 0.0002;
};

// ---- device kernels of the accurate mode -----------------------------------
//
// Division: the IEEE quotient from the hardware reciprocal and Markstein's
// corrections -- y = 1/b refined by one Newton step, q = a y corrected twice
// by the exact residual a - b q (fma).  These are the steps of the compiler's
// own IEEE division without its operand scaling and special-value fix-up, so
// where that scaling is the identity -- a, b and a/b normal, |b| < 2^126 --
// the result is the same correctly rounded quotient, bit for bit.  A zero,
// infinite or NaN operand gives NaN; subnormal or huge operands give a wrong
// value.  Callers make sure of the range (the fixed-point evaluation below
// re-evaluates with IEEE division whenever its result is not finite).
__device__ __forceinline__ float div_raw(float a, float b)
{
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    const float y = __builtin_fmaf(e, y0, y0);
    const float q0 = a * y;
    const float r0 = __builtin_fmaf(-b, q0, a);
    const float q1 = __builtin_fmaf(r0, y, q0);
    const float r1 = __builtin_fmaf(-b, q1, a);
    return __builtin_fmaf(r1, y, q1);
}

// Two quotients over one divisor, sharing the refined reciprocal (each
// quotient's steps are div_raw's).
__device__ __forceinline__ void div2_raw(float a1, float a2, float b, float &q1o, float &q2o)
{
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    const float y = __builtin_fmaf(e, y0, y0);
    float q = a1 * y;
    float r = __builtin_fmaf(-b, q, a1);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-b, q, a1);
    q1o = __builtin_fmaf(r, y, q);
    q = a2 * y;
    r = __builtin_fmaf(-b, q, a2);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-b, q, a2);
    q2o = __builtin_fmaf(r, y, q);
}

// a / c for a divisor fixed per call, with rc = RN(1/c) computed once:
// q0 = RN(a rc) is within an ulp of a/c, and one correction with the exact
// residual gives RN(a/c) (Markstein's theorem for a correctly rounded
// reciprocal), for finite a with a normal quotient -- checked against IEEE
// division over every float of magnitude 1e-30 to 1e30 for the divisors 40,
// 100, 500, 1000, 2500 and others: all agree but -0, which gives +0.
__device__ __forceinline__ float div_const(float a, float c, float rc)
{
    const float q0 = a * rc;
    const float r = __builtin_fmaf(-c, q0, a);
    return __builtin_fmaf(r, rc, q0);
}

// exp(x): the device library's reduction -- x log2(e) as a double-float,
// its nearest integer n split off exactly, 2^f by the hardware exp2 on
// |f| <= 1/2, scaled by 2^n -- without the library's overflow and underflow
// selects: the library's value for -87 < x < 88 (a NaN stays NaN).
__device__ __forceinline__ float exp_core(float x)
{
    // log2(e) as the double-float the library uses (its bit patterns
    // 0x3fb8aa3b, 0x32a5705f)
    constexpr float L2E = 0x1.715476p+0f;
    constexpr float L2E_LO = 0x1.4ae0bep-26f;
    const float ph = x * L2E;
    float pl = __builtin_fmaf(x, L2E, -ph);
    pl = __builtin_fmaf(x, L2E_LO, pl);
    const float n = __builtin_rintf(ph);
    const float f = (ph - n) + pl;
    return __builtin_ldexpf(__builtin_amdgcn_exp2f(f), (int)n);
}

// x^y for 0 < x < 2^126, |y| <= 1 (the Exner function and theta-e's power):
// x = m 2^k with m in [sqrt(1/2), sqrt(2)); log2(m) from the hardware log2
// (|log2 m| <= 1/2); y (k + log2 m) as a double-float (exact products by
// fma, a Fast2Sum: |y k| >= |y log2 m| whenever k != 0); 2^t as in exp_core.
// Its error is the hardware exp2's plus about 0.15 ulp; other operands take
// the device library's powf.
__device__ __forceinline__ float pow_core(float x, float y)
{
    if (__builtin_expect(!(x > 0.0f && x < 8.5e37f), 0)) return powf(x, y);
    float m = __builtin_amdgcn_frexp_mantf(x);          // [1/2, 1)
    int k = __builtin_amdgcn_frexp_expf(x);
    if (m < 0.70710678118654752f) {
        m = m + m;
        k = k - 1;
    }
    const float l = __builtin_amdgcn_logf(m);           // log2(m)
    const float kf = (float)k;
    const float ah = y * kf;
    const float al = __builtin_fmaf(y, kf, -ah);
    const float bh = y * l;
    const float bl = __builtin_fmaf(y, l, -bh);
    const float s = ah + bh;
    const float se = bh - (s - ah);
    const float lo = se + (al + bl);
    const float n = __builtin_rintf(s);
    const float f = (s - n) + lo;
    return __builtin_ldexpf(__builtin_amdgcn_exp2f(f), (int)n);
}

// FP64 division by the steps of the compiler's IEEE sequence (the hardware
// reciprocal, two Newton steps, one correction by the exact residual) without
// its operand scaling and special-value fix-up: the same quotient bit for bit
// wherever that scaling is the identity (a, b and a/b normal, |b| far from
// overflow); a zero, infinite or NaN operand gives NaN.  Callers re-evaluate
// with IEEE division when their result is not finite.
__device__ __forceinline__ double ddiv_raw(double a, double b)
{
    double y = __builtin_amdgcn_rcp(b);
    y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
    y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
    const double q = a * y;
    return __builtin_fma(__builtin_fma(-b, q, a), y, q);
}

// A constant the compiler materialises where it is used (in scalar
// registers) instead of holding it across a loop.
__device__ __forceinline__ double xc_sconst(double c)
{
    __asm__ volatile("" : "+s"(c));
    return c;
}

// ln x in FP64 for the helicity chain's heights, by the reduction of Sun's
// fdlibm log (1993): x = 2^k m with m in [sqrt(1/2), sqrt(2)), f = m - 1
// (exact), s = f / (2 + f), ln m = 2 atanh s = f - (f^2/2 - s (f^2/2 + R(s^2)))
// with R(z) = sum 2 z^j / (2j + 1) (j >= 1) as z P(z), P the degree-6
// Chebyshev fit of R(z)/z on [0, ((sqrt 2 - 1)/(sqrt 2 + 1))^2] (relative
// error of the fit in ln m below 5e-18), and k ln 2 as k (LN2_HI + LN2_LO),
// LN2_HI with 29 significant bits so that k LN2_HI is exact.  s by the
// hardware reciprocal refined by two Newton steps (s enters only the
// correction term).  Within an ulp of ln x for positive finite x (the
// hardware's frexp normalises subnormals); ln 0 = -inf, ln inf = inf, and
// NaN for negative x or NaN, as the C library's.
__device__ __forceinline__ double log_core_f64(double x)
{
    if (__builtin_expect(!(x > 0.0 && x <= 0x1.fffffffffffffp+1023), 0))
        return x == 0.0 ? -__builtin_inf() : (x > 0.0 ? x : __builtin_nan(""));
    constexpr double LN2_HI = 0x1.62e42fe000000p-1;
    constexpr double LN2_LO = 0x1.f473de6af278fp-30;
    constexpr double C1 = 0x1.5555555555558p-1, C2 = 0x1.99999999952e2p-2,
                     C3 = 0x1.2492492df148dp-2, C4 = 0x1.c71c62e5800a1p-3,
                     C5 = 0x1.7462b4ab2ef6bp-3, C6 = 0x1.39fe606542ddep-3,
                     C7 = 0x1.2b584aae78a57p-3;
    double m = __builtin_amdgcn_frexp_mant(x);            // [1/2, 1)
    int k = __builtin_amdgcn_frexp_exp(x);
    if (m < 0x1.6a09e667f3bcdp-1) {
        m = m + m;
        k = k - 1;
    }
    const double f = m - 1.0;
    const double b = 2.0 + f;
    double y = __builtin_amdgcn_rcp(b);
    y = __builtin_fma(__builtin_fma(-b, y, 1.0), y, y);
    y = __builtin_fma(__builtin_fma(-b, y, 1.0), y, y);
    const double s = f * y;
    const double z = s * s;
    double p = __builtin_fma(z, xc_sconst(C7), xc_sconst(C6));
    p = __builtin_fma(z, p, xc_sconst(C5));
    p = __builtin_fma(z, p, xc_sconst(C4));
    p = __builtin_fma(z, p, xc_sconst(C3));
    p = __builtin_fma(z, p, xc_sconst(C2));
    p = __builtin_fma(z, p, xc_sconst(C1));
    const double r = z * p;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)k;
    return dk * LN2_HI - ((hfsq - (s * (hfsq + r) + dk * LN2_LO)) - f);
}

// e^x in FP64 for the helicity chain's heights: x = k ln 2 + r with k the
// nearest integer to x / ln 2 and r = x - k (LN2_HI + LN2_LO) (|r| <= ln 2 /
// 2; x - k LN2_HI is exact, and c carries the rounding error of r),
// e^r = 1 + (r + (r^2 Q(r) + c)) with Q the degree-10 Chebyshev fit of
// (e^r - 1 - r) / r^2 on that interval (relative error of the fit in e^r
// below 4e-19), scaled by 2^k.  Within an ulp of e^x wherever the result
// is normal (-708.3 < x < 709.7); subnormal results, overflow to inf,
// underflow to 0, NaN as the C library's (up to an ulp of the subnormal).
__device__ __forceinline__ double exp_core_f64(double x)
{
    if (__builtin_expect(!(x > -745.2 && x < 709.8), 0))
        return x != x ? x : (x > 0.0 ? __builtin_inf() : 0.0);
    constexpr double INV_LN2 = 0x1.71547652b82fep+0;
    constexpr double LN2_HI = 0x1.62e42fe000000p-1;
    constexpr double LN2_LO = 0x1.f473de6af278fp-30;
    constexpr double E0 = 0x1.0000000000000p-1, E1 = 0x1.5555555555557p-3,
                     E2 = 0x1.5555555555556p-5, E3 = 0x1.11111111100dfp-7,
                     E4 = 0x1.6c16c16c162d6p-10, E5 = 0x1.a01a01abe63b3p-13,
                     E6 = 0x1.a01a01a6d7883p-16, E7 = 0x1.71de02372bdd3p-19,
                     E8 = 0x1.27e4db679bec4p-22, E9 = 0x1.af4ddddcf06d6p-26,
                     E10 = 0x1.1f72fca58cc92p-29;
    const double k = __builtin_rint(x * xc_sconst(INV_LN2));
    const double hi = __builtin_fma(-k, xc_sconst(LN2_HI), x);    // exact
    const double lo = k * xc_sconst(LN2_LO);
    const double r = hi - lo;
    const double c = (hi - r) - lo;             // r's rounding error
    double q = __builtin_fma(r, xc_sconst(E10), xc_sconst(E9));
    q = __builtin_fma(r, q, xc_sconst(E8));
    q = __builtin_fma(r, q, xc_sconst(E7));
    q = __builtin_fma(r, q, xc_sconst(E6));
    q = __builtin_fma(r, q, xc_sconst(E5));
    q = __builtin_fma(r, q, xc_sconst(E4));
    q = __builtin_fma(r, q, xc_sconst(E3));
    q = __builtin_fma(r, q, xc_sconst(E2));
    q = __builtin_fma(r, q, xc_sconst(E1));
    q = __builtin_fma(r, q, E0);
    const double p = 1.0 + (r + __builtin_fma(r * r, q, c));
    return __builtin_ldexp(p, (int)k);
}

// ---- elementary functions per precision and mode ---------------------------
template <typename T, int MATH> struct Fn;

// Accurate mode: IEEE-rounded + - * / (the translation units compile with
// -ffp-contract=off), so those round as the reference's do; exp and log the
// device library's, pow the kernel above (within an ulp; the host build: the
// C library's).  The fixed-point evaluation has a faster path of its own
// (fp_eval in cape_column.h).
template <> struct Fn<float, MATH_ACCURATE> {
    static XC_HD __forceinline__ float exp(float x) { return expf(x); }
    static XC_HD __forceinline__ float log(float x) { return logf(x); }
    static XC_HD __forceinline__ float pow(float x, float y)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return pow_core(x, y);
#else
        return powf(x, y);
#endif
    }
    static XC_HD __forceinline__ float div(float a, float b) { return a / b; }
    // exp(c * a / b) with c*a formed first, as the reference does
    static XC_HD __forceinline__ float exp_ratio(float c, float a, float b)
    {
        return expf(c * a / b);
    }
};

template <> struct Fn<float, MATH_LIB> {
    static XC_HD __forceinline__ float exp(float x) { return expf(x); }
    static XC_HD __forceinline__ float log(float x) { return logf(x); }
    static XC_HD __forceinline__ float pow(float x, float y) { return powf(x, y); }
    static XC_HD __forceinline__ float div(float a, float b) { return a / b; }
    static XC_HD __forceinline__ float exp_ratio(float c, float a, float b)
    {
        return expf(c * a / b);
    }
};

template <> struct Fn<float, MATH_FAST> {
    static constexpr float LOG2E = 1.4426950408889634f;
    static constexpr float LN2 = 0.6931471805599453f;
    static __device__ __forceinline__ float rcp(float b) { return __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float exp(float x)
    {
        return __builtin_amdgcn_exp2f(x * LOG2E);
    }
    static __device__ __forceinline__ float log(float x)
    {
        return __builtin_amdgcn_logf(x) * LN2;
    }
    static __device__ __forceinline__ float pow(float x, float y)
    {
        return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
    }
    // one Newton step on the hardware reciprocal, then a residual correction:
    // error below 1 ulp away from overflow/underflow, no special-case handling
    static __device__ __forceinline__ float div(float a, float b)
    {
        float r = rcp(b);
        float q = a * r;
        float e = __builtin_fmaf(-b, q, a);
        return __builtin_fmaf(e, r, q);
    }
    static __device__ __forceinline__ float exp_ratio(float c, float a, float b)
    {
        return __builtin_amdgcn_exp2f((c * LOG2E) * a * rcp(b));
    }
};

// Division for arithmetic that never reaches a result, only a starting
// guess: the hardware reciprocal plus one correction on the device.
template <typename T>
XC_HD __forceinline__ T pred_div(T a, T b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 4) return Fn<float, MATH_FAST>::div(a, b);
#endif
    return a / b;
}

// exp for the same kind of arithmetic: the hardware exp2 on the device.
template <typename T>
XC_HD __forceinline__ T pred_exp(T x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 4) return Fn<float, MATH_FAST>::exp(x);
#endif
    return ::exp(x);
}

template <int MATH> struct Fn<double, MATH> {
    static XC_HD __forceinline__ double exp(double x) { return ::exp(x); }
    static XC_HD __forceinline__ double log(double x) { return ::log(x); }
    static XC_HD __forceinline__ double pow(double x, double y) { return ::pow(x, y); }
    static XC_HD __forceinline__ double div(double a, double b) { return a / b; }
    static XC_HD __forceinline__ double exp_ratio(double c, double a, double b)
    {
        return ::exp(c * a / b);
    }
};

// Fortran MAX/MIN as the reference evaluates them: the second operand wins
// when the comparison is false (NaN included).
template <typename T> XC_HD __forceinline__ T fmx(T a, T b) { return a > b ? a : b; }
template <typename T> XC_HD __forceinline__ T fmn(T a, T b) { return a < b ? a : b; }

// int(x) for the sub-step count: NaN/out-of-range map to INT32_MIN as the
// reference build's truncating conversion does (the sub-step loop then runs
// zero times).
template <typename T> XC_HD __forceinline__ int ftrunc(T x)
{
    if (!(x > T(-2147483648.0) && x < T(2147483648.0))) return INT32_MIN;
    return (int)x;
}

// getqvs: Bolton saturation mixing ratio over liquid (CAPE_CODE_model_lev.f90
// :570-581); es = 611.2 exp(17.67 (t-273.15)/(t-29.65)), r = eps es/(p-es)
template <typename T, int MATH>
XC_HD __forceinline__ T qvs(T p, T t)
{
    using F = Fn<T, MATH>;
    T es = T(611.2) * F::exp_ratio(T(17.67), t - T(273.15), t - T(29.65));
    return F::div(Phys<T>::eps * es, p - es);
}

// getqvi: over ice (:587-598)
template <typename T, int MATH>
XC_HD __forceinline__ T qvi(T p, T t)
{
    using F = Fn<T, MATH>;
    T es = T(611.2) * F::exp_ratio(T(21.8745584), t - T(273.15), t - T(7.66));
    return F::div(Phys<T>::eps * es, p - es);
}

// getthe: Bolton (1980) theta-e with the LCL temperature (:604-620)
template <typename T, int MATH>
XC_HD __forceinline__ T thetae(T p, T t, T td, T q)
{
    using F = Fn<T, MATH>;
    T tlcl;
    if ((td - t) >= T(-0.1)) {
        tlcl = t;
    } else {
        T a = F::div(T(1.0), td - T(56.0));
        tlcl = T(56.0) + F::div(T(1.0), a + T(0.00125) * F::log(F::div(t, td)));
    }
    return t * F::pow(F::div(T(100000.0), p), T(0.2854) * (T(1.0) - T(0.28) * q)) *
           F::exp((F::div(T(3376.0), tlcl) - T(2.54)) * q * (T(1.0) + T(0.81) * q));
}

}  // namespace xcape
