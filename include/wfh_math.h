/*
 * wfh_math.h - elementary functions and Julia-semantics helpers shared by the HIP kernels
 * and by host code (gcc / clang).
 *
 * Why a project-written exp/log: the kinematic-wave Newton solves of the reference stop on
 * an ABSOLUTE residual of 1e-12 (Wflow/src/routing/surface/surface_process.jl:47-58), and the
 * Brooks-Corey drainage loop takes Int(cld(remainder, 2e-4)) sub-iterations
 * (Wflow/src/soil/soil_process.jl:74). Both turn a one-ulp difference in exp/log into a
 * different iteration count. Compiling the same straight-line IEEE-754 code for the host and
 * for gfx950 (no FMA contraction: -ffp-contract=off; fma only where written explicitly, which
 * is correctly rounded on both) gives bit-identical results on both sides.
 *
 * Methods (own implementation, from the mathematics):
 *   exp: x = k ln2 + r, |r| <= ln2/2 (two-part ln2), e^r by its Taylor series to r^13
 *        (truncation < 4e-18), Horner with fma, then scaling by 2^k in two steps so that
 *        subnormal results are produced by one final multiply.
 *   log: x = 2^k m, m in [sqrt(2)/2, sqrt(2)), f = m - 1, s = f / (2 + f);
 *        log(1+f) = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R(s^2))), R(z) = sum_j 2 z^j / (2j+1)
 *        (the identity 2s = f - s f gives the rearrangement); the series is cut at j = 12
 *        (|z| <= 0.0295, truncation < 1e-20 relative).
 *   cbrt: sign(x) exp(log|x| / 3) refined by one Newton step.
 *   pow:  exp(y log x), exactly the composition the reference uses (Wflow/src/utils.jl:470).
 * Accuracy (tests/test_math.py): exp, log and cbrt within 1 ulp of libm over their whole range,
 * pow within 2 + 2|y log x| ulp (the error of the exp(y log x) composition itself), cld exact
 * against rational arithmetic.
 */
#ifndef WFH_MATH_H
#define WFH_MATH_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define WFH_FN __host__ __device__ static inline
#else
#define WFH_FN static inline
#endif

/* KIN_WAVE_MIN_FLOW = 1e-30 and KIN_WAVE_MIN_FLOW^0.2 (Wflow/src/routing/utils.jl:1-2); the
 * root is the correctly rounded value of (binary 1e-30)^(binary 0.2). */
#define WFH_KIN_WAVE_MIN_FLOW 1e-30
#define WFH_KIN_WAVE_MIN_FLOW_QROOT 9.999999999999993e-07

WFH_FN uint64_t wfh_bits(double x) {
    uint64_t u;
    memcpy(&u, &x, sizeof u);
    return u;
}
WFH_FN double wfh_from_bits(uint64_t u) {
    double x;
    memcpy(&x, &u, sizeof x);
    return x;
}
WFH_FN int wfh_isnan(double x) { return x != x; }
WFH_FN int wfh_signbit(double x) { return (int)(wfh_bits(x) >> 63); }
WFH_FN double wfh_fabs(double x) { return wfh_from_bits(wfh_bits(x) & 0x7fffffffffffffffULL); }

#if defined(__HIPCC__)
#define WFH_FMA(a, b, c) __builtin_fma((a), (b), (c))
#define WFH_RINT(x) __builtin_rint(x)
#define WFH_FLOOR(x) __builtin_floor(x)
#define WFH_CEIL(x) __builtin_ceil(x)
#else
#include <math.h>
#define WFH_FMA(a, b, c) fma((a), (b), (c))
#define WFH_RINT(x) rint(x)
#define WFH_FLOOR(x) floor(x)
#define WFH_CEIL(x) ceil(x)
#endif

/* A constant of the series and reductions below. On the GPU it is materialised in scalar
 * registers where it is used (an empty volatile asm takes it as an operand), instead of being
 * hoisted out of a kernel's long loops and kept in vector registers, or spilled, for the whole
 * loop. The value, and so every result, is the same. */
#if defined(__HIP_DEVICE_COMPILE__)
__device__ static inline double wfh_k(double c) {
    asm volatile("" : "+s"(c));
    return c;
}
#define WFH_K(c) wfh_k(c)
#else
#define WFH_K(c) (c)
#endif

/* 1 when any of a, b, c is NaN or infinite (all exponent bits set) */
WFH_FN int wfh_nonfinite3(double a, double b, double c) {
    const uint64_t e = 0x7ff0000000000000ULL;
    return ((wfh_bits(a) & e) == e) | ((wfh_bits(b) & e) == e) | ((wfh_bits(c) & e) == e);
}

/* 2^k for -1022 <= k <= 1023 */
WFH_FN double wfh_pow2i(int k) { return wfh_from_bits((uint64_t)(k + 1023) << 52); }

WFH_FN double wfh_exp(double x) {
    if (wfh_isnan(x)) return x;
    if (x > 709.782712893384) return wfh_from_bits(0x7ff0000000000000ULL);
    if (x < -745.1332191019412) return 0.0;
    const double inv_ln2 = WFH_K(0x1.71547652b82fep+0);
    const double ln2_hi = WFH_K(0x1.62e42ff000000p-1);
    const double ln2_lo = WFH_K(-0x1.718432a1b0e26p-35);
    const double k = WFH_RINT(x * inv_ln2);
    double r = WFH_FMA(-k, ln2_hi, x);
    r = WFH_FMA(-k, ln2_lo, r);
    double p = WFH_K(1.1470745597729725e-11);         /* 1/14! */
    p = WFH_FMA(p, r, WFH_K(1.6059043836821613e-10)); /* 1/13! */
    p = WFH_FMA(p, r, WFH_K(2.08767569878681e-09));   /* 1/12! */
    p = WFH_FMA(p, r, WFH_K(2.505210838544172e-08));  /* 1/11! */
    p = WFH_FMA(p, r, WFH_K(2.755731922398589e-07));  /* 1/10! */
    p = WFH_FMA(p, r, WFH_K(2.7557319223985893e-06)); /* 1/9! */
    p = WFH_FMA(p, r, WFH_K(2.48015873015873e-05));   /* 1/8! */
    p = WFH_FMA(p, r, WFH_K(0.0001984126984126984));  /* 1/7! */
    p = WFH_FMA(p, r, WFH_K(0.001388888888888889));   /* 1/6! */
    p = WFH_FMA(p, r, WFH_K(0.008333333333333333));   /* 1/5! */
    p = WFH_FMA(p, r, WFH_K(0.041666666666666664));   /* 1/4! */
    p = WFH_FMA(p, r, WFH_K(0.16666666666666666));    /* 1/3! */
    p = WFH_FMA(p, r, 0.5);
    p = WFH_FMA(p, r, 1.0);
    p = WFH_FMA(p, r, 1.0);
    const int ki = (int)k;
    if (ki > 1023) return p * 2.0 * wfh_pow2i(ki - 1);
    if (ki < -1021) return (p * wfh_pow2i(ki + 600)) * wfh_pow2i(-600);
    return p * wfh_pow2i(ki);
}

WFH_FN double wfh_log(double x) {
    if (wfh_isnan(x)) return x;
    if (x < 0.0) return wfh_from_bits(0x7ff8000000000000ULL);
    if (x == 0.0) return wfh_from_bits(0xfff0000000000000ULL);
    if (x == wfh_from_bits(0x7ff0000000000000ULL)) return x;
    int k = 0;
    if (x < 0x1p-1022) { /* subnormal: scale into the normal range */
        x *= 0x1p54;
        k = -54;
    }
    uint64_t u = wfh_bits(x);
    k += (int)((u >> 52) & 0x7ff) - 1023;
    u = (u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL; /* m in [1, 2) */
    double m = wfh_from_bits(u);
    if (m > 1.4142135623730951) {
        m *= 0.5;
        k += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = WFH_K(0.08);                           /* 2/25 */
    R = WFH_FMA(R, z, WFH_K(0.08695652173913043));     /* 2/23 */
    R = WFH_FMA(R, z, WFH_K(0.09523809523809523));     /* 2/21 */
    R = WFH_FMA(R, z, WFH_K(0.10526315789473684));     /* 2/19 */
    R = WFH_FMA(R, z, WFH_K(0.11764705882352941));     /* 2/17 */
    R = WFH_FMA(R, z, WFH_K(0.13333333333333333));     /* 2/15 */
    R = WFH_FMA(R, z, WFH_K(0.15384615384615385));     /* 2/13 */
    R = WFH_FMA(R, z, WFH_K(0.18181818181818182));     /* 2/11 */
    R = WFH_FMA(R, z, WFH_K(0.2222222222222222));      /* 2/9 */
    R = WFH_FMA(R, z, WFH_K(0.2857142857142857));      /* 2/7 */
    R = WFH_FMA(R, z, WFH_K(0.4));                     /* 2/5 */
    R = WFH_FMA(R, z, WFH_K(0.6666666666666666));      /* 2/3 */
    R = R * z;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)k;
    const double ln2_hi = WFH_K(0x1.62e42ff000000p-1);
    const double ln2_lo = WFH_K(-0x1.718432a1b0e26p-35);
    return dk * ln2_hi + (f - (hfsq - (s * (hfsq + R) + dk * ln2_lo)));
}

WFH_FN double wfh_pow(double x, double y) { return wfh_exp(y * wfh_log(x)); }

WFH_FN double wfh_cbrt(double x) {
    if (x == 0.0 || wfh_isnan(x) || wfh_fabs(x) == wfh_from_bits(0x7ff0000000000000ULL)) return x;
    const double a = wfh_fabs(x);
    double y = wfh_exp(wfh_log(a) * 0.3333333333333333);
    y = y - (y - a / (y * y)) * 0.3333333333333333;
    return wfh_signbit(x) ? -y : y;
}

/* Julia Base.min / Base.max for Float64: NaN propagates, -0.0 < 0.0. */
WFH_FN double wfh_min(double x, double y) {
    const double d = x - y;
    if (wfh_isnan(x) || wfh_isnan(y)) return d;
    return wfh_signbit(d) ? x : y;
}
WFH_FN double wfh_max(double x, double y) {
    const double d = x - y;
    if (wfh_isnan(x) || wfh_isnan(y)) return d;
    return wfh_signbit(d) ? y : x;
}
/* Julia Base.clamp: ifelse(x > hi, hi, ifelse(x < lo, lo, x)) */
WFH_FN double wfh_clamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

/* Julia cld(x::Float64, y::Float64) for y > 0: the smallest integer q with q*y >= x exactly.
 * The fma residual has the sign of the exact residual, so the correction is exact. */
WFH_FN double wfh_cld(double x, double y) {
    double q = WFH_CEIL(x / y);
    while (WFH_FMA(q - 1.0, y, -x) >= 0.0) q -= 1.0;
    while (WFH_FMA(q, y, -x) < 0.0) q += 1.0;
    return q + 0.0; /* -0.0 -> 0.0 */
}

/* 10.0^n for n >= 0 as Julia evaluates Float64^Integer (Base.Math.pow_body: power by
 * squaring carrying the rounding error of every product in a second word, muladd taken as a
 * fused multiply-add), which gives the correctly rounded power for the exponents used here. */
WFH_FN double wfh_pow10i(int n) {
    if (n == 0) return 1.0;
    double x = 10.0, y = 1.0, xnlo = 0.0, ynlo = 0.0;
    if (n == 3) return x * x * x;
    while (n > 1) {
        if (n & 1) {
            const double err = WFH_FMA(y, xnlo, x * ynlo);
            const double hi = x * y;
            ynlo = WFH_FMA(x, y, -hi) + err;
            y = hi;
        }
        const double err = x * 2.0 * xnlo;
        const double hi = x * x;
        xnlo = WFH_FMA(x, x, -hi) + err;
        x = hi;
        n >>= 1;
    }
    const double err = WFH_FMA(y, xnlo, x * ynlo);
    const double r = x * y;
    return (r - r == 0.0 && err - err == 0.0) ? WFH_FMA(x, y, err) : r;
}
/* Julia round(x; sigdigits = 12) for finite x (Base._round_sigdigits / _round_digits):
 * h = 1 + floor(log10|x|); d = 12 - h; d >= 0 -> round(x * 10^d) / 10^d,
 * d < 0 -> round(x / 10^-d) * 10^-d, RoundNearest ties to even. The decimal exponent is
 * the exact floor(log10|x|); Julia takes the floor of its rounded log10, which differs only
 * for x within a few ulps below a power of ten, where both exponents give the same result as
 * long as the powers of ten involved are exact (|x| and 10^d within 1e-10 .. 1e22). */
WFH_FN double wfh_round_sig12(double x) {
    if (x == 0.0 || wfh_isnan(x)) return x;
    const double ax = wfh_fabs(x);
    /* floor(log10 ax) from the natural log, corrected against exact powers of ten */
    int e = (int)WFH_FLOOR(wfh_log(ax) * 0.43429448190325176);
    if (e < -300) e = -300;
    if (e > 300) e = 300;
    for (int it = 0; it < 2; ++it) {
        const double lo = e >= 0 ? wfh_pow10i(e) : 1.0 / wfh_pow10i(-e);
        const double hi = (e + 1) >= 0 ? wfh_pow10i(e + 1) : 1.0 / wfh_pow10i(-(e + 1));
        if (ax >= hi) e += 1;
        else if (ax < lo) e -= 1;
    }
    const int d = 12 - (1 + e);
    double r;
    if (d >= 0) {
        const double sc = wfh_pow10i(d);
        r = WFH_RINT(x * sc) / sc;
    } else {
        const double isc = wfh_pow10i(-d);
        r = WFH_RINT(x / isc) * isc;
    }
    return (wfh_fabs(r) <= 1.7976931348623157e308) ? r : x;
}

/* Statistics.quantile!(v, p) of n >= 2 values with the default definition (Hyndman and Fan
 * type 7, alpha = beta = 1), as Statistics.jl 1.11.1 (the version pinned in Manifest.toml)
 * evaluates it: m = alpha + p (1 - alpha - beta), aleph = fma(n, p, m),
 * j = clamp(trunc(aleph), 1, n - 1), gamma = clamp(aleph - j, 0, 1); with a = v(j) and
 * b = v(j+1) the j-th and (j+1)-th smallest values, the result is a + gamma (b - a) when both are
 * finite and a ~ b (isapprox: |a - b| <= sqrt(eps) max(|a|, |b|)), otherwise
 * (1 - gamma) a + gamma b. wfh_quantile_index gives j (1-based); the reference calls it at
 * Wflow/src/routing/surface/surface_kinwave.jl:698. */
WFH_FN int64_t wfh_quantile_index(int64_t n, double p, double* gamma) {
    const double m = 1.0 + p * ((1.0 - 1.0) - 1.0);
    const double aleph = WFH_FMA((double)n, p, m);
    int64_t j = (int64_t)aleph; /* trunc */
    if (j < 1) j = 1;
    if (j > n - 1) j = n - 1;
    double g = aleph - (double)j;
    g = g > 1.0 ? 1.0 : (g < 0.0 ? 0.0 : g);
    *gamma = g;
    return j;
}
WFH_FN double wfh_quantile_interp(double a, double b, double gamma) {
    const double inf = 1.7976931348623157e308;
    const int fin = wfh_fabs(a) <= inf && wfh_fabs(b) <= inf;
    const double aa = wfh_fabs(a), ab = wfh_fabs(b);
    const int approx = a == b || (fin && wfh_fabs(a - b) <= 1.4901161193847656e-8 * (aa > ab ? aa : ab));
    if (fin && approx) return a + gamma * (b - a);
    return (1.0 - gamma) * a + gamma * b;
}

#endif /* WFH_MATH_H */
