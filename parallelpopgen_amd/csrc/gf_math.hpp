/*
 * gf_math.hpp -- random numbers and elementary functions shared by the
 * gfx950 kernels and the library's host code.
 *
 * Every function is built from IEEE binary32 +, -, *, / and sqrt (all
 * correctly rounded on gfx950 and on x86-64) in a fixed order; the library
 * is compiled with -ffp-contract=off, so host and device evaluate the same
 * operation sequence and agree bit for bit.  No device libm (OCML) or
 * hardware approximation (v_exp_f32, v_log_f32, v_rcp_f32) is used on any
 * value that reaches a simulation result.
 */
#ifndef GF_MATH_HPP_
#define GF_MATH_HPP_

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GF_HD __host__ __device__ __forceinline__
/* first statement of every floating-point function here: no FMA
 * contraction whatever -ffp-contract the including translation unit uses
 * (hipcc's default honours this pragma), so device results equal the host
 * code's and the CPU oracle's bit for bit */
#define GF_STRICT_FP _Pragma("clang fp contract(off)")

namespace gf {

/* ----- bit casts ----- */
GF_HD float bits2f(uint32_t b) { return __builtin_bit_cast(float, b); }
GF_HD uint32_t f2bits(float f) { return __builtin_bit_cast(uint32_t, f); }

/* ----- Philox4x32-10 -----
 * Salmon et al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11);
 * same round function and key schedule as Random123's philox4x32 with
 * R = 10 (the generator of reference shared.cuh:67-82). */
struct u32x4 { uint32_t x, y, z, w; };

GF_HD void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
    GF_STRICT_FP
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

GF_HD u32x4 philox10(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    GF_STRICT_FP
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xD2511F53u, c0, hi0, lo0);
        mulhilo(0xCD9E8D57u, c2, hi1, lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    return u32x4{c0, c1, c2, c3};
}

GF_HD uint32_t lane_of(const u32x4 &r, int i) {
    GF_STRICT_FP
    return i == 0 ? r.x : (i == 1 ? r.y : (i == 2 ? r.z : r.w));
}

/* u * 2^-32 + 2^-33 in (0, 1] (reference shared.cuh:59-64) */
GF_HD float u01(uint32_t u) {
    GF_STRICT_FP
    float f = (float)u;
    f = f * 2.3283064365386963e-10f;
    return f + 1.1641532182693481e-10f;
}

/* ----- elementary functions ----- */
GF_HD float pow2i(int k) { return bits2f((uint32_t)(k + 127) << 23); }

GF_HD float scale2(float p, int n) {
    GF_STRICT_FP
    if (n > 127) return (p * pow2i(127)) * pow2i(n - 127);
    if (n < -126) return (p * pow2i(n + 100)) * pow2i(-100);
    return p * pow2i(n);
}

/* exp(x): Cody-Waite split of ln 2 (0.693359375 + -2.12194440e-4), degree-7
 * Taylor polynomial on |r| <= ln2/2 by Horner's rule, each step and the
 * reduction one fused multiply-add; exact scaling by 2^n. */
GF_HD float exp_poly(float x, float &n) {
    GF_STRICT_FP
    const float t = x * 1.44269502162933349609375f;
    n = __builtin_floorf(t + 0.5f);
    float r = __builtin_fmaf(-n, 0.693359375f, x);
    r = __builtin_fmaf(n, 2.12194440e-4f, r);
    float p = 1.98412698e-4f;
    p = __builtin_fmaf(p, r, 1.38888889e-3f);
    p = __builtin_fmaf(p, r, 8.33333333e-3f);
    p = __builtin_fmaf(p, r, 4.16666667e-2f);
    p = __builtin_fmaf(p, r, 1.66666667e-1f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    return __builtin_fmaf(p, r, 1.0f);
}

GF_HD float exp_f(float x) {
    GF_STRICT_FP
    if (x != x) return x;
    if (x > 88.72283935546875f) return __builtin_inff();
    if (x < -103.97208404541015625f) return 0.0f;
    float n;
    const float p = exp_poly(x, n);
    return scale2(p, (int)n);
}

/* log(x): x = m 2^e, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
 * s = (m-1)/(m+1), odd series through s^9. */
GF_HD float log_f(float x) {
    GF_STRICT_FP
    if (x != x || x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }
    uint32_t b = f2bits(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = bits2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float s2 = s * s;
    float q = 0.111111111f;
    q = q * s2 + 0.142857143f;
    q = q * s2 + 0.2f;
    q = q * s2 + 0.333333333f;
    float lm = (s + s) + (s + s) * (s2 * q);
    float fe = (float)e;
    return fe * 0.693359375f + (lm + fe * -2.12194440e-4f);
}

GF_HD float pow_f(float a, float b) {
    GF_STRICT_FP
    if (a == 1.0f || b == 0.0f) return 1.0f;
    if (a == 0.0f) return b > 0.0f ? 0.0f : __builtin_inff();
    return exp_f(b * log_f(a));
}

/* Correctly rounded square root of a positive normal float x >= 2^-96 (the
 * variance of the normal approximation, at least 36/N there, and w >= 8 in
 * the inverse normal's tail).  On the
 * device: the hardware estimate (within one ulp) moved by one ulp where the
 * fused residual x - y*y says so, which is the compiler's own correctly
 * rounded expansion without its scaling of tiny inputs and its zero,
 * infinity and NaN cases; the host's sqrtf is correctly rounded. */
GF_HD float sqrt_pos(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float y = __builtin_amdgcn_sqrtf(x);
    const float ydn = bits2f(f2bits(y) - 1u), yup = bits2f(f2bits(y) + 1u);
    float r = __builtin_fmaf(-ydn, y, x) <= 0.0f ? ydn : y;
    r = __builtin_fmaf(-yup, y, x) > 0.0f ? yup : r;
    return r;
#else
    return __builtin_sqrtf(x);
#endif
}

/* Inverse standard normal CDF for p in [2^-33, 1] (the range of u01; p is
 * clamped into [2^-33, 1 - 2^-24], NaN to 2^-33), from a segment
 * table (gf_ncdf_table.h, fitted by parallelpopgen_amd/csrc/fit_normcdfinv.py):
 * with s = min(p, 1 - p) in [2^-33, 1/2], the float bits of s pick one of
 * 16 segments per binade (biased exponent and top 4 mantissa bits; the
 * segments are logarithmically spaced towards the tail) and the low 19
 * mantissa bits, read as an integer t, feed that segment's cubic
 * c0 + t (c1 + t (c2 + t c3)) approximating ndtri(s); the result is r for
 * p < 1/2 and -r above.  No division, logarithm or branch.  Error against
 * scipy.special.ndtri over u01's values: at most 1.5 ulp of the result
 * where |z| >= 0.1, 4.8e-7 absolute everywhere.  `tab` is the table itself
 * on the host and in the library's less frequent kernels, and a copy in LDS
 * in the drift kernels (8 KB). */
#include "gf_ncdf_table.h"

struct alignas(16) NcdfSeg { float c0, c1, c2, c3; };

GF_HD const NcdfSeg *ncdf_table() { return reinterpret_cast<const NcdfSeg *>(kNcdfSeg); }

/* the table's function on its domain, 2^-33 <= p <= 1 - 2^-24 */
GF_HD float normcdfinv_core(const NcdfSeg *tab, float p) {
    GF_STRICT_FP
    const float q = p - 0.5f;
    const float s = q < 0.0f ? p : 1.0f - p;        /* [2^-33, 1/2], exact */
    const uint32_t b = f2bits(s);
    const NcdfSeg c = tab[(b >> GF_NCDF_T_BITS) - GF_NCDF_BASE];
    const float t = (float)(b & ((1u << GF_NCDF_T_BITS) - 1u));
    const float r = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c.c3, t, c.c2), t, c.c1), t, c.c0);
    return q < 0.0f ? r : -r;
}

GF_HD float normcdfinv_tab(const NcdfSeg *tab, float p) {
    GF_STRICT_FP
    /* the table's domain [2^-33, 1 - 2^-24] (u01's range, its top end
     * clamped): anything below 2^-33, and NaN, is taken as 2^-33, so the
     * segment index stays in the table for every input */
    p = p > 0x1p-33f ? p : 0x1p-33f;
    if (p > 0.99999994039535522f) p = 0.99999994039535522f;
    return normcdfinv_core(tab, p);
}

GF_HD float normcdfinv_f(float p) { return normcdfinv_tab(ncdf_table(), p); }

GF_HD float round_away(float v) {
    GF_STRICT_FP
    float t = __builtin_truncf(v);
    float d = v - t;
    if (__builtin_fabsf(d) >= 0.5f) t = t + (v < 0.0f ? -1.0f : 1.0f);
    return t;
}

/* float -> int32 with saturation (NaN -> 0, >= 2^31 -> INT_MAX,
 * <= -2^31 -> INT_MIN), truncating toward zero otherwise.  On the device
 * this is exactly v_cvt_i32_f32's defined behaviour. */
GF_HD int f2i_sat(float v) {
    GF_STRICT_FP
#if defined(__HIP_DEVICE_COMPILE__)
    int r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
#else
    if (v != v) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int)v;
#endif
}

/* early-exit means of the truncated CDF inversions: after term j the
 * search stops once mean <= exit_mean(j); -1 = no early exit there
 * (reference shared.cuh:105-163 and :188-246) */
GF_HD int exit_mean(int j) {
    GF_STRICT_FP
    switch (j) {
    case 11: return 1;  case 14: return 2;  case 17: return 3;  case 20: return 4;
    case 22: return 5;  case 24: return 6;  case 26: return 7;  case 28: return 8;
    case 30: return 9;  case 32: return 10; case 34: return 11; case 36: return 12;
    case 38: return 13; case 40: return 14; case 41: return 15; case 43: return 16;
    case 45: return 17; case 46: return 18; case 48: return 19; case 50: return 20;
    case 51: return 21; case 53: return 22; case 55: return 23; case 56: return 24;
    case 58: return 25; case 59: return 26; case 61: return 27; case 62: return 28;
    case 64: return 29; case 65: return 30; case 67: return 31; case 68: return 32;
    default: return -1;
    }
}

/* correctly rounded reciprocals fl(1/j), j = 1..69, of the Poisson CDF
 * recurrence emu_j = emu_(j-1) * (mean * (1/j)) */
alignas(16) static constexpr float kRecip[72] = {
    0x1.0000000000000p+0f, 0x1.0000000000000p-1f, 0x1.5555560000000p-2f, 0x1.0000000000000p-2f,
    0x1.99999a0000000p-3f, 0x1.5555560000000p-3f, 0x1.24924a0000000p-3f, 0x1.0000000000000p-3f,
    0x1.c71c720000000p-4f, 0x1.99999a0000000p-4f, 0x1.745d180000000p-4f, 0x1.5555560000000p-4f,
    0x1.3b13b20000000p-4f, 0x1.24924a0000000p-4f, 0x1.1111120000000p-4f, 0x1.0000000000000p-4f,
    0x1.e1e1e20000000p-5f, 0x1.c71c720000000p-5f, 0x1.af286c0000000p-5f, 0x1.99999a0000000p-5f,
    0x1.8618620000000p-5f, 0x1.745d180000000p-5f, 0x1.642c860000000p-5f, 0x1.5555560000000p-5f,
    0x1.47ae140000000p-5f, 0x1.3b13b20000000p-5f, 0x1.2f684c0000000p-5f, 0x1.24924a0000000p-5f,
    0x1.1a7b960000000p-5f, 0x1.1111120000000p-5f, 0x1.0842100000000p-5f, 0x1.0000000000000p-5f,
    0x1.f07c200000000p-6f, 0x1.e1e1e20000000p-6f, 0x1.d41d420000000p-6f, 0x1.c71c720000000p-6f,
    0x1.bacf920000000p-6f, 0x1.af286c0000000p-6f, 0x1.a41a420000000p-6f, 0x1.99999a0000000p-6f,
    0x1.8f9c180000000p-6f, 0x1.8618620000000p-6f, 0x1.7d05f40000000p-6f, 0x1.745d180000000p-6f,
    0x1.6c16c20000000p-6f, 0x1.642c860000000p-6f, 0x1.5c98820000000p-6f, 0x1.5555560000000p-6f,
    0x1.4e5e0a0000000p-6f, 0x1.47ae140000000p-6f, 0x1.4141420000000p-6f, 0x1.3b13b20000000p-6f,
    0x1.3521d00000000p-6f, 0x1.2f684c0000000p-6f, 0x1.29e4120000000p-6f, 0x1.24924a0000000p-6f,
    0x1.1f70480000000p-6f, 0x1.1a7b960000000p-6f, 0x1.15b1e60000000p-6f, 0x1.1111120000000p-6f,
    0x1.0c97140000000p-6f, 0x1.0842100000000p-6f, 0x1.0410420000000p-6f, 0x1.0000000000000p-6f,
    0x1.f81f820000000p-7f, 0x1.f07c200000000p-7f, 0x1.e9131a0000000p-7f, 0x1.e1e1e20000000p-7f,
    0x1.dae6080000000p-7f,
    0.0f, 0.0f, 0.0f,   /* padding: the inversion reads four entries at a time */
};

/* First term index at which the truncated inversion stops whatever the
 * CDF: the smallest j with exit_mean(j) >= mean, 70 if none.  exit_mean is
 * increasing in j, so for mean <= 32 this is the j with
 * exit_mean(j) = max(1, ceil(mean)). */
static constexpr signed char kExitIndex[33] = {
    11, 11, 14, 17, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 41, 43,
    45, 46, 48, 50, 51, 53, 55, 56, 58, 59, 61, 62, 64, 65, 67, 68};

GF_HD int exit_index(float mean) {
    GF_STRICT_FP
    if (!(mean <= 32.0f)) return 70;
    int k = (int)__builtin_ceilf(mean);
    k = k < 1 ? 1 : k;
    if (k <= 4) return 8 + 3 * k;
    if (k <= 14) return 12 + 2 * k;
    return kExitIndex[k];
}

/* exp_f without its range checks: the same operations, so the same value,
 * for -87 <= x <= 88 (no NaN, no overflow, the scale 2^n stays normal) */
GF_HD float exp_f_mid(float x) {
    GF_STRICT_FP
    float n;
    const float p = exp_poly(x, n);
    return p * pow2i((int)n);
}

/* exp(-m) for 0 <= m <= 6 (the first term of the Poisson inversion), from
 * a segment table (gf_expneg_table.h, fitted by parallelpopgen_amd/csrc/
 * fit_expneg.py): segment k = min(floor(16 m), 95), a cubic in t =
 * m - k/16 (exact), Horner's rule with fused multiply-adds; exp(-0) = 1.
 * Error against exp at most 1.5 ulp over every float in [0, 6]
 * (tests/cpp/tables.hip).  9 VALU operations against exp_f_mid's 17.
 * `tab` is the table itself, or its LDS copy in the drift kernels. */
#include "gf_expneg_table.h"

struct alignas(16) ExpSeg { float c0, c1, c2, c3; };

GF_HD const ExpSeg *expneg_table() { return reinterpret_cast<const ExpSeg *>(kExpNegSeg); }

GF_HD float expneg_tab(const ExpSeg *tab, float m) {
    GF_STRICT_FP
    int k = f2i_sat(m * 16.0f);   /* m >= 0: truncation is floor */
    /* kept in the table for any input (a negative m, possible only with
     * h outside [0, 1], extrapolates segment 0) */
    k = k < GF_EXPNEG_SEGMENTS - 1 ? k : GF_EXPNEG_SEGMENTS - 1;
    k = k > 0 ? k : 0;
    const float t = __builtin_fmaf((float)k, -0.0625f, m);
    const ExpSeg c = tab[k];
    return __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c.c3, t, c.c2), t, c.c1), t, c.c0);
}

/* Poisson CDF inversion truncated as in the reference (shared.cuh:168-249):
 * the first j with CDF(j) >= r, or the early-exit index of `mean` if that
 * comes first, or 70.  The first term exp(-mean) comes from the table
 * for 0 <= mean <= 6 (the draws' Poisson tails), else from exp_f_mid
 * (equal to exp_f for the means callers pass, up to 87).  The
 * loop stays rolled (lanes leave it as they finish) and takes four
 * reciprocals per trip in one scalar load.  A trip computes its four terms
 * straight through; the partial sums d0 <= d1 <= d2 <= d3 are
 * non-decreasing (every term is >= 0), so the trip holds the answer iff
 * d3 >= r or the early exit falls inside it, and then the answer is the
 * number of partial sums below r, capped at the early exit.  The
 * recurrence and the comparisons are those of one term at a time. */
GF_HD int poiscdfinv_from(float r, float mean, float emu);

GF_HD int poiscdfinv(float r, float mean, const ExpSeg *etab = expneg_table()) {
    GF_STRICT_FP
    return poiscdfinv_from(r, mean, (mean >= 0.0f && mean <= 6.0f) ? expneg_tab(etab, mean) : exp_f_mid(-1.0f * mean));
}

/* the draws' Poisson tails: 0 < mean <= 6, so always the table (the same
 * value as poiscdfinv) */
GF_HD int poiscdfinv_tail(float r, float mean, const ExpSeg *etab) {
    GF_STRICT_FP
    return poiscdfinv_from(r, mean, expneg_tab(etab, mean));
}

/* the inversion given its first term emu = exp(-mean) */
GF_HD int poiscdfinv_from(float r, float mean, float emu) {
    GF_STRICT_FP
    float cdf = emu;
    if (cdf >= r) return 0;
    /* exit_index for mean <= 14 without its branches (callers pass <= 6) */
    const float cm = __builtin_ceilf(mean);
    const int km = cm < 1.0f ? 1 : (int)cm;
    const int over = km > 4 ? km - 4 : 0;
    int jstop = 8 + 3 * km - over;   /* 8 + 3 km up to 4, 12 + 2 km beyond */
    if (!(mean <= 14.0f)) jstop = exit_index(mean);
#pragma nounroll
    for (int j = 1;; j += 4) {
        const float c0 = kRecip[j - 1], c1 = kRecip[j], c2 = kRecip[j + 1], c3 = kRecip[j + 2];
        const float e0 = emu * (mean * c0), d0 = cdf + e0;
        const float e1 = e0 * (mean * c1), d1 = d0 + e1;
        const float e2 = e1 * (mean * c2), d2 = d1 + e2;
        const float e3 = e2 * (mean * c3), d3 = d2 + e3;
        const int rel = jstop - j;   /* the early exit, relative to this trip (>= 0; <= 1 at j = 69,
                                      * whose three padding terms are 0) */
        if (d3 >= r || rel <= 3) {
            const int k = (int)(d0 < r) + (int)(d1 < r) + (int)(d2 < r) + (int)(d3 < r);
            return j + (k < rel ? k : rel);
        }
        emu = e3;
        cdf = d3;
    }
}

/* binomial CDF inversion (N < 100 only): kept out of line, it is off the
 * hot path */
static __host__ __device__ __attribute__((noinline)) int binomcdfinv(float r, float mean, float x, float n) {
    GF_STRICT_FP
    float emu = pow_f(1.0f - x, n);
    if (emu == 1.0f) emu = exp_f(-1.0f * mean);
    float cdf = emu;
    if (cdf >= r) return 0;
    for (int j = 1; j < 70; j++) {
        float fj = (float)j;
        emu = emu * (((n + 1.0f - fj) * x) / (fj * (1.0f - x)));
        cdf = cdf + emu;
        if (cdf >= r) return j;
        int em = exit_mean(j);
        if (em >= 0 && mean <= (float)em) return j;
    }
    return 70;
}

/* round_away(v) as an integer, for the caller's clamp to [0, N]: equal to
 * f2i_sat(round_away(v)) for v >= 0; <= 0 for v < 0, like round_away */
GF_HD int round_away_nonneg(float v) {
    GF_STRICT_FP
    const float d = v - __builtin_truncf(v);
    return f2i_sat(v) + (d >= 0.5f ? 1 : 0);
}

/* approximate binomial draw (reference shared.cu:10-18): the two Poisson
 * (or binomial, N < 100) tails share one CDF inversion; the rest is the
 * rounded normal approximation.  A mean of exactly 0 or N (lost or fixed:
 * the Poisson inversion of mean 0 returns 0) is answered without entering
 * the inversion.  The result equals the reference's draw wherever that is
 * >= 0; where the normal approximation goes below 0 it is some value <= 0
 * (round_away_nonneg), and the caller clamps the result to [0, N]. */
GF_HD int binom_draw(uint32_t bits, float mean, float var, float N, const NcdfSeg *ncdf = ncdf_table(),
                     const ExpSeg *etab = expneg_table()) {
    GF_STRICT_FP
    const float u = u01(bits);
    const bool lo = mean <= 6.0f;
    const bool hi = !lo && (mean >= N - 6.0f);
    if (lo || hi) {
        const float m = lo ? mean : N - mean;
        int k = 0;
        if (m != 0.0f) k = (N < 100.0f) ? binomcdfinv(u, m, m / N, N) : poiscdfinv_tail(u, m, etab);
        return lo ? k : f2i_sat(N - (float)k);
    }
    /* u01 gives 2^-33 <= u <= 1 (never NaN): of normcdfinv_tab's clamps
     * only the top one can act, and the result is the same */
    return round_away_nonneg(normcdfinv_core(ncdf, u > 0.99999994039535522f ? 0.99999994039535522f : u) *
                                 sqrt_pos(var) + mean);
}

/* approximate Poisson draw (reference shared.cuh:251-256) */
GF_HD int pois_draw(uint32_t bits, float mean, float var, float N) {
    GF_STRICT_FP
    const float u = u01(bits);
    const bool lo = mean <= 6.0f;
    const bool hi = !lo && (mean >= N - 6.0f);
    if (lo || hi) {
        const int k = poiscdfinv(u, lo ? mean : N - mean);
        return lo ? k : f2i_sat(N - (float)k);
    }
    return f2i_sat(round_away(normcdfinv_f(u) * sqrt_pos(var) + mean));
}

/* (float)j / N, correctly rounded, for integers 0 <= j <= N <= 2^22 given
 * r = RN(1/N), in three float operations: q = j r is within 1.5 ulp of
 * j/N; e = j - q N is exact (a multiple of ulp(q) below 3N ulp(q) <
 * 2^24 ulp(q)); q + e r differs from j/N by at most |e| |r - 1/N| <
 * 1.5 ulp 2^-24, less than the distance from j/N to the nearest rounding
 * midpoint (at least ulp / 2N, and j/N with N <= 2^22 is never a
 * midpoint), so the fma's one rounding gives RN(j/N).  Checked
 * exhaustively for N <= 32768 and on sampled larger N by
 * tests/cpp/quotient.hip. */
GF_HD float quot_count(float fj, float N, float r) {
    GF_STRICT_FP
    const float q = fj * r;
    return __builtin_fmaf(__builtin_fmaf(-q, N, fj), r, q);
}

/* ----- model helpers ----- */

/* effective number of chromosomes (int)(2N/(1+F)) (reference go_fish_impl.cuh:659) */
GF_HD int nchrom(int N, float F) {
    GF_STRICT_FP
    float v = (float)(2 * N) / (1.0f + F);
    return f2i_sat(__builtin_truncf(v));
}

/* equilibrium frequency weight (reference go_fish_impl.cuh:20-22) */
GF_HD float mse_weight(float i, int N, float F, float h, float s) {
    GF_STRICT_FP
    float A = ((2.0f * h) + ((1.0f - (2.0f * h)) * i)) * (1.0f - F) + (2.0f * F);
    float e = ((((float)(2 * N) * s) * i) * A) / (1.0f + F);
    return exp_f(e);
}

/* Philox counter words of the drift draw of a mutation in population pop at
 * generation g: {|threadID|, origin_generation, origin_pop<<16 | pop, g>>2};
 * the draw is lane g & 3. */
GF_HD u32x4 drift_block(uint32_t k0, uint32_t k1, uint32_t tid, int ogen, int opop, int pop, int g) {
    GF_STRICT_FP
    return philox10(k0, k1, tid, (uint32_t)ogen, ((uint32_t)opop << 16) | (uint32_t)pop, (uint32_t)(g >> 2));
}

/* new-mutation count of population pop at generation g
 * (reference calc_new_mutations_Index, go_fish_impl.cuh:454-467; counter
 * {0, g, pop, 1}) */
GF_HD int new_mutation_count(int Nc, float mu, float L, uint32_t k0, uint32_t k1, int g, int pop) {
    GF_STRICT_FP
    if (Nc <= 0) return 0;
    float lambda = (mu * (float)Nc) * L;
    if (lambda == 0.0f) return 0;
    u32x4 r = philox10(k0, k1, 0u, (uint32_t)g, (uint32_t)pop, 1u);
    int n = pois_draw(r.x, lambda, lambda, (float)Nc * L);
    return n > 0 ? n : 0;
}

} /* namespace gf */

#endif /* GF_MATH_HPP_ */
