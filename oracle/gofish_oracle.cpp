/*
 * gofish_oracle.cpp -- CPU ORACLE (test infrastructure, not product code).
 *
 * Single-file, plain C++ restatement of the GO_Fish per-generation
 * Wright-Fisher update and of the Spectrum SFS binning of ParallelPopGen
 * 0.3.2, used only by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg as the checker.  The product (parallelpopgen_amd/) never
 * links, loads or calls anything in this directory.
 *
 * Build: g++ -O2 -ffp-contract=off (never -ffast-math / -Ofast), see Makefile.
 * All float arithmetic is IEEE binary32 with round-to-nearest; every
 * transcendental is computed by the functions below from +,-,*,/, sqrt and
 * explicit fused multiply-add (fmaf) only, so the GPU library (which
 * evaluates the same operation sequence) must agree bit for bit.
 *
 * Pinning (tests/test_oracle_math.py, tests/test_oracle_model.py; each pin
 * of the model functions is also shown to fail on a planted mistake, below):
 *   philox4x32_10      Random123 known-answer vectors and vectors produced by
 *                      the reference's vendored philox.h built with g++
 *                      (oracle/Makefile -> oracle/_ref/philox_kat)
 *   u01                closed form (power-of-two scaling)
 *   exp_f, log_f       error bound vs long double (numpy longdouble)
 *   normcdfinv_f       error bound vs scipy.special.ndtri over u01's range;
 *                      symmetry; monotonicity
 *   expneg_tab         error bound vs long double exp over [0, 6] (every
 *                      float there, tests/cpp/tables.hip); exp(-0) = 1;
 *                      non-increasing to within one ulp
 *   poiscdfinv/binomcdfinv  brute-force long-double CDF inversion with the
 *                      same 70-term truncation and mean thresholds
 *   binom_draw, pois_draw   first two moments over many draws
 *   pre_draw           Eq. 13 of the reference (go_fish_impl.cuh:105) in
 *                      float64, and its closed forms (s = 0 -> x; F = 1 ->
 *                      x(1+s)/(1+sx); h = 0, h = 1/2, h = 1; s clamped at
 *                      -1); migration as a float64 weighted sum; mean and
 *                      variance of the draw
 *   drift_one          moments of the drawn count over 1e5 mutation IDs in
 *                      the normal regime and both Poisson tails; 0 and 1
 *                      absorbing
 *   mse_lambda         a float64 restatement of go_fish_impl.cuh:20-71 with
 *                      its quirks, over h, F and gamma grids; the neutral
 *                      total 2 mu L Nc H(Nc-1); class counts ~ Poisson
 *   new_count          Poisson moments over many generations
 *   compact, keep_slot brute force on tiny arrays with present masks
 *                      (keep rule, stable order, preserve negation)
 *   pop_hist           numpy: round half away from zero, bin Nc folded to 0,
 *                      (int)num_sites
 *   sfs, joint_sfs     float64 binomial sums; total mass; marginals
 *   whole path         the reference's published speed-test SNP counts and
 *                      neutral theta/i over many seeds (both through the GPU
 *                      library, which equals this oracle bit for bit);
 *                      dadi-scenario regression against the reference's
 *                      examples/example_dadi/run_980.out and run_980_2.out
 *
 * ORC_PLANT (0 in every build but the pin tests' own): compile-time planted
 * mistakes, one per value, each a plausible slip in a model function; the
 * pin tests build each variant and check that some pin fails on it.
 *
 * Deliberate, documented divergences from the reference (all listed in the
 * project's design notes):
 *   - drift draws are keyed by mutation ID instead of array index, so
 *     results do not depend on compaction or batching;
 *   - the new-mutation count uses Philox counter {0, g, pop, 1} instead of
 *     {1, g, pop, 0} (the latter collides with ID-keyed drift counters);
 *   - the inverse normal CDF (CUDA's normcdfinvf in the reference, a closed
 *     library) is a segmented cubic table over the float bits of
 *     min(u, 1 - u), fitted by parallelpopgen_amd/csrc/fit_normcdfinv.py;
 *     u is clamped to at most 1 - 2^-24 (u == 1.0f would give +inf);
 *   - the Poisson inversion's first term exp(-mean) (CUDA's expf in the
 *     reference) is, for 0 <= mean <= 6, a segmented cubic table in mean,
 *     fitted by parallelpopgen_amd/csrc/fit_expneg.py;
 *   - a population that is not drifted in a generation (not yet arisen,
 *     extinct) gets frequency 0 instead of NaN/stale memory;
 *   - the MSE integral uses a fixed-order two-level float scan
 *     (chunks of 64) instead of CUB's nondeterministic scan;
 *   - the sample SFS accumulates in 2^-32 fixed point instead of float
 *     atomics.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdio.h>
#include <vector>
#include <thread>
#include <algorithm>
#include <string>
#include <functional>

#include "../include/gofish_capi.h"

#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

namespace orc {

/* ===================== RNG: Philox4x32-10 =====================
 * Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11.  Follows Random123 philox.h:284-330 (round, key bump) with the
 * constants of philox.h:233-250, as used by reference shared.cuh:67-82. */
struct U4 { uint32_t x, y, z, w; };

static inline U4 philox10(uint32_t k0, uint32_t k1, U4 c) {
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        U4 o = { hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0 };
        c = o;
    }
    return c;
}

static inline uint32_t lane(const U4 &r, int i) {
    return i == 0 ? r.x : (i == 1 ? r.y : (i == 2 ? r.z : r.w));
}

/* reference shared.cuh:59-64: u * 2^-32 + 2^-33, in (0, 1] */
static inline float u01(uint32_t u) {
    float f = (float)u;
    f = f * 2.3283064365386963e-10f;   /* 2^-32 */
    return f + 1.1641532182693481e-10f;  /* 2^-33 */
}

/* ===================== elementary functions ===================== */
static inline float bits2f(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline uint32_t f2bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/* 2^k for k in [-126, 127] */
static inline float pow2i(int k) { return bits2f((uint32_t)(k + 127) << 23); }

/* p * 2^n for n in [-151, 128], single rounding */
static inline float scale2(float p, int n) {
    if (n > 127) return (p * pow2i(127)) * pow2i(n - 127);
    if (n < -126) return (p * pow2i(n + 100)) * pow2i(-100);
    return p * pow2i(n);
}

/* exp: Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2, degree-7 Taylor
 * polynomial in Horner form (reduction and Horner steps fused
 * multiply-adds), exact power-of-two scaling. */
static float exp_f(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return INFINITY;
    if (x < -103.97208404541015625f) return 0.0f;
    float t = x * 1.44269502162933349609375f;      /* log2(e) rounded to float */
    float n = floorf(t + 0.5f);
    float r = fmaf(-n, 0.693359375f, x);            /* x - n ln2_hi (fused) */
    r = fmaf(n, 2.12194440e-4f, r);                 /* - n (ln2 - ln2_hi) */
    const float c[8] = {1.0f, 1.0f, 0.5f, 1.66666667e-1f, 4.16666667e-2f,   /* 1/k! */
                        8.33333333e-3f, 1.38888889e-3f, 1.98412698e-4f};
    float p = c[7];
    for (int k = 6; k >= 0; k--) p = fmaf(p, r, c[k]);
    return scale2(p, (int)n);
}

/* log: x = m 2^e with m in [sqrt(1/2), sqrt(2)); log(m) = 2 atanh(s),
 * s = (m-1)/(m+1), odd series to s^9. */
static float log_f(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }
    uint32_t b = f2bits(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = bits2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float s2 = s * s;
    float q = 0.111111111f;                          /* 1/9 */
    q = q * s2 + 0.142857143f;                       /* 1/7 */
    q = q * s2 + 0.2f;
    q = q * s2 + 0.333333333f;
    float lm = (s + s) + (s + s) * (s2 * q);
    float fe = (float)e;
    return fe * 0.693359375f + (lm + fe * -2.12194440e-4f);
}

/* a^b for a >= 0 (binomial CDF start term); exp(b log a) */
static float pow_f(float a, float b) {
    if (a == 1.0f || b == 0.0f) return 1.0f;
    if (a == 0.0f) return b > 0.0f ? 0.0f : INFINITY;
    return exp_f(b * log_f(a));
}

/* Inverse standard normal CDF, p in [2^-33, 1] (u01's range), clamped to
 * p <= 1 - 2^-24.  With s = min(p, 1 - p) in [2^-33, 1/2], the float bits
 * of s select a segment of the table kNcdfSeg (biased exponent and top 5
 * mantissa bits: 16 segments per binade, logarithmically spaced towards the
 * tail; the last segment is s = 1/2), and the low 19 mantissa bits, read
 * as an integer t, are the argument of that segment's cubic
 * c0 + t (c1 + t (c2 + t c3)) ~ ndtri(s), by Horner's rule with fused
 * multiply-adds; the result is r below 1/2 and -r above.  The table is the
 * generated header the library also uses (it defines the function, as
 * Philox's constants define the generator); it is pinned here against
 * scipy.special.ndtri. */
#include "../parallelpopgen_amd/csrc/gf_ncdf_table.h"

static float normcdfinv_f(float p) {
    p = p > 0x1p-33f ? p : 0x1p-33f;   /* below the table's domain, and NaN: 2^-33 */
    if (p > 0.99999994039535522f) p = 0.99999994039535522f;
    const float q = p - 0.5f;
    const float s = q < 0.0f ? p : 1.0f - p;
    const uint32_t b = f2bits(s);
    const float *c = kNcdfSeg + 4 * ((b >> GF_NCDF_T_BITS) - GF_NCDF_BASE);
    const float t = (float)(b & ((1u << GF_NCDF_T_BITS) - 1u));
    const float r = fmaf(fmaf(fmaf(c[3], t, c[2]), t, c[1]), t, c[0]);
    return q < 0.0f ? r : -r;
}

/* round half away from zero (CUDA round(), used at shared.cuh:255,266) */
static inline float round_away(float v) {
    float t = truncf(v);
    float d = v - t;
    if (fabsf(d) >= 0.5f) t = t + (v < 0.0f ? -1.0f : 1.0f);
    return t;
}

/* float -> int with saturation; NaN -> 0 */
static inline int f2i_sat(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int)v;
}

/* Early-exit table of the truncated CDF inversions (reference
 * shared.cuh:105-163, :188-246): after term j the search also stops when
 * mean <= exit_mean[j]; no early exit where the entry is negative. */
static const signed char exit_mean[70] = {
    -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,     /* 0-9   */
    -1,  1, -1, -1,  2, -1, -1,  3, -1, -1,     /* 10-19 */
     4, -1,  5, -1,  6, -1,  7, -1,  8, -1,     /* 20-29 */
     9, -1, 10, -1, 11, -1, 12, -1, 13, -1,     /* 30-39 */
    14, 15, -1, 16, -1, 17, 18, -1, 19, -1,     /* 40-49 */
    20, 21, -1, 22, -1, 23, 24, -1, 25, 26,     /* 50-59 */
    -1, 27, 28, -1, 29, 30, -1, 31, 32, -1      /* 60-69 */
};

/* correctly rounded reciprocals fl(1/j), j = 1..69 */
static const float recip[69] = {
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
};

/* exp(-m) for 0 <= m <= 6: the segment table of the library
 * (gf_expneg_table.h, fitted by parallelpopgen_amd/csrc/fit_expneg.py):
 * segment k = min(floor(16 m), 95), cubic in t = m - k/16 by Horner's rule
 * with fused multiply-adds.  Pinned against exp in long double. */
#include "../parallelpopgen_amd/csrc/gf_expneg_table.h"

static float expneg_tab(float m) {
    int k = f2i_sat(m * 16.0f);
    k = k < GF_EXPNEG_SEGMENTS - 1 ? k : GF_EXPNEG_SEGMENTS - 1;
    k = k > 0 ? k : 0;
    const float t = fmaf((float)k, -0.0625f, m);
    const float *c = kExpNegSeg + 4 * k;
    return fmaf(fmaf(fmaf(c[3], t, c[2]), t, c[1]), t, c[0]);
}

/* Poisson CDF inversion, reference shared.cuh:168-249.  The recurrence
 * multiplies by fl(1/j) where the reference divides by j (its
 * --use_fast_math build evaluates that quotient as x * rcp(j) as well).
 * The first term exp(-mean) is expneg_tab for 0 <= mean <= 6 (the draws'
 * Poisson tails), exp_f beyond. */
static int poiscdfinv_from(float r, float mean, float emu);

static int poiscdfinv(float r, float mean) {
    return poiscdfinv_from(r, mean, (mean >= 0.0f && mean <= 6.0f) ? expneg_tab(mean) : exp_f(-1.0f * mean));
}

/* the approximate binomial's Poisson tails (mean <= 6, or N - mean <= 6):
 * the table whatever the sign (a negative mean needs h outside [0, 1]) */
static int poiscdfinv_tail(float r, float mean) { return poiscdfinv_from(r, mean, expneg_tab(mean)); }

/* the inversion given its first term emu = exp(-mean) */
static int poiscdfinv_from(float r, float mean, float emu) {
    float cdf = emu;
    if (cdf >= r) return 0;
    for (int j = 1; j < 70; j++) {
        emu = emu * (mean * recip[j - 1]);
        cdf = cdf + emu;
        if (cdf >= r) return j;
        if (exit_mean[j] >= 0 && mean <= (float)exit_mean[j]) return j;
    }
    return 70;
}

/* Binomial CDF inversion, reference shared.cuh:84-166 */
static int binomcdfinv(float r, float mean, float x, float n) {
    float emu = pow_f(1.0f - x, n);
    if (emu == 1.0f) emu = exp_f(-1.0f * mean);
    float cdf = emu;
    if (cdf >= r) return 0;
    for (int j = 1; j < 70; j++) {
        float fj = (float)j;
        emu = emu * (((n + 1.0f - fj) * x) / (fj * (1.0f - x)));
        cdf = cdf + emu;
        if (cdf >= r) return j;
        if (exit_mean[j] >= 0 && mean <= (float)exit_mean[j]) return j;
    }
    return 70;
}

static const float MEAN_BOUNDARY_NORM = 6.0f;   /* shared.cuh:44 */
static const float N_BOUNDARY_POIS_BINOM = 100.0f; /* shared.cuh:45 */

/* Approximate binomial draw (reference shared.cu:10-18, ApproxRandBinomHelper) */
static int binom_draw(uint32_t rbits, float mean, float var, float N) {
    float u = u01(rbits);
    if (mean <= MEAN_BOUNDARY_NORM) {
        if (N < N_BOUNDARY_POIS_BINOM) return binomcdfinv(u, mean, mean / N, N);
        return poiscdfinv_tail(u, mean);
    } else if (mean >= N - MEAN_BOUNDARY_NORM) {
        if (N < N_BOUNDARY_POIS_BINOM) {
            float nm = N - mean;
            return f2i_sat(N - (float)binomcdfinv(u, nm, nm / N, N));
        }
        return f2i_sat(N - (float)poiscdfinv_tail(u, N - mean));
    }
#if ORC_PLANT == 15
    return f2i_sat(round_away(normcdfinv_f(u) * var + mean));   /* planted: variance for sd */
#else
    return f2i_sat(round_away(normcdfinv_f(u) * sqrtf(var) + mean));
#endif
}

/* Approximate Poisson draw (reference shared.cuh:251-256, ApproxRandPois1) */
static int pois_draw(uint32_t rbits, float mean, float var, float N) {
    float u = u01(rbits);
    if (mean <= MEAN_BOUNDARY_NORM) return poiscdfinv(u, mean);
    if (mean >= N - MEAN_BOUNDARY_NORM) return f2i_sat(N - (float)poiscdfinv(u, N - mean));
    return f2i_sat(round_away(normcdfinv_f(u) * sqrtf(var) + mean));
}

/* ===================== model helpers ===================== */

/* Nchrom_e = (int)(2*N/(1+F)), reference go_fish_impl.cuh:659 */
static inline int nchrom(int N, float F) {
    float v = (float)(2 * N) / (1.0f + F);
    return f2i_sat(truncf(v));
}

/* mse(), reference go_fish_impl.cuh:20-22 */
static inline float mse(float i, int N, float F, float h, float s) {
#if ORC_PLANT == 12
    float A = ((2.0f * h) + ((1.0f - (2.0f * h)) * i)) * (1.0f + F) + (2.0f * F);   /* planted: F-term sign */
#else
    float A = ((2.0f * h) + ((1.0f - (2.0f * h)) * i)) * (1.0f - F) + (2.0f * F);
#endif
    float e = ((((float)(2 * N) * s) * i) * A) / (1.0f + F);
    return exp_f(e);
}

static inline float sel_eval(float slope, float intercept, float x) {
    return slope * x + intercept;
}

/* One generation for one mutation across all P populations: migration,
 * selection (Eq. 13 of the GO_Fish paper), approximate binomial drift.
 * Reference go_fish_impl.cuh:93-124 (float4 body :97-109).  The draw is
 * lane (g & 3) of Philox({seed1,seed2}, {|tid|, origin_gen,
 * (origin_pop << 16) | pop, g >> 2}). */
struct GenRow {
    int P;
    int Nc[GF_MAX_POPULATIONS];          /* <= 0: population not drifted -> 0 */
    int present[GF_MAX_POPULATIONS];     /* demography(pop,g) > 0 */
    float F[GF_MAX_POPULATIONS], h[GF_MAX_POPULATIONS];
    float sa[GF_MAX_POPULATIONS], sb[GF_MAX_POPULATIONS];
    float mig[GF_MAX_POPULATIONS][GF_MAX_POPULATIONS];  /* [from][to] */
};

/* Migration and selection of population p (Nc > 0): the frequency after
 * selection xs and the mean and variance of the drift draw.  Migration is
 * the conservative weighted sum over the source populations; selection is
 * Eq. 13 of the reference, (s x^2 + x + (F+h-hF) s x (1-x)) /
 * (x^2 s + (F+2h-2hF) s x (1-x) + 1) with s = max(s, -1)
 * (go_fish_impl.cuh:97-106), restated as x + s x (x + c1 (1 - x)) over
 * 1 + s x (x + c2 (1 - x)) with fused multiply-adds. */
static inline void pre_draw(const GenRow &t, const float *x, int p, float &xs, float &mean, float &var) {
    float im = 0.0f;
    for (int k = 0; k < t.P; k++) im = im + t.mig[k][p] * x[k];
    float s = sel_eval(t.sa[p], t.sb[p], im);
#if ORC_PLANT != 11
    s = s > -1.0f ? s : -1.0f;
#endif
    float F = t.F[p], h = t.h[p];
#if ORC_PLANT == 2
    float c1 = (F + h) + (h * F);                         /* planted: F-term sign */
#else
    float c1 = (F + h) - (h * F);
#endif
    float c2 = (F + (2.0f * h)) - ((2.0f * h) * F);
#if ORC_PLANT == 1
    std::swap(c1, c2);                                    /* planted: c1 and c2 swapped */
#endif
    float om = 1.0f - im;
    float sx = s * im;
    float num = fmaf(sx, fmaf(c1, om, im), im);
    float den = fmaf(sx, fmaf(c2, om, im), 1.0f);
    xs = num / den;
    float N = (float)t.Nc[p];
    mean = xs * N;
#if ORC_PLANT == 10
    var = xs * mean;                                      /* planted: x for 1 - x */
#else
    var = (1.0f - xs) * mean;
#endif
}

static inline void drift_one(const GenRow &t, float *x, const gf_mutid &id, int g,
                             uint32_t k0, uint32_t k1, U4 *cache, int *cache_blk) {
    float nx[GF_MAX_POPULATIONS];
    uint32_t tid = (uint32_t)(id.origin_threadID < 0 ? -id.origin_threadID : id.origin_threadID);
    for (int p = 0; p < t.P; p++) {
        int Nc = t.Nc[p];
        if (Nc <= 0) { nx[p] = 0.0f; continue; }
        float xs, mean, var;
        pre_draw(t, x, p, xs, mean, var);
        float N = (float)Nc;
        int blk = g >> 2;
        U4 r;
        if (cache && cache_blk[p] == blk) r = cache[p];
        else {
            U4 c = { tid, (uint32_t)id.origin_generation,
                     ((uint32_t)id.origin_population << 16) | (uint32_t)p, (uint32_t)blk };
            r = philox10(k0, k1, c);
            if (cache) { cache[p] = r; cache_blk[p] = blk; }
        }
        int j = binom_draw(lane(r, g & 3), mean, var, N);
        j = j < 0 ? 0 : (j > Nc ? Nc : j);
        nx[p] = (float)j / N;
    }
    for (int p = 0; p < t.P; p++) x[p] = nx[p];
}

/* keep rule of flag_segregating_mutations, reference go_fish_impl.cuh:137-176 */
static inline bool keep_slot(const GenRow &t, const float *x, const gf_mutid &id) {
    bool zero = true, one = true;
    for (int p = 0; p < t.P; p++) {
#if ORC_PLANT == 8
        if (true) {                                       /* planted: present mask ignored */
#else
        if (t.present[p]) {
#endif
            zero = zero && (x[p] <= 0.0f);
            one = one && (x[p] >= 1.0f);
        }
    }
    return !(zero || one) || (id.origin_threadID < 0);
}

/* ===================== simulation ===================== */
struct Sample {
    int num_mutations = 0, generation = 0;
    std::vector<int> Nchrom_e;
    std::vector<uint8_t> extinct;
    std::vector<float> freq;   /* [P][M] */
};

struct Result {
    int P = 0, G = 0;
    std::vector<Sample> samples;
    std::vector<gf_mutid> ids;
    std::vector<int64_t> live;
    double total_ms = 0;
};

static int g_threads = 1;
static std::string g_err;

/* State in slot-major layout x[i*P + p] (the oracle's own layout). */
struct Sim {
    int P;
    float L;
    uint32_t k0, k1;
    std::vector<float> x;
    std::vector<gf_mutid> id;
    std::vector<U4> cache;
    std::vector<int> cache_blk;
    std::vector<uint8_t> extinct;
    size_t M() const { return id.size(); }
};

static void parallel_for(size_t n, const std::function<void(size_t, size_t, int)> &fn);

static void fill_row(const gf_sim_desc *d, int t, int g, const std::vector<uint8_t> &extinct_now,
                     const std::vector<int> &nc_now, GenRow &row) {
    int P = d->num_populations;
    row.P = P;
    const gf_pop_gen *pg = d->params + (size_t)t * P;
    for (int p = 0; p < P; p++) {
        row.Nc[p] = nc_now[p];
        row.present[p] = pg[p].N > 0;
        row.F[p] = pg[p].F; row.h[p] = pg[p].h;
        row.sa[p] = pg[p].sel_slope; row.sb[p] = pg[p].sel_intercept;
        for (int k = 0; k < P; k++) row.mig[k][p] = d->migration[(size_t)t * P * P + (size_t)k * P + p];
    }
    (void)g; (void)extinct_now;
}

/* Deterministic two-level inclusive scan (chunks of 64, sequential float
 * sums); replaces the CUB scan of integrate_mse (go_fish_impl.cuh:215-217). */
static void scan_chunked(const std::vector<float> &a, std::vector<float> &S) {
    const int CH = 64;
    size_t n = a.size();
    S.resize(n);
    std::vector<float> tot((n + CH - 1) / CH);
    for (size_t c = 0; c * CH < n; c++) {
        float acc = 0.0f;
        for (size_t i = c * CH; i < std::min(n, (c + 1) * CH); i++) {
            acc = (i == c * CH) ? a[i] : acc + a[i];
            S[i] = acc;
        }
        tot[c] = acc;
    }
    float pre = 0.0f;
    for (size_t c = 0; c < tot.size(); c++) {
        if (c > 0) {
            for (size_t i = c * CH; i < std::min(n, (c + 1) * CH); i++) S[i] = pre + S[i];
        }
        pre = pre + tot[c];
    }
}

/* Expected mutation count of every frequency class k/Nc, k = 1..Nc-1, at
 * mutation-selection equilibrium for one population (reference
 * integrate_mse go_fish_impl.cuh:201-223, initialize_mse_frequency_array
 * :58-71, with the quirks kept: the upper integral of class id starts at
 * id/Nc, j = ((Nc - id) + 1)/Nc, the integrand's exponent uses
 * -max(s, -1)). */
static void mse_lambda(const gf_pop_gen &pg, float L, std::vector<float> &lam) {
    int Nind = pg.N;
    float F = pg.F, h = pg.h, mu = pg.mu;
    int Nc = nchrom(Nind, F);
    lam.assign(Nc > 1 ? Nc - 1 : 0, 0.0f);
    if (Nc <= 1) return;
    /* the integral is only needed when some class has s != 0 */
    bool neutral = (pg.sel_slope == 0.0f && pg.sel_intercept == 0.0f);
    std::vector<float> I;
    if (!neutral) {
        float step = 1.0f / (float)Nc;
        std::vector<float> area(Nc);
        for (int id = 0; id < Nc; id++) {
            float a = 1.0f - (float)id * step;
            float s1 = sel_eval(pg.sel_slope, pg.sel_intercept, a);
            s1 = s1 > -1.0f ? s1 : -1.0f;
            float a2 = a - step;
            float s2 = sel_eval(pg.sel_slope, pg.sel_intercept, a2);
            s2 = s2 > -1.0f ? s2 : -1.0f;
#if ORC_PLANT == 5
            float f1 = mse(a, Nind, F, h, s1), f2 = mse(a2, Nind, F, h, s2);   /* planted: integrand sign */
#else
            float f1 = mse(a, Nind, F, h, -1.0f * s1);
            float f2 = mse(a2, Nind, F, h, -1.0f * s2);
#endif
            area[id] = (step * (f1 + f2)) / 2.0f;
        }
        std::vector<float> S;
        scan_chunked(area, S);
        I.resize(Nc + 1);
        for (int id = 0; id < Nc; id++) I[id] = S[Nc - 1 - id];
        I[Nc] = 0.0f;
    }
    float total = neutral ? 0.0f : I[0];
    float twomuL = (2.0f * mu) * L;
    for (int id = 0; id < Nc - 1; id++) {
        float i = ((float)id + 1.0f) / (float)Nc;
#if ORC_PLANT == 3
        float j = 1.0f - i;                               /* planted: j without the quirk */
#else
        float j = ((float)(Nc - id) + 1.0f) / (float)Nc;
#endif
        float s = sel_eval(pg.sel_slope, pg.sel_intercept, i);
#if ORC_PLANT == 4
        const float Iu = neutral ? 0.0f : I[id + 1];      /* planted: integral from (id+1)/Nc */
#else
        const float Iu = neutral ? 0.0f : I[id];
#endif
        if (s == 0.0f) lam[id] = twomuL / i;
        else lam[id] = (twomuL * (mse(i, Nind, F, h, s) * Iu)) / ((total * i) * j);
    }
}

/* Mutation-selection-equilibrium class counts for one population: a
 * Poisson draw of each class's expected count with counter {0, id, pop, 0}
 * (reference initialize_mse_frequency_array go_fish_impl.cuh:58-71). */
static void mse_counts(const gf_pop_gen &pg, int pop, float L, uint32_t k0, uint32_t k1,
                       std::vector<int> &counts) {
    std::vector<float> lam;
    mse_lambda(pg, L, lam);
    counts.assign(lam.size(), 0);
    const float Nn = L * (float)nchrom(pg.N, pg.F);
    for (size_t id = 0; id < lam.size(); id++) {
        U4 c = { 0u, (uint32_t)id, (uint32_t)pop, 0u };
        U4 r = philox10(k0, k1, c);
        int n = pois_draw(r.x, lam[id], lam[id], Nn);
        counts[id] = n > 0 ? n : 0;
    }
}

/* new mutations of one population at generation g
 * (reference calc_new_mutations_Index go_fish_impl.cuh:454-467) */
static int new_count(int Nc, float mu, float L, uint32_t k0, uint32_t k1, int g, int pop) {
    if (Nc <= 0) return 0;
    float lambda = (mu * (float)Nc) * L;
    if (lambda == 0.0f) return 0;
    U4 c = { 0u, (uint32_t)g, (uint32_t)pop, 1u };
    U4 r = philox10(k0, k1, c);
    int n = pois_draw(r.x, lambda, lambda, (float)Nc * L);
    return n > 0 ? n : 0;
}

static void take_sample(const Sim &S, int g, const gf_sim_desc *d, int t, Result &R, bool with_ids) {
    Sample smp;
    int P = S.P;
    size_t M = S.M();
    smp.num_mutations = (int)M;
    smp.generation = g;
    smp.freq.resize((size_t)P * M);
    for (size_t i = 0; i < M; i++)
        for (int p = 0; p < P; p++) smp.freq[(size_t)p * M + i] = S.x[i * P + p];
    for (int p = 0; p < P; p++) {
        const gf_pop_gen &pg = d->params[(size_t)t * P + p];
        smp.Nchrom_e.push_back(nchrom(pg.N, pg.F));
        smp.extinct.push_back(S.extinct[p]);
    }
    R.samples.push_back(std::move(smp));
    if (with_ids) R.ids = S.id;
}

static void compact(Sim &S, const GenRow &row, bool preserve) {
    int P = S.P;
    size_t w = 0;
    for (size_t i = 0; i < S.M(); i++) {
        if (!keep_slot(row, &S.x[i * P], S.id[i])) continue;
        if (w != i) {
            for (int p = 0; p < P; p++) S.x[w * P + p] = S.x[i * P + p];
            S.id[w] = S.id[i];
        }
        if (preserve && ORC_PLANT != 9) {
            int z = S.id[w].origin_threadID;
            S.id[w].origin_threadID = -(z < 0 ? -z : z);
        }
        w++;
    }
    S.x.resize(w * P);
    S.id.resize(w);
    S.cache.assign(w * P, U4{0, 0, 0, 0});
    S.cache_blk.assign(w * P, -1);
}

static int check_params(const gf_sim_desc *d, int t, int g, const std::vector<uint8_t> &extinct) {
    int P = d->num_populations;
    const gf_pop_gen *pg = d->params + (size_t)t * P;
    const float *mig = d->migration + (size_t)t * P * P;
    for (int p = 0; p < P; p++) {
        double migration = 0;
        if (pg[p].mu < 0) return GF_ERR_PARAM;
        int N = pg[p].N;
        if (N > 0 && extinct[p]) return GF_ERR_PARAM;
        if (pg[p].F < 0 || pg[p].F > 1) return GF_ERR_PARAM;
        for (int p2 = 0; p2 < P; p2++) {
            float m_from = mig[p * P + p2];
            if (m_from < 0) return GF_ERR_PARAM;
            if (m_from > 0 && ((N <= 0 || extinct[p]) && !(pg[p2].N <= 0 || extinct[p2]))) return GF_ERR_PARAM;
            migration += (double)mig[p2 * P + p];
        }
        if ((float)migration != 1.f && !(N <= 0 || extinct[p])) return GF_ERR_PARAM;
    }
    (void)g;
    return GF_OK;
}

static int run(const gf_sim_desc *d, Result &R) {
    int P = d->num_populations;
    if (P < 1 || P > GF_MAX_POPULATIONS || !d->params || !d->migration || !d->preserve || !d->sample)
        return GF_ERR_ARG;
    Sim S;
    S.P = P; S.L = d->num_sites;
    S.k0 = (uint32_t)d->seed1; S.k1 = (uint32_t)d->seed2;
    S.extinct.assign(P, 0);
    R.P = P; R.G = d->num_generations;
    int G = d->num_generations;
    int start = d->init_mode == GF_INIT_PREV ? d->start_generation : 0;
    int final_g = start + G;
    int interval = d->compact_interval;
    R.live.assign(G + 1, 0);

    /* ----- initialisation ----- */
    if (d->init_mode == GF_INIT_MSE) {
        int rc = check_params(d, 0, 0, S.extinct);
        if (rc) return rc;
        bool pres = d->preserve[0] || d->sample[0];
        int idx = 0;
        for (int p = 0; p < P; p++) {
            const gf_pop_gen &pg = d->params[p];
            std::vector<int> counts;
            mse_counts(pg, p, S.L, S.k0, S.k1, counts);
            int Nc = nchrom(pg.N, pg.F);
            for (size_t k = 0; k < counts.size(); k++) {
                float f = ((float)k + 1.0f) / (float)Nc;
                for (int c = 0; c < counts[k]; c++) {
                    for (int q = 0; q < P; q++) S.x.push_back(q == p ? f : 0.0f);
                    int tid = idx + 1;
                    S.id.push_back(gf_mutid{0, p, pres ? -tid : tid, 0});
                    idx++;
                }
            }
        }
    } else if (d->init_mode == GF_INIT_PREV) {
        int n = d->init_num_mutations;
        if (n > 0) {
            if (!d->init_freq || !d->init_ids) return GF_ERR_ARG;
            for (int p = 0; p < P; p++) S.extinct[p] = d->init_extinct ? d->init_extinct[p] : 0;
            bool pres = d->preserve[0] || d->sample[0];
            S.x.resize((size_t)n * P);
            for (int i = 0; i < n; i++) {
                for (int p = 0; p < P; p++) S.x[(size_t)i * P + p] = d->init_freq[(size_t)p * n + i];
                gf_mutid m = d->init_ids[i];
                if (pres) { int z = m.origin_threadID; m.origin_threadID = -(z < 0 ? -z : z); }
                S.id.push_back(m);
            }
        }
        int rc = check_params(d, 0, start, S.extinct);
        if (rc) return rc;
    } else {
        int rc = check_params(d, 0, 0, S.extinct);
        if (rc) return rc;
    }
    S.cache.assign(S.M() * P, U4{0, 0, 0, 0});
    S.cache_blk.assign(S.M() * P, -1);

    /* live count of the initial state */
    {
        GenRow row;
        std::vector<int> nc(P);
        for (int p = 0; p < P; p++) nc[p] = nchrom(d->params[p].N, d->params[p].F);
        fill_row(d, 0, start, S.extinct, nc, row);
        int64_t c = 0;
        for (size_t i = 0; i < S.M(); i++) c += keep_slot(row, &S.x[i * P], S.id[i]);
        R.live[0] = c;
    }

    if (d->sample[0] && start != final_g) take_sample(S, start, d, 0, R, false);

    int next_compact = start + interval;
    std::vector<int> prevN(P);
    for (int p = 0; p < P; p++) prevN[p] = d->params[p].N;

    for (int g = start + 1; g <= final_g; g++) {
        int t = g - start;
        int rc = check_params(d, t, g, S.extinct);
        if (rc) return rc;
        const gf_pop_gen *pg = d->params + (size_t)t * P;
        std::vector<int> nc(P, 0);
        for (int p = 0; p < P; p++) {
            int N = pg[p].N;
            if (S.extinct[p]) { nc[p] = 0; continue; }
            if (N <= 0) {
                if (prevN[p] > 0) S.extinct[p] = 1;
                nc[p] = 0;
                continue;
            }
            nc[p] = nchrom(N, pg[p].F);
        }
        GenRow row;
        fill_row(d, t, g, S.extinct, nc, row);
        size_t M0 = S.M();
        /* migration, selection, drift for every existing mutation */
        std::vector<int64_t> part((size_t)std::max(1, g_threads), 0);
        parallel_for(M0, [&](size_t b, size_t e, int th) {
            int64_t c = 0;
            for (size_t i = b; i < e; i++) {
                drift_one(row, &S.x[i * P], S.id[i], g, S.k0, S.k1, &S.cache[i * P], &S.cache_blk[i * P]);
                c += keep_slot(row, &S.x[i * P], S.id[i]);
            }
            part[th] += c;
        });
        int64_t live = 0;
        for (auto v : part) live += v;
        /* new mutations, appended in population order */
        for (int p = 0; p < P; p++) {
            int Ncn = nc[p];
            int n = new_count(Ncn, pg[p].mu, S.L, S.k0, S.k1, g, p);
            if (n <= 0) continue;
            float f = 1.0f / (float)Ncn;
            for (int k = 0; k < n; k++) {
                for (int q = 0; q < P; q++) S.x.push_back(q == p ? f : 0.0f);
                S.id.push_back(gf_mutid{g, p, k + 1, 0});
                S.cache.push_back(U4{0, 0, 0, 0});
                for (int q = 1; q < P; q++) S.cache.push_back(U4{0, 0, 0, 0});
                for (int q = 0; q < P; q++) S.cache_blk.push_back(-1);
                live += keep_slot(row, &S.x[(S.M() - 1) * P], S.id.back());
            }
        }
        R.live[t] = live;
        for (int p = 0; p < P; p++) prevN[p] = pg[p].N;

        bool preserve = d->preserve[t] || d->sample[t];
        if ((g == next_compact || g == final_g || preserve) && interval > 0) {
            compact(S, row, preserve);
            next_compact = g + interval;
        }
        if (d->sample[t] && g != final_g) take_sample(S, g, d, t, R, false);
    }
    take_sample(S, final_g, d, G, R, true);
    return GF_OK;
}

static void parallel_for(size_t n, const std::function<void(size_t, size_t, int)> &fn) {
    int T = std::max(1, g_threads);
    if (T == 1 || n < 4096) { fn(0, n, 0); return; }
    std::vector<std::thread> th;
    size_t chunk = (n + T - 1) / T;
    for (int t = 0; t < T; t++) {
        size_t b = (size_t)t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=, &fn]() { fn(b, e, t); });
    }
    for (auto &x : th) x.join();
}

/* ===================== Spectrum ===================== */

/* reference population_hist, spectrum.cu:74-83 */
static void pop_hist(const float *freq, int M, int Nc, float num_sites, std::vector<uint32_t> &h) {
    h.assign(Nc, 0);
    for (int i = 0; i < M; i++) {
#if ORC_PLANT == 6
        int idx = f2i_sat(rintf((float)Nc * freq[i]));    /* planted: ties to even */
#else
        int idx = f2i_sat(round_away((float)Nc * freq[i]));
#endif
        if (idx == Nc && ORC_PLANT != 7) idx = 0;        /* ORC_PLANT 7: bin Nc not folded */
        if (idx >= 0 && idx < Nc) h[idx]++;
    }
#if ORC_PLANT == 13
    int ns = f2i_sat(round_away(num_sites));              /* planted: num_sites rounded */
#else
    int ns = f2i_sat(truncf(num_sites));
#endif
    h[0] += (uint32_t)(ns - M);
}

/* log C(n,k) table for k < half_n (reference binom_coeff spectrum.cu:90-95
 * plus inclusive scan :258-260), scanned sequentially in float. */
static void log_binom_table(int n, std::vector<float> &lc, int &half_n) {
    half_n = (n % 2 == 0) ? n / 2 + 1 : (n + 1) / 2;
    lc.assign(half_n, 0.0f);
    float acc = 0.0f;
    for (int k = 1; k < half_n; k++) {
        acc = acc + log_f(((float)n + 1.0f - (float)k) / (float)k);
        lc[k] = acc;
    }
}

/* reference binom<>, spectrum.cu:111-139, with 2^-32 fixed-point sums */
static void sfs(const float *freq, int M, int n, float num_sites, std::vector<float> &out) {
    std::vector<float> lc; int half_n;
    log_binom_table(n, lc, half_n);
    std::vector<uint64_t> acc(n + 1, 0);
    uint64_t nonseg = 0;
    for (int i = 0; i < M; i++) {
        float pf = freq[i];
        if (pf > 0.0f && pf < 1.0f) {
            float lp = log_f(pf), lq = log_f(1.0f - pf);
            for (int k = 0; k <= n; k++) {
                float coeff = k < half_n ? lc[k] : lc[n - k];
                float t = (coeff + (float)k * lp) + (float)(n - k) * lq;
                if (t < -23.0f) continue;
                float v = exp_f(t);
                acc[k] += (uint64_t)(v * 4294967296.0f);
            }
        } else nonseg++;
    }
    out.assign(n, 0.0f);
    const double sc = 1.0 / 4294967296.0;
    double b0 = (double)(acc[0] + acc[n]) * sc + (double)nonseg + ((double)num_sites - (double)M);
    out[0] = (float)b0;
    for (int k = 1; k < n; k++) out[k] = (float)((double)acc[k] * sc);
}

/* per-population sample probabilities of one mutation, the float sequence
 * of sfs() above: P(k | n, p) = exp(log C(n,k) + k log p + (n-k) log q),
 * 0 where the exponent is below -23; p <= 0 or p >= 1: all mass at k = 0
 * or k = n */
static void sample_pmf(float pf, int n, const std::vector<float> &lc, int half_n, float *pmf) {
    for (int k = 0; k <= n; k++) pmf[k] = 0.0f;
    if (pf <= 0.0f) { pmf[0] = 1.0f; return; }
    if (pf >= 1.0f) { pmf[n] = 1.0f; return; }
    float lp = log_f(pf), lq = log_f(1.0f - pf);
    for (int k = 0; k <= n; k++) {
        float coeff = k < half_n ? lc[k] : lc[n - k];
        float t = (coeff + (float)k * lp) + (float)(n - k) * lq;
        if (t < -23.0f) continue;
        pmf[k] = exp_f(t);
    }
}

/* joint sample SFS of two populations (no reference counterpart; the
 * reference's spectrum is per population, spectrum.h:23):
 * out[ka*(nb+1)+kb] = sum over mutations of Pa(ka) Pb(kb), each product
 * truncated to 2^-32 fixed point and summed exactly; monomorphic sites not
 * in the array (num_sites - M) go to [0][0] */
static void joint_sfs(const float *fa, const float *fb, int M, int na, int nb, float num_sites, std::vector<float> &out) {
    std::vector<float> la, lb; int ha, hb;
    log_binom_table(na, la, ha);
    log_binom_table(nb, lb, hb);
    const int B = (na + 1) * (nb + 1);
    std::vector<uint64_t> acc(B, 0);
    std::vector<float> pa(na + 1), pb(nb + 1);
    for (int i = 0; i < M; i++) {
        sample_pmf(fa[i], na, la, ha, pa.data());
        sample_pmf(fb[i], nb, lb, hb, pb.data());
        for (int a = 0; a <= na; a++) {
            if (pa[a] == 0.0f) continue;
            for (int b = 0; b <= nb; b++) acc[(size_t)a * (nb + 1) + b] += (uint64_t)((pa[a] * pb[b]) * 4294967296.0f);
        }
    }
    out.assign(B, 0.0f);
    const double sc = 1.0 / 4294967296.0;
    for (int k = 0; k < B; k++) out[k] = (float)((double)acc[k] * sc + (k == 0 ? (double)num_sites - (double)M : 0.0));
}

static void fill_test_row(GenRow &row, int32_t P, const int32_t *Nc, const float *F, const float *h, const float *sa,
                          const float *sb, const float *mig) {
    row.P = P;
    for (int p = 0; p < P; p++) {
        row.Nc[p] = Nc[p]; row.present[p] = Nc[p] > 0; row.F[p] = F[p]; row.h[p] = h[p];
        row.sa[p] = sa[p]; row.sb[p] = sb[p];
        for (int k = 0; k < P; k++) row.mig[k][p] = mig[k * P + p];
    }
}

} /* namespace orc */

/* ===================== C API (test hooks) ===================== */
using namespace orc;

struct orc_result { Result r; };

extern "C" {

int32_t orc_set_threads(int32_t n) { g_threads = n < 1 ? 1 : n; return 0; }

void orc_philox(uint32_t k0, uint32_t k1, const uint32_t *ctr, uint32_t *out) {
    U4 r = philox10(k0, k1, U4{ctr[0], ctr[1], ctr[2], ctr[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
void orc_u01_array(const uint32_t *in, float *out, int64_t n) { for (int64_t i = 0; i < n; i++) out[i] = u01(in[i]); }
void orc_exp_array(const float *in, float *out, int64_t n) { for (int64_t i = 0; i < n; i++) out[i] = exp_f(in[i]); }
void orc_log_array(const float *in, float *out, int64_t n) { for (int64_t i = 0; i < n; i++) out[i] = log_f(in[i]); }
void orc_normcdfinv_array(const float *in, float *out, int64_t n) { for (int64_t i = 0; i < n; i++) out[i] = normcdfinv_f(in[i]); }
void orc_expneg_array(const float *in, float *out, int64_t n) { for (int64_t i = 0; i < n; i++) out[i] = expneg_tab(in[i]); }
void orc_poiscdfinv_array(const float *r, const float *mean, int32_t *out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = poiscdfinv(r[i], mean[i]);
}
void orc_binomcdfinv_array(const float *r, const float *mean, const float *x, const float *N, int32_t *out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = binomcdfinv(r[i], mean[i], x[i], N[i]);
}
void orc_binom_draw_array(const uint32_t *bits, const float *mean, const float *var, const float *N, int32_t *out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = binom_draw(bits[i], mean[i], var[i], N[i]);
}
void orc_pois_draw_array(const uint32_t *bits, const float *mean, const float *N, int32_t *out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = pois_draw(bits[i], mean[i], mean[i], N[i]);
}
void orc_round_away_array(const float *in, float *out, int64_t n) { for (int64_t i = 0; i < n; i++) out[i] = round_away(in[i]); }
void orc_f2i_array(const float *in, int32_t *out, int64_t n) { for (int64_t i = 0; i < n; i++) out[i] = f2i_sat(in[i]); }

/* one generation of drift for a set of mutations (P populations) */
int32_t orc_drift_step(int32_t P, int32_t M, float *x /*[M][P]*/, const gf_mutid *ids, int32_t g,
                       int32_t seed1, int32_t seed2, const int32_t *Nc, const float *F, const float *h,
                       const float *sa, const float *sb, const float *mig /*[P][P] from,to*/) {
    if (P < 1 || P > GF_MAX_POPULATIONS) return GF_ERR_ARG;
    GenRow row;
    fill_test_row(row, P, Nc, F, h, sa, sb, mig);
    for (int i = 0; i < M; i++) drift_one(row, &x[(size_t)i * P], ids[i], g, (uint32_t)seed1, (uint32_t)seed2, nullptr, nullptr);
    return 0;
}

/* migration + selection of M mutations x[M][P] (no draw): the frequency
 * after selection and the mean and variance of the drift draw, [M][P] each;
 * populations with Nc <= 0 get 0 */
int32_t orc_pre_draw(int32_t P, int32_t M, const float *x, const int32_t *Nc, const float *F, const float *h,
                     const float *sa, const float *sb, const float *mig, float *xs, float *mean, float *var) {
    if (P < 1 || P > GF_MAX_POPULATIONS || M < 0) return GF_ERR_ARG;
    GenRow row;
    fill_test_row(row, P, Nc, F, h, sa, sb, mig);
    for (int i = 0; i < M; i++)
        for (int p = 0; p < P; p++) {
            size_t o = (size_t)i * P + p;
            xs[o] = mean[o] = var[o] = 0.0f;
            if (Nc[p] > 0) pre_draw(row, &x[(size_t)i * P], p, xs[o], mean[o], var[o]);
        }
    return 0;
}

/* expected equilibrium count of each frequency class (Nc - 1 of them) */
int32_t orc_mse_lambda(const gf_pop_gen *pg, float L, float *out, int32_t cap) {
    std::vector<float> lam;
    mse_lambda(*pg, L, lam);
    for (size_t i = 0; i < lam.size() && (int)i < cap; i++) out[i] = lam[i];
    return (int32_t)lam.size();
}

/* one compaction of M slots x[M][P] / ids (in place): keep rule over the
 * populations with present[p] != 0, stable order, preserve negation;
 * returns the number kept */
int32_t orc_compact_step(int32_t P, int32_t M, float *x, gf_mutid *ids, const int32_t *present, int32_t preserve) {
    if (P < 1 || P > GF_MAX_POPULATIONS || M < 0) return GF_ERR_ARG;
    Sim S;
    S.P = P;
    S.x.assign(x, x + (size_t)M * P);
    S.id.assign(ids, ids + M);
    GenRow row;
    row.P = P;
    for (int p = 0; p < P; p++) row.present[p] = present[p] != 0;
    compact(S, row, preserve != 0);
    memcpy(x, S.x.data(), S.x.size() * sizeof(float));
    memcpy(ids, S.id.data(), S.id.size() * sizeof(gf_mutid));
    return (int32_t)S.M();
}

int32_t orc_mse_counts(const gf_pop_gen *pg, int32_t pop, float L, int32_t seed1, int32_t seed2, int32_t *out, int32_t cap) {
    std::vector<int> c;
    mse_counts(*pg, pop, L, (uint32_t)seed1, (uint32_t)seed2, c);
    for (size_t i = 0; i < c.size() && (int)i < cap; i++) out[i] = c[i];
    return (int32_t)c.size();
}

int32_t orc_new_count(int32_t Nc, float mu, float L, int32_t seed1, int32_t seed2, int32_t g, int32_t pop) {
    return new_count(Nc, mu, L, (uint32_t)seed1, (uint32_t)seed2, g, pop);
}

int32_t orc_run(const gf_sim_desc *d, orc_result **out) {
    auto *r = new orc_result;
    int rc = run(d, r->r);
    if (rc) { delete r; *out = nullptr; return rc; }
    *out = r;
    return GF_OK;
}
int32_t orc_result_num_samples(const orc_result *r) { return (int32_t)r->r.samples.size(); }
int32_t orc_result_sample_info(const orc_result *r, int32_t s, int32_t *M, int32_t *gen, int32_t *Nc, uint8_t *ext) {
    if (s < 0 || s >= (int)r->r.samples.size()) return GF_ERR_ARG;
    const Sample &x = r->r.samples[s];
    if (M) *M = x.num_mutations;
    if (gen) *gen = x.generation;
    for (int p = 0; p < r->r.P; p++) { if (Nc) Nc[p] = x.Nchrom_e[p]; if (ext) ext[p] = x.extinct[p]; }
    return 0;
}
int32_t orc_result_sample_freq(const orc_result *r, int32_t s, float *out) {
    if (s < 0 || s >= (int)r->r.samples.size()) return GF_ERR_ARG;
    const Sample &x = r->r.samples[s];
    memcpy(out, x.freq.data(), x.freq.size() * sizeof(float));
    return 0;
}
int32_t orc_result_num_ids(const orc_result *r) { return (int32_t)r->r.ids.size(); }
int32_t orc_result_ids(const orc_result *r, gf_mutid *out) {
    memcpy(out, r->r.ids.data(), r->r.ids.size() * sizeof(gf_mutid));
    return 0;
}
int32_t orc_result_live_counts(const orc_result *r, int64_t *out) {
    memcpy(out, r->r.live.data(), r->r.live.size() * sizeof(int64_t));
    return 0;
}
void orc_result_free(orc_result *r) { delete r; }

int32_t orc_population_histogram(const float *freq, int32_t M, int32_t Nc, float num_sites, float *out) {
    if (Nc <= 0 || M < 0) return GF_ERR_ARG;
    std::vector<uint32_t> h;
    pop_hist(freq, M, Nc, num_sites, h);
    for (int i = 0; i < Nc; i++) out[i] = (float)h[i];
    return 0;
}
int32_t orc_joint_site_frequency_spectrum(const float *fa, const float *fb, int32_t M, int32_t na, int32_t nb,
                                          float num_sites, float *out) {
    if (na <= 0 || nb <= 0 || M < 0) return GF_ERR_ARG;
    std::vector<float> o;
    joint_sfs(fa, fb, M, na, nb, num_sites, o);
    memcpy(out, o.data(), o.size() * sizeof(float));
    return 0;
}
int32_t orc_site_frequency_spectrum(const float *freq, int32_t M, int32_t n, float num_sites, float *out) {
    if (n <= 0 || M < 0) return GF_ERR_ARG;
    std::vector<float> o;
    sfs(freq, M, n, num_sites, o);
    memcpy(out, o.data(), n * sizeof(float));
    return 0;
}

} /* extern "C" */
