/*
 * gf_spectrum.hpp -- site-frequency-spectrum kernels.
 *
 * k_pop_hist: integer histogram of round(Nc * x) with bin Nc folded into 0
 *   (reference population_hist, spectrum.cu:74-83), LDS-privatised in passes
 *   over bin ranges of up to 40,960 bins.
 * k_sfs: expected SFS of an n-chromosome sample, sum over segregating
 *   mutations of exp(log C(n,k) + k log p + (n-k) log q) for k = 0..n
 *   (reference binom, spectrum.cu:111-139).  Each term is truncated to 2^-32
 *   fixed point and summed in 64-bit integers (LDS, then global), so the
 *   result is independent of summation order.  Terms with exponent < -23
 *   truncate to 0 and are skipped; each lane visits only the levels around
 *   its mode where they can occur.
 */
#ifndef GF_SPECTRUM_HPP_
#define GF_SPECTRUM_HPP_

#include "gf_kernels.hpp"
#include "gf_runtime.hpp"
#include <vector>

namespace gf {

constexpr int SFS_LDS_BINS = 4096;

/* Population histogram in LDS passes over bin ranges: workgroup (x, y)
 * counts the mutations of strip x (a grid-stride share of the row) whose bin
 * lies in range y, [y * chunk, y * chunk + chunk), in its own LDS copy of
 * that range (dynamic LDS, chunk <= 40,960 bins = 160 KB), then adds its
 * non-zero bins to the global histogram.  Every row is read once per bin
 * range (Nc = 4e5 takes ten), and global atomics touch only non-zero
 * per-strip bins. */
constexpr int HIST_WG = 1024;
constexpr int HIST_MAX_CHUNK = 40960;

static __global__ __launch_bounds__(HIST_WG) void k_pop_hist(const float *__restrict__ freq, int M, int Nc, int chunk,
                                                      unsigned int *__restrict__ hist) {
    extern __shared__ unsigned int sh_hist[];
    const int b0 = (int)blockIdx.y * chunk;
    const int nb = Nc - b0 < chunk ? Nc - b0 : chunk;
    for (int k = threadIdx.x; k < nb; k += HIST_WG) sh_hist[k] = 0;
    __syncthreads();
    /* 64-bit index: i + stride must not wrap for M near INT_MAX */
    for (int64_t i = (int64_t)blockIdx.x * HIST_WG + threadIdx.x; i < M; i += (int64_t)gridDim.x * HIST_WG) {
        int idx = f2i_sat(round_away((float)Nc * freq[i]));
        if (idx == Nc) idx = 0;
        const unsigned int r = (unsigned int)(idx - b0);
        if (idx >= 0 && r < (unsigned int)nb) atomicAdd(&sh_hist[r], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nb; k += HIST_WG)
        if (sh_hist[k]) atomicAdd(&hist[b0 + k], sh_hist[k]);
}

/* Adds one mutation of frequency pf to the fixed-point level sums of a
 * sample of n (sh in LDS when LDS, else out in global memory): for
 * 0 < pf < 1 the terms exp(log C(n,k) + k log p + (n-k) log q), each
 * truncated to 2^-32 fixed point; a frequency of 0 or 1 counts one
 * non-segregating mutation at [n + 1]. */
template <bool LDS>
__device__ __forceinline__ void sfs_add_mutation(float pf, int n, int half_n, const float *__restrict__ lc,
                                                 unsigned long long *sh, unsigned long long *out) {
    if (pf > 0.0f && pf < 1.0f) {
        const float lp = log_f(pf);
        const float lq = log_f(1.0f - pf);
        /* The terms with t >= -23 (the only ones that survive the 2^-32
         * truncation) form one run of levels around the mode: t is a
         * concave function of k (log C(n,k) is concave, the rest linear)
         * evaluated with rounding errors far below 1, so walking out from
         * the mode until t < -24 on each side visits every one of them, and
         * adds exactly the terms a loop over all n + 1 levels would. */
        auto level = [&](int k) {
            const float coeff = k < half_n ? lc[k] : lc[n - k];
            return (coeff + (float)k * lp) + (float)(n - k) * lq;
        };
        auto add = [&](int k, float t) {
            if (t < -23.0f) return;
            const unsigned long long v = (unsigned long long)(exp_f(t) * 4294967296.0f);
            if (v) {
                if (LDS) atomicAdd(&sh[k], v);
                else atomicAdd(&out[k], v);
            }
        };
        int km = (int)((float)(n + 1) * pf);
        km = km > n ? n : km;
        for (int k = km; k <= n; k++) {
            const float t = level(k);
            if (t < -24.0f) break;
            add(k, t);
        }
        for (int k = km - 1; k >= 0; k--) {
            const float t = level(k);
            if (t < -24.0f) break;
            add(k, t);
        }
    } else {
        if (LDS) atomicAdd(&sh[n + 1], 1ull);
        else atomicAdd(&out[n + 1], 1ull);
    }
}

/* acc layout per segment: [0..n] fixed-point level sums, [n+1] count of
 * mutations that are not segregating (p <= 0 or p >= 1) */
template <bool LDS>
__global__ __launch_bounds__(WG) void k_sfs(const float *__restrict__ freq, const int64_t *__restrict__ seg_base,
                                            const int *__restrict__ seg_M, const int *__restrict__ tile_seg,
                                            const int *__restrict__ tile_first, int n, int half_n,
                                            const float *__restrict__ lc, unsigned long long *__restrict__ acc) {
    __shared__ unsigned long long sh[LDS ? SFS_LDS_BINS : 1];
    const int tile = blockIdx.x;
    const int s = tile_seg[tile];
    const int li = (tile - tile_first[s]) * WG + (int)threadIdx.x;
    unsigned long long *out = acc + (size_t)s * (n + 2);
    if (LDS) {
        for (int k = threadIdx.x; k < n + 2; k += WG) sh[k] = 0;
        __syncthreads();
    }
    if (li < seg_M[s]) sfs_add_mutation<LDS>(freq[seg_base[s] + li], n, half_n, lc, sh, out);
    if (LDS) {
        __syncthreads();
        for (int k = threadIdx.x; k < n + 2; k += WG)
            if (sh[k]) atomicAdd(&out[k], sh[k]);
    }
}

/* The sample SFS of one device row with every step on the device (no host
 * copy, no allocation): k_sfs_prep writes the log-binomial table lc[0 ..
 * half_n) with the host table's sequential float sum (log_binom_table) and
 * clears acc[0 .. n + 1]; k_sfs_row accumulates, a grid-stride loop over
 * the M mutations; k_sfs_finish converts to the float spectrum as
 * finish_sfs does. */
__global__ void k_sfs_prep(int n, int half_n, float *__restrict__ lc, unsigned long long *__restrict__ acc);
template <bool LDS>
__global__ __launch_bounds__(WG) void k_sfs_row(const float *__restrict__ freq, int M, int n, int half_n,
                                                const float *__restrict__ lc, unsigned long long *__restrict__ acc) {
    __shared__ unsigned long long sh[LDS ? SFS_LDS_BINS : 1];
    if (LDS) {
        for (int k = threadIdx.x; k < n + 2; k += WG) sh[k] = 0;
        __syncthreads();
    }
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < M; i += (int64_t)gridDim.x * WG)
        sfs_add_mutation<LDS>(freq[i], n, half_n, lc, sh, acc);
    if (LDS) {
        __syncthreads();
        for (int k = threadIdx.x; k < n + 2; k += WG)
            if (sh[k]) atomicAdd(&acc[k], sh[k]);
    }
}
__global__ void k_sfs_finish(const unsigned long long *__restrict__ acc, int n, int M, float num_sites,
                             float *__restrict__ out);

void log_binom_table(int n, std::vector<float> &lc, int &half_n);
void finish_sfs(const unsigned long long *acc, unsigned long long nonseg, int n, int M, float num_sites, float *out);
/* sample SFS of one population row already in device memory (M floats) */
/* device buffers of the sample SFS of one row, kept between calls by their
 * owner (grow-only) */
struct SfsScratch {
    DevBuf<float> lc;
    DevBuf<unsigned long long> acc;
    DevBuf<int> meta;
    DevBuf<int64_t> base;
};
int sfs_of_device_row(const float *d_row, int M, int n, float num_sites, float *out, hipStream_t st,
                      SfsScratch *scratch = nullptr);
int sfs_device(const float *d_freq, const int64_t *d_seg_base, const int *d_seg_M, const int *d_tile_seg,
               const int *d_tile_first, int num_tiles, int nseg, int n, unsigned long long *d_acc,
               hipStream_t st, DevBuf<float> &d_lc);

/* ---------------- joint sample SFS of two populations ----------------
 * (no reference counterpart: the reference's spectrum is per population,
 * spectrum.h:23).  out[ka (nb+1) + kb] = sum over mutations of
 * Pa(ka) Pb(kb), with the per-population sample probabilities of k_sfs
 * (mass at k = 0 or n for a frequency of 0 or 1); each product is
 * truncated to 2^-32 fixed point and summed exactly, so the result does
 * not depend on order.  Output-stationary: a workgroup owns
 * JS_PER_THREAD * 256 bins (grid y) and a range of JS_MUT_PER_WG mutations
 * (grid x), streams them through LDS in batches with their probability
 * vectors, and adds its sums to the global bins once. */
constexpr int JS_PER_THREAD = 12;
constexpr int JS_MUT_PER_WG = 2048;
constexpr int JS_LDS_FLOATS = 16384;

__device__ __forceinline__ float sample_prob(float pf, int n, int k, const float *__restrict__ lc, int half_n) {
    GF_STRICT_FP
    if (pf <= 0.0f) return k == 0 ? 1.0f : 0.0f;
    if (pf >= 1.0f) return k == n ? 1.0f : 0.0f;
    const float lp = log_f(pf), lq = log_f(1.0f - pf);
    const float coeff = k < half_n ? lc[k] : lc[n - k];
    const float t = (coeff + (float)k * lp) + (float)(n - k) * lq;
    return t < -23.0f ? 0.0f : exp_f(t);
}

static __global__ __launch_bounds__(WG) void k_joint_sfs(const float *__restrict__ fa, const float *__restrict__ fb, int M,
                                                  int na, int nb, int half_a, int half_b,
                                                  const float *__restrict__ la, const float *__restrict__ lb,
                                                  int batch, unsigned long long *__restrict__ acc) {
    GF_STRICT_FP
    extern __shared__ float sh[];
    float *sa = sh;                          /* [batch][na + 1] */
    float *sb = sh + (size_t)batch * (na + 1);   /* [batch][nb + 1] */
    const int B = (na + 1) * (nb + 1);
    const int bin0 = blockIdx.y * WG * JS_PER_THREAD;
    unsigned long long a[JS_PER_THREAD];
    int ka[JS_PER_THREAD], kb[JS_PER_THREAD];
#pragma unroll
    for (int j = 0; j < JS_PER_THREAD; j++) {
        a[j] = 0;
        const int bin = bin0 + (int)threadIdx.x + WG * j;
        ka[j] = bin < B ? bin / (nb + 1) : -1;
        kb[j] = bin < B ? bin % (nb + 1) : 0;
    }
    const int m_begin = blockIdx.x * JS_MUT_PER_WG;
    const int m_end = m_begin + JS_MUT_PER_WG < M ? m_begin + JS_MUT_PER_WG : M;
    for (int mb = m_begin; mb < m_end; mb += batch) {
        const int nm = m_end - mb < batch ? m_end - mb : batch;
        __syncthreads();
        for (int idx = threadIdx.x; idx < nm * (na + 1); idx += WG) {
            const int m = idx / (na + 1), k = idx % (na + 1);
            sa[idx] = sample_prob(fa[mb + m], na, k, la, half_a);
        }
        for (int idx = threadIdx.x; idx < nm * (nb + 1); idx += WG) {
            const int m = idx / (nb + 1), k = idx % (nb + 1);
            sb[idx] = sample_prob(fb[mb + m], nb, k, lb, half_b);
        }
        __syncthreads();
        for (int m = 0; m < nm; m++) {
            const float *pa = sa + (size_t)m * (na + 1);
            const float *pb = sb + (size_t)m * (nb + 1);
#pragma unroll
            for (int j = 0; j < JS_PER_THREAD; j++) {
                if (ka[j] < 0) continue;
                const float x = pa[ka[j]];
                if (x != 0.0f) a[j] += (unsigned long long)((x * pb[kb[j]]) * 4294967296.0f);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < JS_PER_THREAD; j++)
        if (ka[j] >= 0 && a[j]) atomicAdd(&acc[bin0 + (int)threadIdx.x + WG * j], a[j]);
}

/* joint sample SFS of two population rows already in device memory */
int joint_sfs_of_device_rows(const float *d_fa, const float *d_fb, int M, int na, int nb, float num_sites, float *out,
                             hipStream_t st);

} /* namespace gf */

#endif /* GF_SPECTRUM_HPP_ */
