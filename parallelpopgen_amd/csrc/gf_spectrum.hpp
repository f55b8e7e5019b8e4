/*
 * gf_spectrum.hpp -- site-frequency-spectrum kernels.
 *
 * k_pop_hist: integer histogram of round(Nc * x) with bin Nc folded into 0
 *   (reference population_hist, spectrum.cu:74-83); LDS-privatised when the
 *   histogram fits (Nc <= 16384), global 32-bit atomics otherwise.
 * k_sfs: expected SFS of an n-chromosome sample, sum over segregating
 *   mutations of exp(log C(n,k) + k log p + (n-k) log q) for k = 0..n
 *   (reference binom, spectrum.cu:111-139).  Each term is truncated to 2^-32
 *   fixed point and summed in 64-bit integers (LDS, then global), so the
 *   result is independent of summation order.  Terms with exponent < -23
 *   truncate to 0 and are skipped.
 */
#ifndef GF_SPECTRUM_HPP_
#define GF_SPECTRUM_HPP_

#include "gf_kernels.hpp"
#include "gf_runtime.hpp"
#include <vector>

namespace gf {

constexpr int HIST_LDS_BINS = 16384;
constexpr int SFS_LDS_BINS = 4096;

template <bool LDS>
__global__ __launch_bounds__(WG) void k_pop_hist(const float *__restrict__ freq, int M, int Nc,
                                                 unsigned int *__restrict__ hist) {
    __shared__ unsigned int sh[LDS ? HIST_LDS_BINS : 1];
    if (LDS) {
        for (int k = threadIdx.x; k < Nc; k += WG) sh[k] = 0;
        __syncthreads();
    }
    for (int i = blockIdx.x * WG + threadIdx.x; i < M; i += gridDim.x * WG) {
        int idx = f2i_sat(round_away((float)Nc * freq[i]));
        if (idx == Nc) idx = 0;
        if (idx >= 0 && idx < Nc) {
            if (LDS) atomicAdd(&sh[idx], 1u);
            else atomicAdd(&hist[idx], 1u);
        }
    }
    if (LDS) {
        __syncthreads();
        for (int k = threadIdx.x; k < Nc; k += WG)
            if (sh[k]) atomicAdd(&hist[k], sh[k]);
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
    if (li < seg_M[s]) {
        const float pf = freq[seg_base[s] + li];
        if (pf > 0.0f && pf < 1.0f) {
            const float lp = log_f(pf);
            const float lq = log_f(1.0f - pf);
            for (int k = 0; k <= n; k++) {
                const float coeff = k < half_n ? lc[k] : lc[n - k];
                const float t = (coeff + (float)k * lp) + (float)(n - k) * lq;
                if (t < -23.0f) continue;
                const unsigned long long v = (unsigned long long)(exp_f(t) * 4294967296.0f);
                if (v) {
                    if (LDS) atomicAdd(&sh[k], v);
                    else atomicAdd(&out[k], v);
                }
            }
        } else {
            if (LDS) atomicAdd(&sh[n + 1], 1ull);
            else atomicAdd(&out[n + 1], 1ull);
        }
    }
    if (LDS) {
        __syncthreads();
        for (int k = threadIdx.x; k < n + 2; k += WG)
            if (sh[k]) atomicAdd(&out[k], sh[k]);
    }
}

void log_binom_table(int n, std::vector<float> &lc, int &half_n);
void finish_sfs(const unsigned long long *acc, unsigned long long nonseg, int n, int M, float num_sites, float *out);
int sfs_device(const float *d_freq, const int64_t *d_seg_base, const int *d_seg_M, const int *d_tile_seg,
               const int *d_tile_first, int num_tiles, int nseg, int n, unsigned long long *d_acc,
               hipStream_t st, DevBuf<float> &d_lc);

} /* namespace gf */

#endif /* GF_SPECTRUM_HPP_ */
