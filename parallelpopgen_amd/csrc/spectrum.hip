/*
 * spectrum.hip -- Spectrum API kernels and C ABI:
 *   gf_population_histogram     (reference spectrum.cu:153-210)
 *   gf_site_frequency_spectrum  (reference spectrum.cu:213-294)
 *
 * Both results are integer (or 2^-32 fixed-point) sums, so they do not depend
 * on thread or workgroup order: the reference's float atomics
 * (spectrum.cu:134-138) are replaced by 64-bit integer LDS/global atomics.
 */
#include "gf_spectrum.hpp"
#include "gf_runtime.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace gf {

void log_binom_table(int n, std::vector<float> &lc, int &half_n) {
    half_n = (n % 2 == 0) ? n / 2 + 1 : (n + 1) / 2;
    lc.assign(half_n, 0.0f);
    float acc = 0.0f;
    for (int k = 1; k < half_n; k++) {
        acc = acc + log_f(((float)n + 1.0f - (float)k) / (float)k);
        lc[k] = acc;
    }
}

void finish_sfs(const unsigned long long *acc, unsigned long long nonseg, int n, int M, float num_sites, float *out) {
    const double sc = 1.0 / 4294967296.0;
    const double b0 = (double)(acc[0] + acc[n]) * sc + (double)nonseg + ((double)num_sites - (double)M);
    out[0] = (float)b0;
    for (int k = 1; k < n; k++) out[k] = (float)((double)acc[k] * sc);
}

int sfs_device(const float *d_freq, const int64_t *d_seg_base, const int *d_seg_M, const int *d_tile_seg,
               const int *d_tile_first, int num_tiles, int nseg, int n, unsigned long long *d_acc,
               hipStream_t st, DevBuf<float> &d_lc) {
    std::vector<float> lc;
    int half_n;
    log_binom_table(n, lc, half_n);
    GF_TRY(d_lc.alloc(lc.size()));
    GF_HIP(hipMemcpyAsync(d_lc.p, lc.data(), lc.size() * sizeof(float), hipMemcpyHostToDevice, st));
    GF_HIP(hipMemsetAsync(d_acc, 0, (size_t)nseg * (n + 2) * sizeof(unsigned long long), st));
    if (num_tiles <= 0) return GF_OK;
    const bool lds = (n + 2) <= SFS_LDS_BINS;
    if (lds)
        hipLaunchKernelGGL((k_sfs<true>), dim3(num_tiles), dim3(WG), 0, st, d_freq, d_seg_base, d_seg_M,
                           d_tile_seg, d_tile_first, n, half_n, d_lc.p, d_acc);
    else
        hipLaunchKernelGGL((k_sfs<false>), dim3(num_tiles), dim3(WG), 0, st, d_freq, d_seg_base, d_seg_M,
                           d_tile_seg, d_tile_first, n, half_n, d_lc.p, d_acc);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

int joint_sfs_of_device_rows(const float *d_fa, const float *d_fb, int M, int na, int nb, float num_sites, float *out,
                             hipStream_t st) {
    if ((int64_t)(na + 1) * (nb + 1) > ((int64_t)1 << 26) || na + nb + 2 > JS_LDS_FLOATS) {
        set_error("joint_site_frequency_spectrum: sample sizes too large");
        return GF_ERR_UNSUPPORTED;
    }
    std::vector<float> la, lb;
    int ha, hb;
    log_binom_table(na, la, ha);
    log_binom_table(nb, lb, hb);
    const int B = (na + 1) * (nb + 1);
    DevBuf<float> d_l;
    DevBuf<unsigned long long> d_acc;
    GF_TRY(d_l.alloc(la.size() + lb.size()));
    GF_TRY(d_acc.alloc((size_t)B));
    GF_HIP(hipMemcpyAsync(d_l.p, la.data(), la.size() * sizeof(float), hipMemcpyHostToDevice, st));
    GF_HIP(hipMemcpyAsync(d_l.p + la.size(), lb.data(), lb.size() * sizeof(float), hipMemcpyHostToDevice, st));
    GF_HIP(hipMemsetAsync(d_acc.p, 0, (size_t)B * sizeof(unsigned long long), st));
    if (M > 0) {
        const int batch = std::max(1, std::min(32, JS_LDS_FLOATS / (na + nb + 2)));
        const dim3 grid((unsigned)((M + JS_MUT_PER_WG - 1) / JS_MUT_PER_WG), (unsigned)((B + WG * JS_PER_THREAD - 1) / (WG * JS_PER_THREAD)));
        const size_t lds = (size_t)batch * (na + nb + 2) * sizeof(float);
        hipLaunchKernelGGL(k_joint_sfs, grid, dim3(WG), lds, st, d_fa, d_fb, M, na, nb, ha, hb, d_l.p,
                           d_l.p + la.size(), batch, d_acc.p);
        GF_HIP(hipGetLastError());
    }
    std::vector<unsigned long long> acc((size_t)B);
    GF_HIP(hipMemcpyAsync(acc.data(), d_acc.p, acc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    GF_HIP(hipStreamSynchronize(st));
    const double sc = 1.0 / 4294967296.0;
    for (int k = 0; k < B; k++) out[k] = (float)((double)acc[k] * sc + (k == 0 ? (double)num_sites - (double)M : 0.0));
    return GF_OK;
}

__global__ void k_sfs_prep(int n, int half_n, float *__restrict__ lc, unsigned long long *__restrict__ acc) {
    GF_STRICT_FP
    for (int k = threadIdx.x; k < n + 2; k += blockDim.x) acc[k] = 0ull;
    for (int k = 1 + (int)threadIdx.x; k < half_n; k += blockDim.x)
        lc[k] = log_f(((float)n + 1.0f - (float)k) / (float)k);
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.0f;
        lc[0] = 0.0f;
        for (int k = 1; k < half_n; k++) {
            a = a + lc[k];
            lc[k] = a;
        }
    }
}

__global__ void k_sfs_finish(const unsigned long long *__restrict__ acc, int n, int M, float num_sites,
                             float *__restrict__ out) {
    GF_STRICT_FP
    const double sc = 1.0 / 4294967296.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        if (k == 0)
            out[0] = (float)((double)(acc[0] + acc[n]) * sc + (double)acc[n + 1] + ((double)num_sites - (double)M));
        else
            out[k] = (float)((double)acc[k] * sc);
    }
}

/* sample SFS of one population row already in device memory (M floats) */
int sfs_of_device_row(const float *d_row, int M, int n, float num_sites, float *out, hipStream_t st,
                      SfsScratch *scratch) {
    SfsScratch local;
    SfsScratch &sc = scratch ? *scratch : local;
    DevBuf<float> &d_lc = sc.lc;
    DevBuf<unsigned long long> &d_acc = sc.acc;
    DevBuf<int> &d_meta = sc.meta;
    DevBuf<int64_t> &d_base = sc.base;
    const int tiles = (M + WG - 1) / WG;
    GF_TRY(d_acc.alloc((size_t)n + 2));
    GF_TRY(d_meta.alloc((size_t)tiles + 2));
    GF_TRY(d_base.alloc(1));
    std::vector<int> meta((size_t)tiles + 2, 0);
    meta[tiles] = M;      /* seg_M */
    meta[tiles + 1] = 0;  /* tile_first */
    const int64_t base0 = 0;
    GF_HIP(hipMemcpyAsync(d_meta.p, meta.data(), meta.size() * sizeof(int), hipMemcpyHostToDevice, st));
    GF_HIP(hipMemcpyAsync(d_base.p, &base0, sizeof(int64_t), hipMemcpyHostToDevice, st));
    GF_TRY(sfs_device(d_row, d_base.p, d_meta.p + tiles, d_meta.p, d_meta.p + tiles + 1, tiles, 1, n, d_acc.p, st, d_lc));
    std::vector<unsigned long long> acc((size_t)n + 2);
    GF_HIP(hipMemcpyAsync(acc.data(), d_acc.p, acc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    GF_HIP(hipStreamSynchronize(st));
    finish_sfs(acc.data(), acc[n + 1], n, M, num_sites, out);
    return GF_OK;
}

/* k_pop_hist over every bin range of [0, Nc): bin ranges as large as the
 * workgroup's LDS allows (at most HIST_MAX_CHUNK bins), and as many strips
 * per range as fill the device about twice over, each strip at least 8K
 * mutations */
static hipError_t launch_pop_hist(const float *d_freq, int M, int Nc, unsigned int *d_hist, hipStream_t st) {
    static int lds_bytes = -1;
    if (lds_bytes < 0) {
        int dev = 0, v = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e != hipSuccess) return e;
        v = std::min(v, HIST_MAX_CHUNK * (int)sizeof(unsigned int));
        e = hipFuncSetAttribute((const void *)k_pop_hist, hipFuncAttributeMaxDynamicSharedMemorySize, v);
        if (e != hipSuccess) return e;
        lds_bytes = v;
    }
    const int chunk = std::min(Nc, lds_bytes / (int)sizeof(unsigned int));
    const int ranges = (Nc + chunk - 1) / chunk;
    const int want = std::max(1, 2 * num_cus() / ranges);
    const int strips = std::max(1, std::min(want, (M + 8191) / 8192));
    hipLaunchKernelGGL(k_pop_hist, dim3(strips, ranges), dim3(HIST_WG), (size_t)chunk * sizeof(unsigned int), st,
                       d_freq, M, Nc, chunk, d_hist);
    return hipGetLastError();
}

/* ---------------- the sample SFS without allocation ----------------
 * Scratch of gf_site_frequency_spectrum_noalloc: the fixed-point level sums
 * (n + 2 words of 64 bits) and the log-binomial table (n/2 + 1 floats) of
 * one call, for sample sizes up to SFS_NOALLOC_MAX_N: 656 KiB in a
 * file-scope device array, one copy on every device the library runs on,
 * held until the process ends.  Host code never passes the array's name to
 * a kernel or copy (there it only stands for the device symbol): it takes
 * the device address from hipGetSymbolAddress, once per device with that
 * device current, and uses that pointer. */
constexpr int SFS_NOALLOC_MAX_N = 65536;
constexpr size_t SFS_NOALLOC_ACC_WORDS = (size_t)SFS_NOALLOC_MAX_N + 2;
constexpr size_t SFS_NOALLOC_LC_FLOATS = (size_t)SFS_NOALLOC_MAX_N / 2 + 1;
constexpr size_t SFS_NOALLOC_BYTES = SFS_NOALLOC_ACC_WORDS * sizeof(unsigned long long) +
                                     SFS_NOALLOC_LC_FLOATS * sizeof(float);
static __device__ __align__(256) unsigned char scratch[SFS_NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[MAX_DEVICES];

/* device address of `scratch` on the current device `dev` */
static int scratch_on_device(int dev, void **out) {
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("site_frequency_spectrum_noalloc: device index beyond " + std::to_string(MAX_DEVICES - 1));
        return GF_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("site_frequency_spectrum_noalloc: hipGetSymbolAddress failed: ") +
                      hipGetErrorString(e));
            return GF_ERR_DEVICE;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GF_OK;
}

} /* namespace gf */

using namespace gf;

extern "C" {

int32_t gf_population_histogram(const float *freq, int32_t M, int32_t Nc, float num_sites, float *out, int32_t device) {
    if (!out || Nc <= 0 || M < 0 || (M > 0 && !freq)) {
        set_error("population_frequency_histogram: invalid arguments");
        return GF_ERR_ARG;
    }
    GF_TRY(select_device(device));
    OwnedStream st;
    GF_HIP(st.create());
    StreamScope scope(st);
    int rc = GF_OK;
    {
        DevBuf<float> d_f;
        DevBuf<unsigned int> d_h;
        rc = d_f.alloc(std::max(M, 1));
        if (!rc) rc = d_h.alloc(Nc);
        if (!rc) {
            std::vector<unsigned int> h(Nc);
            hipError_t e = hipSuccess;
            if (M > 0) e = hipMemcpyAsync(d_f.p, freq, (size_t)M * sizeof(float), hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipMemsetAsync(d_h.p, 0, (size_t)Nc * sizeof(unsigned int), st);
            if (e == hipSuccess && M > 0) {
                e = launch_pop_hist(d_f.p, M, Nc, d_h.p, st);
            }
            if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_h.p, (size_t)Nc * sizeof(unsigned int), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                set_error(std::string("population_frequency_histogram: ") + hipGetErrorString(e));
                rc = GF_ERR_DEVICE;
            } else {
                /* bin 0 also holds the monomorphic sites (spectrum.cu:82) */
                const int ns = f2i_sat(__builtin_truncf(num_sites));
                h[0] += (unsigned int)(ns - M);
                for (int i = 0; i < Nc; i++) out[i] = (float)h[i];
            }
        }
    }
    return rc;
}

int32_t gf_site_frequency_spectrum(const float *freq, int32_t M, int32_t n, float num_sites, float *out, int32_t device) {
    if (!out || n <= 0 || M < 0 || (M > 0 && !freq)) {
        set_error("site_frequency_spectrum: invalid arguments");
        return GF_ERR_ARG;
    }
    GF_TRY(select_device(device));
    OwnedStream st;
    GF_HIP(st.create());
    StreamScope scope(st);
    int rc = GF_OK;
    {
        DevBuf<float> d_f;
        rc = d_f.alloc(std::max(M, 1));
        if (!rc && M > 0) {
            const hipError_t e = hipMemcpyAsync(d_f.p, freq, (size_t)M * sizeof(float), hipMemcpyHostToDevice, st);
            if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = GF_ERR_DEVICE; }
        }
        if (!rc) rc = sfs_of_device_row(d_f.p, M, n, num_sites, out, st);
    }
    return rc;
}

int32_t gf_site_frequency_spectrum_noalloc(const float *d_freq, int32_t M, int32_t n, float num_sites, float *d_out,
                                           void *stream) {
    if (!d_out || n <= 0 || M < 0 || (M > 0 && !d_freq)) {
        set_error("site_frequency_spectrum_noalloc: invalid arguments");
        return GF_ERR_ARG;
    }
    if (n > SFS_NOALLOC_MAX_N) {
        set_error("site_frequency_spectrum_noalloc: sample size " + std::to_string(n) + " above the maximum " +
                  std::to_string(SFS_NOALLOC_MAX_N) + " of the allocation-free entry point");
        return GF_ERR_ARG;
    }
    int dev = 0;
    GF_HIP(hipGetDevice(&dev));
    void *base = nullptr;
    GF_TRY(scratch_on_device(dev, &base));
    unsigned long long *acc = (unsigned long long *)base;
    float *lc = (float *)(acc + SFS_NOALLOC_ACC_WORDS);
    const int half_n = (n % 2 == 0) ? n / 2 + 1 : (n + 1) / 2;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_sfs_prep, dim3(1), dim3(WG), 0, st, n, half_n, lc, acc);
    if (M > 0) {
        const int grid = std::max(1, std::min((M + WG - 1) / WG, num_cus() * 8));
        if (n + 2 <= SFS_LDS_BINS)
            hipLaunchKernelGGL((k_sfs_row<true>), dim3(grid), dim3(WG), 0, st, d_freq, M, n, half_n, lc, acc);
        else
            hipLaunchKernelGGL((k_sfs_row<false>), dim3(grid), dim3(WG), 0, st, d_freq, M, n, half_n, lc, acc);
    }
    hipLaunchKernelGGL(k_sfs_finish, dim3(1), dim3(WG), 0, st, acc, n, M, num_sites, d_out);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

int32_t gf_joint_site_frequency_spectrum(const float *freq_a, const float *freq_b, int32_t M, int32_t n_a, int32_t n_b,
                                         float num_sites, float *out, int32_t device) {
    if (!out || n_a <= 0 || n_b <= 0 || M < 0 || (M > 0 && (!freq_a || !freq_b))) {
        set_error("joint_site_frequency_spectrum: invalid arguments");
        return GF_ERR_ARG;
    }
    GF_TRY(select_device(device));
    OwnedStream st;
    GF_HIP(st.create());
    StreamScope scope(st);
    int rc = GF_OK;
    {
        DevBuf<float> d_f;
        rc = d_f.alloc(2 * (size_t)std::max(M, 1));
        if (!rc && M > 0) {
            hipError_t e = hipMemcpyAsync(d_f.p, freq_a, (size_t)M * sizeof(float), hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipMemcpyAsync(d_f.p + M, freq_b, (size_t)M * sizeof(float), hipMemcpyHostToDevice, st);
            if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = GF_ERR_DEVICE; }
        }
        if (!rc) rc = joint_sfs_of_device_rows(d_f.p, d_f.p + M, M, n_a, n_b, num_sites, out, st);
    }
    return rc;
}

} /* extern "C" */
