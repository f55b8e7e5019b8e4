/*
 * gf_kernels.hpp -- library-only gfx950 kernels of the GO_Fish hot path:
 * stream compaction (batched path and initial counts), the
 * mutation-selection-equilibrium start and scans.  The drift kernel and the
 * shared types are in gf_drift.hpp.
 */
#ifndef GF_KERNELS_HPP_
#define GF_KERNELS_HPP_

#include "gf_drift.hpp"

namespace gf {

/* ---------------- prefix of the fused launch's tile counts ----------------
 * T = tiles of the launch (from the device slot count M_in + new);
 * pre[s] = survivors of tiles before s, pre[T] = total = *M_out; for every
 * tile j of the next launch (256 j < total) src0[j] = the tile holding
 * compacted index 256 j.
 * One workgroup walks the counts in chunks of SCAN_CHUNK.  For large
 * launches (chunk_sum != nullptr) k_tile_chunk_sums first sums every chunk
 * and workgroup b scans chunk b alone, starting from the sum of the chunks
 * before it; the grid then covers every chunk (the host's tile bound is at
 * least T). */
constexpr int SCAN_CHUNK = 16 * 1024;

static __global__ __launch_bounds__(1024) void k_tile_chunk_sums(const unsigned int *__restrict__ cnt,
                                                                 const int *__restrict__ M_in, int new_in_block,
                                                                 int *__restrict__ chunk_sum) {
    __shared__ int sh_w[16];
    const int T = (*M_in + new_in_block + WG - 1) / WG;
    const int base = (int)blockIdx.x * SCAN_CHUNK;
    if (base >= T) return;   /* the whole workgroup */
    const int b = base + (int)threadIdx.x * 16;
    int sum = 0;
    if (b + 16 <= T) {
        const uint4 *q = reinterpret_cast<const uint4 *>(cnt + b);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint4 u = q[k];
            sum += (int)(u.x + u.y + u.z + u.w);
        }
    } else {
        for (int i = 0; i < 16; i++) sum += b + i < T ? (int)cnt[b + i] : 0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if ((threadIdx.x & 63) == 0) sh_w[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < 16; k++) t += sh_w[k];
        chunk_sum[blockIdx.x] = t;
    }
}

static __global__ __launch_bounds__(1024) void k_tile_scan(const unsigned int *__restrict__ cnt, const int *__restrict__ M_in,
                                                           int new_in_block, int *__restrict__ pre, int *__restrict__ src0,
                                                           int *__restrict__ M_out, int *__restrict__ T_out,
                                                           int *M_host, const int *__restrict__ chunk_sum) {
    /* 16 counts per thread and round (four 16-byte loads), wave scans by
     * shuffles, one LDS exchange of the 16 wave totals */
    __shared__ int sh_w[16];
    __shared__ int sh_carry;
    const int T = (*M_in + new_in_block + WG - 1) / WG;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int first = (int)blockIdx.x * SCAN_CHUNK;
    int carry = 0;
    if (chunk_sum != nullptr && first < T) {
        /* survivors of the chunks before this workgroup's */
        if (w == 0) {
            int c = 0;
            for (int k = lane; k < (int)blockIdx.x; k += 64) c += chunk_sum[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
            if (lane == 0) sh_carry = c;
        }
        __syncthreads();
        carry = sh_carry;
    }
    const int stride = chunk_sum != nullptr ? (int)gridDim.x * SCAN_CHUNK : SCAN_CHUNK;
    for (int base = first; base < T; base += stride) {
        const int b = base + (int)threadIdx.x * 16;
        unsigned int v[16];
        if (b + 16 <= T) {
            const uint4 *q = reinterpret_cast<const uint4 *>(cnt + b);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint4 u = q[k];
                v[4 * k] = u.x; v[4 * k + 1] = u.y; v[4 * k + 2] = u.z; v[4 * k + 3] = u.w;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) v[i] = b + i < T ? cnt[b + i] : 0u;
        }
        int sum = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) sum += (int)v[i];
        int x = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) sh_w[w] = x;
        __syncthreads();
        if (w == 0) {
            int t = lane < 16 ? sh_w[lane] : 0;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                const int y = __shfl_up(t, off);
                if (lane >= off) t += y;
            }
            if (lane < 16) sh_w[lane] = t;
        }
        __syncthreads();
        if (threadIdx.x == 0 && base + SCAN_CHUNK >= T) {
            /* the last round: the totals are known, and their stores (one
             * of them to host memory) overlap the prefix stores below */
            const int total = carry + sh_w[15];
            pre[T] = total;
            *M_out = total;
            *T_out = T;
            if (M_host) *M_host = total;   /* pinned host word: the host reads it after the launch's event */
        }
        int run = carry + (w > 0 ? sh_w[w - 1] : 0) + x - sum;
        int pv[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int idx = b + i;
            pv[i] = run;
            if (idx < T) {
                const int c = (int)v[i];
                if (c > 0) {
                    for (int j = (run + WG - 1) / WG; j * WG <= run + c - 1; j++) src0[j] = idx;
                }
                run += c;
            }
        }
        if (b + 16 <= T) {   /* four 16-byte stores (b is a multiple of 16) */
            int4 *q = reinterpret_cast<int4 *>(pre + b);
#pragma unroll
            for (int k = 0; k < 4; k++) q[k] = make_int4(pv[4 * k], pv[4 * k + 1], pv[4 * k + 2], pv[4 * k + 3]);
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) if (b + i < T) pre[b + i] = pv[i];
        }
        carry += sh_w[15];
        __syncthreads();
    }
    if (T == 0 && threadIdx.x == 0 && blockIdx.x == 0) {
        pre[0] = 0;
        *M_out = 0;
        *T_out = 0;
        if (M_host) *M_host = 0;
    }
}

/* live count of each generation: the sum of its 64 counters (the drift
 * kernel spreads its atomics over 64 addresses per generation) */
static __global__ __launch_bounds__(256) void k_sum_live(const unsigned long long *__restrict__ live,
                                                         unsigned long long *__restrict__ out, int rows) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= rows) return;
    const unsigned long long *r = live + (size_t)t * 64;
    unsigned long long s = 0;
#pragma unroll 8
    for (int k = 0; k < 64; k++) s += r[k];
    out[t] = s;
}

/* dense copy of a fragmented state: slot j of the output = compacted index j */
template <int P>
__global__ __launch_bounds__(WG) void k_densify(const float *__restrict__ freq_in, const gf_mutid *__restrict__ ids_in,
                                                float *__restrict__ freq_out, gf_mutid *__restrict__ ids_out, int64_t pitch,
                                                const int *__restrict__ pre, const int *__restrict__ src0,
                                                const int *__restrict__ T_in, const int *__restrict__ M_in) {
    __shared__ int sh_pre[WG + 1];
    const int M = *M_in, T = *T_in;
    for (int tile = blockIdx.x; tile * WG < M; tile += gridDim.x) {
        __syncthreads();
        const int s0 = load_frag_window(tile, src0, pre, T, sh_pre);
        const int li = tile * WG + (int)threadIdx.x;
        if (li < M) {
            const int64_t src = frag_slot(li, s0, sh_pre, pre, T);
#pragma unroll
            for (int p = 0; p < P; p++) freq_out[p * pitch + li] = freq_in[p * pitch + src];
            ids_out[li] = ids_in[src];
        }
    }
}

/* ---------------- compaction (stable stream compaction) ---------------- */

/* flag pass: one bit per slot (wave64 ballot), one count per 256-slot tile
 * (reference flag_segregating_mutations, go_fish_impl.cuh:137-176) */
template <int P>
__global__ __launch_bounds__(WG) void k_flag(const float *__restrict__ freq, int64_t pitch,
                                             const gf_mutid *__restrict__ ids,
                                             const int64_t *__restrict__ seg_base,
                                             const int *__restrict__ seg_M,
                                             const int *__restrict__ seg_present,
                                             const int *__restrict__ tile_first, int nseg,
                                             unsigned long long *__restrict__ masks,
                                             unsigned int *__restrict__ tile_counts) {
    const int tile = blockIdx.x;
    const int s = seg_of_tile(tile_first, nseg, tile);
    const int li = (tile - tile_first[s]) * WG + (int)threadIdx.x;
    const int64_t slot = seg_base[s] + li;
    bool keep = false;
    if (li < seg_M[s]) {
        float x[P];
#pragma unroll
        for (int p = 0; p < P; p++) x[p] = freq[p * pitch + slot];
        keep = keep_rule<P>(x, seg_present[s], ids[slot].origin_threadID < 0);
    }
    const uint64_t b = __ballot(keep);
    __shared__ unsigned int sh[WG / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        masks[(size_t)tile * (WG / 64) + w] = b;
        sh[w] = (unsigned int)__popcll(b);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int c = 0;
        for (int k = 0; k < WG / 64; k++) c += sh[k];
        tile_counts[tile] = c;
    }
}

/* exclusive scan of tile counts within each segment (one workgroup per
 * segment, deterministic integer sums); writes per-tile offsets and the
 * segment's survivor count */
static __global__ __launch_bounds__(1024) void k_scan_tiles(const unsigned int *__restrict__ tile_counts,
                                                     const int *__restrict__ tile_first,
                                                     const int *__restrict__ tile_num,
                                                     unsigned int *__restrict__ tile_off,
                                                     int *__restrict__ seg_total) {
    const int s = blockIdx.x;
    const int first = tile_first[s], n = tile_num[s];
    __shared__ unsigned int sh[1024];
    const int T = blockDim.x;
    const int per = (n + T - 1) / T;
    const int b = threadIdx.x * per;
    const int e = b + per < n ? b + per : n;
    unsigned int sum = 0;
    for (int k = b; k < e; k++) sum += tile_counts[first + k];
    sh[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        unsigned int v = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned int run = sh[threadIdx.x] - sum;
    for (int k = b; k < e; k++) {
        tile_off[first + k] = run;
        run += tile_counts[first + k];
    }
    if (threadIdx.x == T - 1) seg_total[s] = (int)sh[T - 1];
}


/* ---------------- mutation-selection-equilibrium initialisation ---------------- */

struct MseSeg {
    int Nc, Nind;
    float F, h, mu, sa, sb, L;
    uint32_t k0, k1;
    int pop;            /* population index: Philox counter word and freq row */
    int neutral;        /* selection identically 0 at generation 0 */
    int preserve;       /* mark initial mutations preserved */
    int pad;
    int64_t class_off;  /* first class in the count arrays (Nc-1 classes) */
    int64_t area_off;   /* first entry of the integral arrays (Nc entries) */
    int64_t chunk_off;  /* first chunk total ((Nc+63)/64 chunks) */
    int64_t slot_base;  /* first slot of this segment's mutations */
    int64_t sim_base;   /* slot where the owning simulation starts (threadID origin) */
    int64_t tid_base;   /* threadID offset of this segment within its simulation */
    const float *sel_tab;   /* a caller's selection at the equilibrium points (k_mse_sel_table
                               layout), or NULL: the linear form sa * x + sb */
};

constexpr int MSE_CHUNK = 64;

__device__ __forceinline__ int find_seg(const int64_t *__restrict__ offs, int nseg, int64_t idx) {
    int lo = 0, hi = nseg - 1;   /* last s with offs[s] <= idx */
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= idx) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ float mse_integrand(const MseSeg &S, float i, float s) {
    GF_STRICT_FP
    s = s > -1.0f ? s : -1.0f;
    return mse_weight(i, S.Nind, S.F, S.h, -1.0f * s);
}

/* trapezoid areas from 1 down to 0 (reference calculate_area,
 * go_fish_impl.cuh:48-53, with trapezoidal_upper :39-45) */
static __global__ void k_mse_area(const MseSeg *__restrict__ segs, const int64_t *__restrict__ area_offs,
                           int nseg, int64_t total, float *__restrict__ area) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
        const int s = find_seg(area_offs, nseg, k);
        const MseSeg &S = segs[s];
        const int id = (int)(k - S.area_off);
        const float step = 1.0f / (float)S.Nc;
        float a, b;
        mse_area_points(id, S.Nc, a, b);
        const float sa = S.sel_tab ? S.sel_tab[id] : S.sa * a + S.sb;
        const float sb = S.sel_tab ? S.sel_tab[S.Nc + id] : S.sa * b + S.sb;
        const float f1 = mse_integrand(S, a, sa);
        const float f2 = mse_integrand(S, b, sb);
        area[k] = (step * (f1 + f2)) / 2.0f;
    }
}

/* chunk-local inclusive scans (sequential within 64-entry chunks) */
static __global__ void k_mse_chunk_local(const MseSeg *__restrict__ segs, const int64_t *__restrict__ chunk_offs,
                                  int nseg, int64_t total_chunks, float *__restrict__ area,
                                  float *__restrict__ tot) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total_chunks; k += (int64_t)gridDim.x * blockDim.x) {
        const int s = find_seg(chunk_offs, nseg, k);
        const MseSeg &S = segs[s];
        const int c = (int)(k - S.chunk_off);
        const int b = c * MSE_CHUNK;
        const int e = b + MSE_CHUNK < S.Nc ? b + MSE_CHUNK : S.Nc;
        float acc = 0.0f;
        for (int i = b; i < e; i++) {
            acc = (i == b) ? area[S.area_off + i] : acc + area[S.area_off + i];
            area[S.area_off + i] = acc;
        }
        tot[k] = acc;
    }
}

/* sequential exclusive prefix of chunk totals, one thread per segment */
static __global__ void k_mse_chunk_prefix(const MseSeg *__restrict__ segs, int nseg, float *__restrict__ tot) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const MseSeg &S = segs[s];
    if (S.neutral) return;
    const int nc = (S.Nc + MSE_CHUNK - 1) / MSE_CHUNK;
    float pre = 0.0f;
    for (int c = 0; c < nc; c++) {
        const float t = tot[S.chunk_off + c];
        tot[S.chunk_off + c] = pre;
        pre = pre + t;
    }
}

static __global__ void k_mse_chunk_add(const MseSeg *__restrict__ segs, const int64_t *__restrict__ area_offs,
                                int nseg, int64_t total, float *__restrict__ area,
                                const float *__restrict__ tot) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
        const int s = find_seg(area_offs, nseg, k);
        const MseSeg &S = segs[s];
        const int id = (int)(k - S.area_off);
        const int c = id / MSE_CHUNK;
        if (c > 0) area[k] = tot[S.chunk_off + c] + area[k];
    }
}

/* expected count per frequency class and its Poisson draw (reference
 * initialize_mse_frequency_array, go_fish_impl.cuh:58-71) */
static __global__ void k_mse_counts(const MseSeg *__restrict__ segs, const int64_t *__restrict__ class_offs,
                             int nseg, int64_t total, const float *__restrict__ S_int,
                             int *__restrict__ counts) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
        const int s = find_seg(class_offs, nseg, k);
        const MseSeg &S = segs[s];
        const int id = (int)(k - S.class_off);
        const int Nc = S.Nc;
        const float i = mse_class_point(id, Nc);
        const float j = ((float)(Nc - id) + 1.0f) / (float)Nc;
        const float sel = S.sel_tab ? S.sel_tab[2 * Nc + id] : S.sa * i + S.sb;
        const float twomuL = (2.0f * S.mu) * S.L;
        float lambda;
        if (sel == 0.0f) {
            lambda = twomuL / i;
        } else {
            /* integral from i to 1 is S[Nc-1-id]; total integral S[Nc-1] */
            const float I_id = S_int[S.area_off + (Nc - 1 - id)];
            const float I_0 = S_int[S.area_off + (Nc - 1)];
            lambda = (twomuL * (mse_weight(i, S.Nind, S.F, S.h, sel) * I_id)) / ((I_0 * i) * j);
        }
        const u32x4 r = philox10(S.k0, S.k1, 0u, (uint32_t)id, (uint32_t)S.pop, 0u);
        const int n = pois_draw(r.x, lambda, lambda, S.L * (float)Nc);
        counts[k] = n > 0 ? n : 0;
    }
}

/* materialise the equilibrium mutations: slot j of segment s belongs to the
 * class whose exclusive offset is the last one <= j (reference
 * initialize_mse_mutation_array go_fish_impl.cu:45-57 and mse_set_mutID
 * :59-70) */
template <int P>
static __global__ void k_mse_fill(const MseSeg *__restrict__ segs, const int64_t *__restrict__ out_offs,
                           int nseg, int64_t total, const int64_t *__restrict__ class_scan,
                           float *__restrict__ freq, int64_t pitch, gf_mutid *__restrict__ ids) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x) {
        const int s = find_seg(out_offs, nseg, j);
        const MseSeg &S = segs[s];
        const int64_t local = j - out_offs[s];
        /* class: last c in [0, Nc-2] with scan[c] - scan[0] <= local */
        const int64_t c0 = S.class_off;
        const int64_t base = class_scan[c0];
        int lo = 0, hi = S.Nc - 2;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (class_scan[c0 + mid] - base <= local) lo = mid; else hi = mid - 1;
        }
        const float f = ((float)lo + 1.0f) / (float)S.Nc;
        const int64_t slot = S.slot_base + local;
#pragma unroll
        for (int p = 0; p < P; p++) freq[p * pitch + slot] = (p == S.pop) ? f : 0.0f;
        const int tid = (int)(slot - S.sim_base + 1);
        gf_mutid id;
        id.origin_generation = 0;
        id.origin_population = S.pop;
        id.origin_threadID = S.preserve ? -tid : tid;
        id.reserved = 0;
        ids[slot] = id;
    }
}

/* generic exclusive scan of int32 counts into int64 offsets (three-phase,
 * deterministic): per-block sums, single-block scan of the sums, per-block
 * rescan */
static __global__ __launch_bounds__(256) void k_scan_partial(const int *__restrict__ in, int64_t n, int64_t per_block,
                                                      int64_t *__restrict__ block_sums) {
    const int64_t b = blockIdx.x * per_block;
    const int64_t e = b + per_block < n ? b + per_block : n;
    int64_t s = 0;
    for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) s += in[k];
    __shared__ int64_t sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = sh[0];
}

static __global__ __launch_bounds__(1024) void k_scan_sums(int64_t *__restrict__ sums, int nb, int64_t *__restrict__ total) {
    __shared__ int64_t sh[1024];
    const int T = blockDim.x;
    const int per = (nb + T - 1) / T;
    const int b = threadIdx.x * per, e = b + per < nb ? b + per : nb;
    int64_t s = 0;
    for (int k = b; k < e; k++) s += sums[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        int64_t v = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t run = sh[threadIdx.x] - s;
    for (int k = b; k < e; k++) { int64_t t = sums[k]; sums[k] = run; run += t; }
    if (threadIdx.x == T - 1) *total = sh[T - 1];
}

static __global__ __launch_bounds__(256) void k_scan_final(const int *__restrict__ in, int64_t n, int64_t per_block,
                                                    const int64_t *__restrict__ block_off,
                                                    int64_t *__restrict__ out) {
    /* each block rescans its range in 256-element rounds */
    const int64_t b = blockIdx.x * per_block;
    const int64_t e = b + per_block < n ? b + per_block : n;
    __shared__ int64_t sh[256];
    int64_t run = block_off[blockIdx.x];
    for (int64_t r = b; r < e; r += 256) {
        const int64_t k = r + threadIdx.x;
        const int64_t v = k < e ? in[k] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            int64_t t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (k < e) out[k] = run + sh[threadIdx.x] - v;
        const int64_t tot = sh[255];
        __syncthreads();
        run += tot;
    }
}

} /* namespace gf */

#endif /* GF_KERNELS_HPP_ */
