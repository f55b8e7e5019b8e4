/*
 * gf_kernels.hpp -- gfx950 kernels of the GO_Fish hot path.
 *
 * Layout in HBM (one "segment" = one simulation; the plain run_sim path is a
 * batch of one segment):
 *   freq  : float [P][cap]  structure of arrays, row p = population p
 *   ids   : gf_mutid [cap]  16 B per slot (origin gen, origin pop, threadID, 0)
 * A thread owns one mutation slot for all P populations, so migration is a
 * register-local dot product and the update is done in place.
 *
 * k_advance runs up to KMAX generations per launch ("temporal blocking"):
 * each slot is read once, advanced through K generations of migration,
 * selection and drift in registers, and written once.  Mutations born inside
 * the block are materialised in their slots at the generation they arise and
 * drifted from the next generation on, exactly as the generation-by-
 * generation order of the reference (go_fish_impl.cuh:645-714) prescribes.
 */
#ifndef GF_KERNELS_HPP_
#define GF_KERNELS_HPP_

#include "gf_math.hpp"
#include "../../include/gofish_capi.h"

namespace gf {

constexpr int KMAX = 64;   /* generations per advance launch (lane-indexed counters) */
constexpr int WG = 256;    /* threads per workgroup = slots per tile */

/* Per-generation model row of one segment (host-built, read through scalar
 * loads: every field is uniform across a tile). */
template <int P>
struct alignas(16) GenRow {
    double invN[P];       /* 1/(double)fN.  For floats a and b, (float)((double)a * (1/(double)b))
                           * equals the correctly rounded a/b: a/b has a denominator below 2^24, so
                           * it is never a float rounding midpoint and lies at least 2^-49 (relative)
                           * from one, while the double product is within 2^-52 of a/b. */
    int   Nc[P];          /* effective chromosomes; <= 0: population not drifted -> 0 */
    float fN[P];          /* (float)Nc */
    float sa[P], sb[P];   /* selection s(x) = sa*x + sb */
    float c1[P], c2[P];   /* (F+h)-(h*F), (F+2h)-(2h*F) */
    float mig[P][P];      /* [from][to] */
    float newfreq[P];     /* 1/Nc of new mutations */
    int   newcum[P + 1];  /* new mutations since run start before (g,p); [P] = after g */
    int   present_mask;   /* bit p: demography(p,g) > 0 */
    int   neutral_mask;   /* bit p: selection identically 0 at g */
};

/* Per-launch description of one segment. */
template <int P>
struct SegLaunch {
    const GenRow<P> *rows;        /* rows[t] = generation g0+1+t */
    unsigned long long *live;     /* plain: [run gens][64]; batch: single counter */
    int64_t base;                 /* first slot of the segment */
    int   M0;                     /* slots in use before the block */
    int   Mend;                   /* slots in use after the block */
    int   cum_base;               /* rows[0].newcum[0] */
    int   K;                      /* generations of this segment in the block */
    int   g0;                     /* generation before the block */
    int   t0;                     /* run index (g0+1 - start) of rows[0] */
    uint32_t k0, k1;              /* Philox key */
    int   pad;
};

/* Model rows and segment descriptors are read-only for a whole launch.
 * Reading them through the constant address space lets the compiler use
 * scalar loads (and uniform branches) where the index is wave-uniform,
 * which it cannot prove for a generic pointer next to the kernel's own
 * stores. */
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *as_const(const T *p) {
    return (const __attribute__((address_space(4))) T *)p;
}

/* a segment descriptor read with scalar loads (see as_const) */
template <int P>
__device__ __forceinline__ SegLaunch<P> load_seg(const SegLaunch<P> *p) {
    const auto &c = *as_const(p);
    SegLaunch<P> S;
    S.rows = c.rows; S.live = c.live; S.base = c.base; S.M0 = c.M0; S.Mend = c.Mend; S.cum_base = c.cum_base;
    S.K = c.K; S.g0 = c.g0; S.t0 = c.t0; S.k0 = c.k0; S.k1 = c.k1; S.pad = c.pad;
    return S;
}

/* keep rule of the compaction flag (reference go_fish_impl.cuh:153-166) */
template <int P>
__device__ __forceinline__ bool keep_rule(const float *x, int present_mask, bool preserved) {
    bool zero = true, one = true;
#pragma unroll
    for (int p = 0; p < P; p++) {
        if (present_mask & (1 << p)) {
            zero = zero && (x[p] <= 0.0f);
            one = one && (x[p] >= 1.0f);
        }
    }
    return !(zero || one) || preserved;
}

/* ---------------- migration, selection and drift of one generation ----------------
 * Reference migration_selection_drift (go_fish_impl.cuh:93-124): each
 * population's frequency after migration (weighted sum over the source
 * populations), selection (Eq. 13 of the reference's model, with F and h)
 * and an approximate binomial(N, x') drift draw, clamped to [0, N].
 *
 * rc[p] caches the lane's Philox block for generations 4b..4b+3 (the
 * reference's float4 packing, go_fish_impl.cuh:107).  It is fetched when
 * `refresh` is set (at a block boundary, or in the first generation a slot
 * is drifted in this launch), also for a population that is empty in this
 * generation, so that it is current if the population is founded later in
 * the same block. */
template <int P, typename Row>
__device__ __forceinline__ void generation_update(float *x, const Row &R, int g, bool refresh, u32x4 *rc,
                                                  uint32_t k0, uint32_t k1, uint32_t tid, int ogen, int opop) {
    float nx[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int Nc = R.Nc[p];
        if (refresh) rc[p] = drift_block(k0, k1, tid, ogen, opop, p, g);
        if (Nc <= 0) { nx[p] = 0.0f; continue; }
        float im = 0.0f;
#pragma unroll
        for (int k = 0; k < P; k++) im = im + R.mig[k][p] * x[k];
        float xs;
        if (R.neutral_mask & (1 << p)) {
            xs = im;   /* s == 0: Eq. 13 reduces exactly to x */
        } else {
            float s = R.sa[p] * im + R.sb[p];
            s = s > -1.0f ? s : -1.0f;
            const float om = 1.0f - im;
            const float num = (((s * im) * im) + im) + (((R.c1[p] * s) * im) * om);
            const float den = (((im * im) * s) + (((R.c2[p] * s) * im) * om)) + 1.0f;
            xs = num / den;
        }
        const float N = R.fN[p];
        const float mean = xs * N;
        const float var = (1.0f - xs) * mean;
        int j = binom_draw(lane_of(rc[p], g & 3), mean, var, N);
        j = j < 0 ? 0 : (j > Nc ? Nc : j);
        nx[p] = (float)((double)(float)j * R.invN[p]);   /* == (float)j / N */
    }
#pragma unroll
    for (int p = 0; p < P; p++) x[p] = nx[p];
}

/* locate a slot born inside the block: generation index t and population */
template <int P>
__device__ __forceinline__ void find_birth(const GenRow<P> *rows, int K, int rel, int &t, int &p) {
    int lo = 0, hi = K - 1;          /* last t with rows[t].newcum[0] <= rel */
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (rows[mid].newcum[0] <= rel) lo = mid; else hi = mid - 1;
    }
    t = lo;
    p = 0;
#pragma unroll
    for (int q = 1; q < P; q++) if (rows[lo].newcum[q] <= rel) p = q;
}

/* Advance every slot of every segment by its K generations.
 * BATCH=false: one segment, per-generation live counts (lane t of each wave
 * accumulates generation t, KMAX <= 64).  BATCH=true: tile_seg maps tiles to
 * segments and only the total of live mutation-generations is counted. */
template <int P, bool BATCH>
__global__ __launch_bounds__(WG) void k_advance(float *__restrict__ freq, int64_t pitch,
                                                gf_mutid *__restrict__ ids,
                                                const SegLaunch<P> seg0,
                                                const SegLaunch<P> *__restrict__ segs,
                                                const int *__restrict__ tile_seg,
                                                const int *__restrict__ tile_first,
                                                int num_tiles) {
    const int lane = threadIdx.x & 63;
    uint64_t cnt = 0;
    for (int tile = blockIdx.x; tile < num_tiles; tile += gridDim.x) {
        const int s = BATCH ? tile_seg[tile] : 0;
        SegLaunch<P> S = seg0;
        if (BATCH) S = load_seg<P>(segs + s);
        const int li = (BATCH ? (tile - tile_first[s]) : tile) * WG + (int)threadIdx.x;
        const bool valid = li < S.Mend;
        const int64_t slot = S.base + li;
        const GenRow<P> *rows = S.rows;
        float x[P];
        int ogen = 0, opop = 0, born = -1;
        uint32_t tid = 0;
        bool preserved = false;
#pragma unroll
        for (int p = 0; p < P; p++) x[p] = 0.0f;
        if (valid) {
            if (li < S.M0) {
#pragma unroll
                for (int p = 0; p < P; p++) x[p] = freq[p * pitch + slot];
                const gf_mutid id = ids[slot];
                ogen = id.origin_generation;
                opop = id.origin_population;
                preserved = id.origin_threadID < 0;
                tid = (uint32_t)(preserved ? -id.origin_threadID : id.origin_threadID);
            } else {
                const int rel = li - S.M0 + S.cum_base;
                int tb, pb;
                find_birth<P>(rows, S.K, rel, tb, pb);
                born = tb;
                ogen = S.g0 + 1 + tb;
                opop = pb;
                tid = (uint32_t)(rel - rows[tb].newcum[pb] + 1);
#pragma unroll
                for (int p = 0; p < P; p++) x[p] = (p == pb) ? rows[tb].newfreq[p] : 0.0f;
                gf_mutid id;
                id.origin_generation = ogen;
                id.origin_population = opop;
                id.origin_threadID = (int32_t)tid;
                id.reserved = 0;
                ids[slot] = id;
            }
        }
        u32x4 rc[P];
        const auto *crows = as_const(rows);
        for (int t = 0; t < S.K; t++) {
            const auto &R = crows[t];
            const int g = S.g0 + 1 + t;
            if (valid && born < t) {
                const bool refresh = ((g & 3) == 0) || (t == born + 1);
                generation_update<P>(x, R, g, refresh, rc, S.k0, S.k1, tid, ogen, opop);
            }
            const bool keep = valid && born <= t && keep_rule<P>(x, R.present_mask, preserved);
            const uint64_t b = __ballot(keep);
            if (BATCH) {
                if (lane == 0) cnt += (uint64_t)__popcll(b);
            } else {
                if (lane == t) cnt += (uint64_t)__popcll(b);
            }
        }
        if (valid) {
#pragma unroll
            for (int p = 0; p < P; p++) freq[p * pitch + slot] = x[p];
        }
        if (BATCH) {
            /* flush per tile: segments differ between tiles */
            if (lane == 0 && cnt) atomicAdd(S.live, (unsigned long long)cnt);
            cnt = 0;
        }
    }
    if (!BATCH) {
        __shared__ unsigned long long sh[64];
        if (threadIdx.x < 64) sh[threadIdx.x] = 0;
        __syncthreads();
        if (cnt) atomicAdd(&sh[lane], (unsigned long long)cnt);
        __syncthreads();
        const SegLaunch<P> &S = seg0;
        if ((int)threadIdx.x < S.K && sh[threadIdx.x])
            atomicAdd(&S.live[(size_t)(S.t0 + threadIdx.x) * 64 + (blockIdx.x & 63)], sh[threadIdx.x]);
    }
}

/* ---------------- fused advance + stable compaction ----------------
 * One launch advances every slot of a segment by K generations and, when
 * `compact` is set, writes only the survivors (compaction keep rule at the
 * last generation of the block) in their original order into the output
 * buffers; otherwise every slot is written to the same position.  The order
 * is kept by a single-pass decoupled look-back scan: workgroups take tiles
 * in increasing order from a ticket counter, publish their tile's survivor
 * count, and read their predecessors' published (aggregate or inclusive)
 * counts.  Each published word is (epoch << 34) | (flag << 32) | count with
 * flag 1 = aggregate, 2 = inclusive prefix; the epoch identifies the launch,
 * so the status array never needs clearing.  The slot count before the
 * block is read from the device (M_in) and the count after it is written
 * there (M_out): the host never needs it to launch the next block. */
struct FusedLaunch {
    const float *freq_in;
    const gf_mutid *ids_in;
    float *freq_out;
    gf_mutid *ids_out;
    int64_t pitch;
    const int *M_in;
    int *M_out;
    unsigned long long *status;   /* one word per tile */
    unsigned int *ticket;         /* zero before the launch */
    int *error;                   /* set if a look-back wait exceeded its bound */
    unsigned long long epoch;
    int ntiles;                   /* tiles launched (upper bound of the needed ones) */
    int new_in_block;
    int compact;
    int preserve;
};

constexpr unsigned long long FLAG_AGG = 1ull, FLAG_INC = 2ull;

__device__ __forceinline__ unsigned long long status_word(unsigned long long epoch, unsigned long long flag, unsigned int v) {
    return (epoch << 34) | (flag << 32) | (unsigned long long)v;
}

template <int P>
__global__ __launch_bounds__(WG) void k_advance_fused(const FusedLaunch F, const SegLaunch<P> S) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    __shared__ int sh_tile;
    __shared__ unsigned int sh_wcount[WG / 64];
    __shared__ unsigned int sh_excl;
    __shared__ unsigned long long sh_live[64];
    if (threadIdx.x < 64) sh_live[threadIdx.x] = 0;
    const int M0 = *F.M_in;
    const int Mend = M0 + F.new_in_block;
    const int need_tiles = (Mend + WG - 1) / WG;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        /* without compaction every slot is kept in place */
        if (Mend == 0 || !F.compact) *F.M_out = Mend;
        if (need_tiles > F.ntiles) *F.error = 2;   /* host bound too small: slots would be dropped */
    }
    uint64_t cnt = 0;
    const GenRow<P> *rows = S.rows;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) sh_tile = (int)atomicAdd(F.ticket, 1u);
        __syncthreads();
        const int tile = sh_tile;
        if (tile >= F.ntiles || tile >= need_tiles) break;
        const int li = tile * WG + (int)threadIdx.x;
        const bool valid = li < Mend;
        float x[P];
        int ogen = 0, opop = 0, born = -1;
        uint32_t tid = 0;
        bool preserved = false;
#pragma unroll
        for (int p = 0; p < P; p++) x[p] = 0.0f;
        gf_mutid id;
        id.origin_generation = 0; id.origin_population = 0; id.origin_threadID = 0; id.reserved = 0;
        if (valid) {
            if (li < M0) {
#pragma unroll
                for (int p = 0; p < P; p++) x[p] = F.freq_in[p * F.pitch + li];
                id = F.ids_in[li];
                ogen = id.origin_generation;
                opop = id.origin_population;
                preserved = id.origin_threadID < 0;
                tid = (uint32_t)(preserved ? -id.origin_threadID : id.origin_threadID);
            } else {
                const int rel = li - M0 + S.cum_base;
                int tb, pb;
                find_birth<P>(rows, S.K, rel, tb, pb);
                born = tb;
                ogen = S.g0 + 1 + tb;
                opop = pb;
                tid = (uint32_t)(rel - rows[tb].newcum[pb] + 1);
#pragma unroll
                for (int p = 0; p < P; p++) x[p] = (p == pb) ? rows[tb].newfreq[p] : 0.0f;
                id.origin_generation = ogen;
                id.origin_population = opop;
                id.origin_threadID = (int32_t)tid;
                id.reserved = 0;
            }
        }
        u32x4 rc[P];
        bool keep = false;
        const auto *crows = as_const(rows);
        for (int t = 0; t < S.K; t++) {
            const auto &R = crows[t];
            const int g = S.g0 + 1 + t;
            if (valid && born < t) {
                const bool refresh = ((g & 3) == 0) || (t == born + 1);
                generation_update<P>(x, R, g, refresh, rc, S.k0, S.k1, tid, ogen, opop);
            }
            keep = valid && born <= t && keep_rule<P>(x, R.present_mask, preserved);
            const uint64_t b = __ballot(keep);
            if (lane == t) cnt += (uint64_t)__popcll(b);
        }
        if (!F.compact) {
            if (valid) {
#pragma unroll
                for (int p = 0; p < P; p++) F.freq_out[p * F.pitch + li] = x[p];
                F.ids_out[li] = id;
            }
            continue;
        }
        /* stable compaction of this tile */
        const uint64_t kb = __ballot(keep);
        if (lane == 0) sh_wcount[w] = (unsigned int)__popcll(kb);
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned int c = 0;
            for (int k = 0; k < WG / 64; k++) c += sh_wcount[k];
            unsigned int excl = 0;
            if (tile == 0) {
                __hip_atomic_store(&F.status[0], status_word(F.epoch, FLAG_INC, c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                __hip_atomic_store(&F.status[tile], status_word(F.epoch, FLAG_AGG, c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                int p = tile - 1;
                long spins = 0;
                while (p >= 0) {
                    const unsigned long long v = __hip_atomic_load(&F.status[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const unsigned long long flag = (v >> 32) & 3ull;
                    if ((v >> 34) != F.epoch || flag == 0ull) {
                        if (++spins > (1l << 26)) { *F.error = 1; break; }
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    excl += (unsigned int)v;
                    if (flag == FLAG_INC) break;
                    p--;
                }
                __hip_atomic_store(&F.status[tile], status_word(F.epoch, FLAG_INC, excl + c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            sh_excl = excl;
            if (tile == need_tiles - 1) *F.M_out = (int)(excl + c);
        }
        __syncthreads();
        if (keep) {
            unsigned int pos = sh_excl;
            for (int k = 0; k < w; k++) pos += sh_wcount[k];
            pos += __builtin_amdgcn_mbcnt_hi((uint32_t)(kb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)kb, 0u));
#pragma unroll
            for (int p = 0; p < P; p++) F.freq_out[p * F.pitch + pos] = x[p];
            if (F.preserve) {
                const int z = id.origin_threadID;
                id.origin_threadID = -(z < 0 ? -z : z);
            }
            F.ids_out[pos] = id;
        }
    }
    /* per-generation live counts: lane t of every wave holds generation t */
    if (cnt) atomicAdd(&sh_live[lane], (unsigned long long)cnt);
    __syncthreads();
    if ((int)threadIdx.x < S.K && sh_live[threadIdx.x])
        atomicAdd(&S.live[(size_t)(S.t0 + threadIdx.x) * 64 + (blockIdx.x & 63)], sh_live[threadIdx.x]);
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
                                             const int *__restrict__ tile_seg,
                                             const int *__restrict__ tile_first,
                                             unsigned long long *__restrict__ masks,
                                             unsigned int *__restrict__ tile_counts) {
    const int tile = blockIdx.x;
    const int s = tile_seg[tile];
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

/* scatter pass: write survivors in order into the new arrays; preserved
 * survivors get threadID = -|threadID| (reference scatter_arrays,
 * go_fish_impl.cu:147-189) */
template <int P>
__global__ __launch_bounds__(WG) void k_scatter(const float *__restrict__ freq, int64_t pitch,
                                                const gf_mutid *__restrict__ ids,
                                                float *__restrict__ nfreq, int64_t npitch,
                                                gf_mutid *__restrict__ nids,
                                                const int64_t *__restrict__ seg_base,
                                                const int64_t *__restrict__ seg_nbase,
                                                const int *__restrict__ seg_preserve,
                                                const int *__restrict__ tile_seg,
                                                const int *__restrict__ tile_first,
                                                const unsigned long long *__restrict__ masks,
                                                const unsigned int *__restrict__ tile_off) {
    const int tile = blockIdx.x;
    const int s = tile_seg[tile];
    const int li = (tile - tile_first[s]) * WG + (int)threadIdx.x;
    const int w = threadIdx.x >> 6;
    const uint64_t b = masks[(size_t)tile * (WG / 64) + w];
    if (!((b >> (threadIdx.x & 63)) & 1ull)) return;
    unsigned int pos = tile_off[tile];
    for (int k = 0; k < w; k++) pos += (unsigned int)__popcll(masks[(size_t)tile * (WG / 64) + k]);
    pos += __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
    const int64_t src = seg_base[s] + li;
    const int64_t dst = seg_nbase[s] + pos;
#pragma unroll
    for (int p = 0; p < P; p++) nfreq[p * npitch + dst] = freq[p * pitch + src];
    gf_mutid id = ids[src];
    if (seg_preserve[s]) {
        int z = id.origin_threadID;
        id.origin_threadID = -(z < 0 ? -z : z);
    }
    nids[dst] = id;
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

__device__ __forceinline__ float mse_integrand(const MseSeg &S, float i) {
    float s = S.sa * i + S.sb;
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
        const float a = 1.0f - (float)id * step;
        const float f1 = mse_integrand(S, a);
        const float f2 = mse_integrand(S, a - step);
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
        const float i = ((float)id + 1.0f) / (float)Nc;
        const float j = ((float)(Nc - id) + 1.0f) / (float)Nc;
        const float sel = S.sa * i + S.sb;
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
