/*
 * gf_drift.hpp -- the GO_Fish drift kernel (gfx950) and the types it shares
 * with the host code.
 *
 * Layout in HBM (one "segment" = one simulation; the plain run_sim path is a
 * batch of one segment):
 *   freq  : float [P][cap]  structure of arrays, row p = population p
 *   ids   : gf_mutid [cap]  16 B per slot (origin gen, origin pop, threadID, 0)
 * A thread owns one mutation slot for all P populations, so migration is a
 * register-local dot product and the update is done in place.
 *
 * The advance kernels run up to KMAX generations per launch ("temporal
 * blocking"): each slot is read once, advanced through K generations of
 * migration, selection and drift in registers, and written once.  Mutations
 * born inside the block are materialised in their slots at the generation
 * they arise and drifted from the next generation on, exactly as the
 * generation-by-generation order of the reference (go_fish_impl.cuh:645-714)
 * prescribes.
 *
 * This header is also compiled into programs that pass their own selection
 * functor to GO_Fish::run_sim (include/go_fish.cuh): k_advance_fused and
 * k_mse_sel_table are then instantiated there with that functor, and every
 * floating-point function here keeps FMA contraction off itself, so the
 * results do not depend on the caller's -ffp-contract setting.
 */
#ifndef GF_DRIFT_HPP_
#define GF_DRIFT_HPP_

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

/* Selection evaluators: s(population, generation, frequency after
 * migration) for the row R of that generation.  LinearSel is the library's
 * own, the per-(population, generation) linear form the host lowers every
 * Sim_Model selection functor to; FunctorSel wraps a caller's functor with
 * the reference signature __device__ float(int population, int generation,
 * float freq) (template_inline_simulation_functors.cuh:35-37). */
struct LinearSel {
    template <typename Row>
    __device__ __forceinline__ float operator()(int p, int, float x, const Row &R) const {
        GF_STRICT_FP
        return R.sa[p] * x + R.sb[p];
    }
};

template <typename F>
struct FunctorSel {
    F f;
    template <typename Row>
    __device__ __forceinline__ float operator()(int p, int g, float x, const Row &) const { return f(p, g, x); }
};

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
 * the same block.  ncdf and etab are the workgroup's LDS copies of the
 * inverse normal's and exp(-m)'s segment tables (load_draw_tables). */
template <int P, typename Row, typename Sel>
__device__ __forceinline__ void generation_update(float *x, const Row &R, int g, bool refresh, u32x4 *rc,
                                                  uint32_t k0, uint32_t k1, uint32_t tid, int ogen, int opop,
                                                  const Sel &sel, const NcdfSeg *ncdf, const ExpSeg *etab) {
    GF_STRICT_FP
    float nx[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int Nc = R.Nc[p];
        if (refresh) rc[p] = drift_block(k0, k1, tid, ogen, opop, p, g);
        if (Nc <= 0) { nx[p] = 0.0f; continue; }
        /* one population: the migration weight is 1 (conservative
         * migration), and 0 + 1 * x == x */
        float im = P == 1 ? x[0] : 0.0f;
        if (P > 1) {
#pragma unroll
            for (int k = 0; k < P; k++) im = im + R.mig[k][p] * x[k];
        }
        float xs;
        if (R.neutral_mask & (1 << p)) {
            xs = im;   /* s == 0: Eq. 13 reduces exactly to x */
        } else {
            float s = sel(p, g, im, R);
            s = s > -1.0f ? s : -1.0f;
            /* Eq. 13: (x + s x (x + c1 (1-x))) / (1 + s x (x + c2 (1-x))) */
            const float om = 1.0f - im;
            const float sx = s * im;
            const float num = __builtin_fmaf(sx, __builtin_fmaf(R.c1[p], om, im), im);
            const float den = __builtin_fmaf(sx, __builtin_fmaf(R.c2[p], om, im), 1.0f);
            xs = num / den;
        }
        const float N = R.fN[p];
        const float mean = xs * N;
        const float var = (1.0f - xs) * mean;
        int j = binom_draw(lane_of(rc[p], g & 3), mean, var, N, ncdf, etab);
        j = min(max(j, 0), Nc);
        /* (float)j / N, correctly rounded: up to N = 2^22 in three float
         * operations with r = RN(1/N), the row's newfreq (quot_count),
         * above through double (invN); the test is on the integer Nc, a
         * scalar comparison */
        const float fj = (float)j;
        nx[p] = Nc <= 4194304 ? quot_count(fj, N, R.newfreq[p]) : (float)((double)fj * R.invN[p]);
    }
#pragma unroll
    for (int p = 0; p < P; p++) x[p] = nx[p];
}

/* copies the inverse normal's and exp(-m)'s segment tables into LDS (all
 * threads of the workgroup; the caller's next barrier publishes them) */
__device__ __forceinline__ void load_draw_tables(NcdfSeg *sh, ExpSeg *she) {
    const NcdfSeg *tab = ncdf_table();
    for (int k = threadIdx.x; k < GF_NCDF_SEGMENTS; k += WG) sh[k] = tab[k];
    const ExpSeg *etab = expneg_table();
    for (int k = threadIdx.x; k < GF_EXPNEG_SEGMENTS; k += WG) she[k] = etab[k];
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

/* segment of a tile: the last segment whose first tile is <= tile
 * (segments without tiles share their first tile with the next one) */
__device__ __forceinline__ int seg_of_tile(const int *__restrict__ tile_first, int nseg, int tile) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_first[mid] <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* The drift kernels ask for 8 waves per SIMD (amdgpu_waves_per_eu): the
 * compiler then keeps them within 96 SGPRs (spilling a few to VGPR lanes),
 * which measured 5 % faster than 7 waves with no spills (and 6 % faster
 * than 6 waves without spills for three populations). */

/* ---------------- fused advance + tile-local compaction ----------------
 * One launch advances every slot by K generations and writes the slots to
 * keep, in their order, packed at the start of the slot's own output tile
 * (tile i of the output holds its survivors in [256 i, 256 i + cnt_out[i])).
 * No tile waits for another.  k_tile_scan then turns the counts into the
 * exclusive prefix pre[] (the compacted index of each tile's first
 * survivor), the slot count, and for every tile j of the next launch the
 * input tile src0[j] that holds compacted index 256 j.  The next launch
 * reads its compacted slot li from input tile s, offset li - pre[s], with s
 * the last tile whose pre[s] <= li: the stable order of the reference's
 * compaction (go_fish_impl.cuh:370-429) without a separate scatter pass.
 * Samples and capacity growth first make the state dense again
 * (k_densify). */
struct FusedLaunch {
    const float *freq_in;
    const gf_mutid *ids_in;
    float *freq_out;
    gf_mutid *ids_out;
    int64_t pitch;
    const int *M_in;              /* compacted slots in the input */
    const int *pre_in;            /* fragmented input: exclusive prefix per input tile, [T_in + 1] */
    const int *src0_in;           /* fragmented input: first input tile of each tile of this launch */
    const int *T_in;              /* fragmented input: number of input tiles */
    unsigned int *cnt_out;        /* survivors per output tile */
    unsigned int *ticket;         /* tile tickets (several populations); zero before the launch */
    int *error;                   /* set if the host's tile bound was too small */
    int ntiles;                   /* tiles launched (upper bound of the needed ones) */
    int new_in_block;
    int dense;                    /* input is dense: compacted slot li is slot li */
    int compact;
    int preserve;
};

/* Slot of compacted index li (< M_in) in a fragmented input, for every lane
 * of one tile.  sh_pre is a 257-entry LDS window pre[s0 .. s0 + 256] filled
 * by the caller (INT_MAX past the end); s0 = the input tile holding the
 * tile's first compacted index, so sh_pre[0] <= li for every lane. */
__device__ __forceinline__ int64_t frag_slot(int li, int s0, const int *sh_pre, const int *__restrict__ pre, int T_in) {
    int lo = 0, hi = WG;                 /* last k with sh_pre[k] <= li */
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sh_pre[mid] <= li) lo = mid; else hi = mid - 1;
    }
    int s = s0 + lo;
    if (lo == WG) {                      /* more than 256 input tiles before li: search the rest */
        int l2 = s, h2 = T_in - 1;
        while (l2 < h2) {
            const int mid = (l2 + h2 + 1) >> 1;
            if (pre[mid] <= li) l2 = mid; else h2 = mid - 1;
        }
        s = l2;
        return (int64_t)s * WG + (li - pre[s]);
    }
    return (int64_t)s * WG + (li - sh_pre[lo]);
}

/* loads the LDS window of frag_slot for tile `tile` (all threads; ends with
 * a barrier) */
__device__ __forceinline__ int load_frag_window(int tile, const int *__restrict__ src0, const int *__restrict__ pre,
                                                int T_in, int *sh_pre) {
    const int s0 = src0[tile];
    for (int k = threadIdx.x; k <= WG; k += WG) sh_pre[k] = (s0 + k <= T_in) ? pre[s0 + k] : 0x7fffffff;
    __syncthreads();
    return s0;
}

template <int P, typename Sel>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8))) void k_advance_fused(const FusedLaunch F, const SegLaunch<P> S,
                                                                                          const Sel sel) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    __shared__ int sh_tile;
    __shared__ unsigned int sh_wcount[WG / 64];
    __shared__ int sh_pre[WG + 1];
    __shared__ unsigned long long sh_live[64];
    __shared__ int sh_diff[KMAX + 1];
    __shared__ NcdfSeg sh_ncdf[GF_NCDF_SEGMENTS];
    __shared__ ExpSeg sh_eneg[GF_EXPNEG_SEGMENTS];
    load_draw_tables(sh_ncdf, sh_eneg);   /* published by the tile loop's first barrier */
    if (threadIdx.x < 64) sh_live[threadIdx.x] = 0;
    if (threadIdx.x <= KMAX) sh_diff[threadIdx.x] = 0;
    const int M0 = *F.M_in;
    const int Mend = M0 + F.new_in_block;
    const int need_tiles = (Mend + WG - 1) / WG;
    const int T_in = F.dense ? 0 : *F.T_in;
    if (blockIdx.x == 0 && threadIdx.x == 0 && need_tiles > F.ntiles) *F.error = 2;   /* slots would be dropped */
    uint64_t cnt = 0;
    const GenRow<P> *rows = S.rows;
    /* Tile order.  One population: a fixed stride over the tiles (their
     * costs are alike, and a ticket's atomic round trip was measured to cost
     * more than the imbalance).  Several populations: an atomic ticket per
     * tile (the cost of a tile varies with the populations its slots hold). */
    const int last_tile = need_tiles < F.ntiles ? need_tiles : F.ntiles;
    for (int it = 0;; it++) {
        __syncthreads();
        int tile;
        if constexpr (P == 1) {
            tile = (int)blockIdx.x + it * (int)gridDim.x;
        } else {
            if (threadIdx.x == 0) sh_tile = (int)atomicAdd(F.ticket, 1u);
            __syncthreads();
            tile = sh_tile;
        }
        if (tile >= last_tile) break;
        const int li = tile * WG + (int)threadIdx.x;
        const bool valid = li < Mend;
        int s0 = 0;
        if (!F.dense && tile * WG < M0) s0 = load_frag_window(tile, F.src0_in, F.pre_in, T_in, sh_pre);
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
                const int64_t src = F.dense ? (int64_t)li : frag_slot(li, s0, sh_pre, F.pre_in, T_in);
#pragma unroll
                for (int p = 0; p < P; p++) x[p] = F.freq_in[p * F.pitch + src];
                id = F.ids_in[src];
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
        const auto *crows = as_const(rows);
        u32x4 rc[P];
        bool keep = false;
        int kfirst = KMAX, klast = -1;
        for (int t = 0; t < S.K; t++) {
            const auto &R = crows[t];
            const int g = S.g0 + 1 + t;
            if (valid && born < t) {
                const bool refresh = ((g & 3) == 0) || (t == born + 1);
                generation_update<P>(x, R, g, refresh, rc, S.k0, S.k1, tid, ogen, opop, sel, sh_ncdf, sh_eneg);
            }
            keep = valid && born <= t && keep_rule<P>(x, R.present_mask, preserved);
            if constexpr (P == 1) {
                /* one population: a slot is kept over one run of generations
                 * (lost and fixed are absorbing), recorded by its ends */
                if (keep) { kfirst = t < kfirst ? t : kfirst; klast = t; }
            } else {
                const uint64_t b = __ballot(keep);
                if (lane == t) cnt += (uint64_t)__popcll(b);
            }
        }
        if constexpr (P == 1) {
            /* per-generation live counts as a difference array: +1 at the
             * run's first generation, -1 after its last */
            if (klast >= 0) {
                atomicAdd(&sh_diff[kfirst], 1);
                if (klast + 1 < S.K) atomicAdd(&sh_diff[klast + 1], -1);
            }
        }
        /* without compaction every slot of the block stays */
        if (!F.compact) keep = valid;
        /* tile-local stable packing */
        const uint64_t kb = __ballot(keep);
        if (lane == 0) sh_wcount[w] = (unsigned int)__popcll(kb);
        __syncthreads();
        unsigned int pos = 0;
        for (int k = 0; k < w; k++) pos += sh_wcount[k];
        pos += __builtin_amdgcn_mbcnt_hi((uint32_t)(kb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)kb, 0u));
        if (keep) {
            const int64_t o = (int64_t)tile * WG + pos;
#pragma unroll
            for (int p = 0; p < P; p++) __builtin_nontemporal_store(x[p], &F.freq_out[p * F.pitch + o]);
            if (F.preserve) {
                const int z = id.origin_threadID;
                id.origin_threadID = -(z < 0 ? -z : z);
            }
            /* streamed out (the next launch reads them): non-temporal stores
             * leave the L2 to this launch's inputs */
            typedef int i4v __attribute__((ext_vector_type(4)));
            const i4v v = {id.origin_generation, id.origin_population, id.origin_threadID, id.reserved};
            __builtin_nontemporal_store(v, reinterpret_cast<i4v *>(&F.ids_out[o]));
        }
        if (threadIdx.x == 0) {
            unsigned int c = 0;
            for (int k = 0; k < WG / 64; k++) c += sh_wcount[k];
            F.cnt_out[tile] = c;
        }
    }
    /* per-generation live counts: lane t of every wave holds generation t */
    if constexpr (P == 1) {
        __syncthreads();
        if (threadIdx.x == 0) {   /* prefix of the difference array */
            long long run = 0;
            for (int t = 0; t < S.K; t++) {
                run += sh_diff[t];
                sh_live[t] = (unsigned long long)run;
            }
        }
    } else {
        if (cnt) atomicAdd(&sh_live[lane], (unsigned long long)cnt);
    }
    __syncthreads();
    if ((int)threadIdx.x < S.K && sh_live[threadIdx.x])
        atomicAdd(&S.live[(size_t)(S.t0 + threadIdx.x) * 64 + (blockIdx.x & 63)], sh_live[threadIdx.x]);
}

/* ---------------- selection at the equilibrium start ----------------
 * The frequencies at which the mutation-selection-equilibrium start
 * evaluates selection (generation 0 of the run): the end points a, a - 1/Nc
 * of trapezoid id (reference calculate_area / mse_integrand,
 * go_fish_impl.cuh:24-53) and the class frequency (id + 1) / Nc (reference
 * initialize_mse_frequency_array, :58-71), in the library's float order. */
__host__ __device__ __forceinline__ void mse_area_points(int id, int Nc, float &a, float &b) {
    GF_STRICT_FP
    const float step = 1.0f / (float)Nc;
    a = 1.0f - (float)id * step;
    b = a - step;
}

__host__ __device__ __forceinline__ float mse_class_point(int id, int Nc) {
    GF_STRICT_FP
    return ((float)id + 1.0f) / (float)Nc;
}

/* Table of a caller's selection functor at those points, for one
 * population with Nc chromosomes:
 *   tab[id]          = s(a_id),        id in [0, Nc)
 *   tab[Nc + id]     = s(a_id - 1/Nc), id in [0, Nc)
 *   tab[2 Nc + id]   = s((id + 1)/Nc), id in [0, Nc - 1) */
template <typename F>
__global__ void k_mse_sel_table(const F f, int pop, int gen, int Nc, float *__restrict__ tab) {
    for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < Nc; id += gridDim.x * blockDim.x) {
        float a, b;
        mse_area_points(id, Nc, a, b);
        tab[id] = f(pop, gen, a);
        tab[Nc + id] = f(pop, gen, b);
        if (id < Nc - 1) tab[2 * Nc + id] = f(pop, gen, mse_class_point(id, Nc));
    }
}

} /* namespace gf */

#endif /* GF_DRIFT_HPP_ */
