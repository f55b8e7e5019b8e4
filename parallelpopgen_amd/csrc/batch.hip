/*
 * batch.hip -- batched simulation: thousands of independent single-
 * population simulations advanced together (no reference counterpart; it
 * replaces the back-to-back run_sim + site_frequency_spectrum loop of
 * examples/example_dadi/run.cu:83-94, which recreates streams and buffers
 * and synchronises the host several times per simulation).
 *
 * Layout.  Two state buffers (frequencies + mutation IDs); simulation s owns
 * the tiles [t0_s, t0_s + nt_s) of both (256 slots per tile), so every tile
 * belongs to one simulation and its parameters are wave-uniform scalar
 * loads.  Each simulation advances on its own: its generation counter, the
 * buffer holding its current state and that state's layout (dense, or
 * fragmented: tile j holds its survivors packed at its start, described by
 * the prefix pre_s and the source-tile table src0_s) live in a device
 * array of BSim records.
 *
 * One block (up to 32 generations) is four kernels over every simulation,
 * with no host round trip:
 *   k_bt_rows     per simulation: the block's generation rows and new-
 *                 mutation counts; a simulation whose survivors and new
 *                 mutations would not fit its tiles stalls (its state stays)
 *   k_bt_advance  the drift kernel with tile-local compaction
 *                 (migration_selection_drift + add_new_mutations +
 *                 compaction, as k_advance_fused): reads the current
 *                 buffer, writes each tile's survivors packed into the other
 *   k_bt_scan     per simulation: prefix of its tiles' survivor counts (the
 *                 next block's gather table), new slot count, generation
 *   k_bt_status   active / stalled counts and the live mutation-generations
 *                 into pinned host memory
 * The block is captured once in a hipGraph and replayed; the host looks at
 * the status every few replays.  Stalled simulations get larger regions
 * (the host re-lays out the live ones, densely, into new buffers, which also
 * drops finished simulations once they hold most of the tiles).  Finished
 * simulations are finalised (SFS of n chromosomes, optionally the state)
 * from wherever their last block left them.
 *
 * Block length: a small online autotuner measures the batch's own
 * throughput (live mutation-generations per second) with several lengths
 * on the first replays and keeps the best for the rest of the batch and
 * for later batches of the same shape.  Results do not depend on it: draws
 * are keyed by (seed, generation, mutation ID), so every simulation's
 * result is bit for bit the result gf_run gives for it alone.
 */
#include "gf_kernels.hpp"
#include "gf_runtime.hpp"
#include "gf_spectrum.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

struct gf_batch_result {
    int32_t S = 0, n = 0;
    std::vector<float> sfs;                  /* [S][n] */
    std::vector<int32_t> num_mutations;      /* final compacted count */
    std::vector<std::vector<float>> freq;    /* optional final state */
    std::vector<std::vector<gf_mutid>> ids;
    int64_t live_mutation_generations = 0;
    gf_timing timing{};
    int32_t generations_per_block = 0;       /* chosen by the autotuner (or the caller's cap) */
    int32_t blocks = 0;                      /* graph replays */
    int32_t relayouts = 0;
};

namespace gf {

constexpr int BK_MAX = 32;                   /* longest block the autotuner tries */
static float g_region_scale = 1.0f;          /* gf_set_batch_region_scale */
enum { BS_ACTIVE = 0, BS_DONE = 1, BS_STALLED = 2 };

/* one simulation of a batch, device-resident */
struct alignas(16) BSim {
    /* model */
    double invN;
    int Nc, N, G, neutral;
    float F, h, mu, s, L, c1, c2, newfreq, fN;
    uint32_t k0, k1;
    /* layout: tiles [t0, t0 + nt) of both buffers, prefix entries [p0, p0 + nt] */
    int t0, nt, p0;
    /* progress */
    int g;        /* generations done */
    int M;        /* slots (all segregating) in the current buffer */
    int T;        /* tiles of the current fragmented state; -1: dense */
    int cur;      /* buffer of the current state */
    int status;   /* BS_* */
    /* this block (k_bt_rows) */
    int Kb, nnew, tiles;
    int need;     /* tiles a stalled simulation's block needs */
};

/* device pointers of a layout (kernel argument, fixed while a graph lives) */
struct BBufs {
    float *freq[2];
    gf_mutid *ids[2];
    int *pre[2];                  /* per simulation nt + 1 entries from p0 */
    int *src0[2];                 /* per simulation nt entries from t0 */
    unsigned int *cnt;            /* survivors per output tile */
    GenRow<1> *rows;              /* [S][BK_MAX] */
    const int *tile_sim;          /* tile -> simulation */
    unsigned long long *live;     /* 64 partial sums of live mutation-generations */
    int *K;                       /* block length (device word: the tuner changes it between replays) */
    int ntiles, S;
};

/* a BSim read through scalar loads (uniform per tile) */
__device__ __forceinline__ BSim load_bsim(const BSim *p) {
    const auto &c = *as_const(p);
    BSim b;
    b.invN = c.invN; b.Nc = c.Nc; b.N = c.N; b.G = c.G; b.neutral = c.neutral;
    b.F = c.F; b.h = c.h; b.mu = c.mu; b.s = c.s; b.L = c.L; b.c1 = c.c1; b.c2 = c.c2;
    b.newfreq = c.newfreq; b.fN = c.fN; b.k0 = c.k0; b.k1 = c.k1;
    b.t0 = c.t0; b.nt = c.nt; b.p0 = c.p0; b.g = c.g; b.M = c.M; b.T = c.T; b.cur = c.cur;
    b.status = c.status; b.Kb = c.Kb; b.nnew = c.nnew; b.tiles = c.tiles;
    return b;
}

/* ---- 1. rows and new-mutation counts of the next block, one thread per simulation ---- */
static __global__ void k_bt_rows(BSim *__restrict__ sims, BBufs B) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B.S) return;
    BSim &b = sims[s];
    b.Kb = 0;
    b.tiles = 0;
    if (b.status != BS_ACTIVE || b.g >= b.G) return;
    const int K = min(*B.K, b.G - b.g);
    GenRow<1> *rows = B.rows + (size_t)s * BK_MAX;
    int cum = 0;
    for (int t = 0; t < K; t++) {
        const int g = b.g + 1 + t;
        GenRow<1> r;
        r.Nc[0] = b.Nc;
        r.fN[0] = b.fN;
        r.invN[0] = b.invN;
        r.sa[0] = 0.0f;
        r.sb[0] = b.s;
        r.c1[0] = b.c1;
        r.c2[0] = b.c2;
        r.mig[0][0] = 1.0f;
        r.newfreq[0] = b.newfreq;
        r.newcum[0] = cum;
        cum += new_mutation_count(b.Nc, b.mu, b.L, b.k0, b.k1, g, 0);
        r.newcum[1] = cum;
        r.present_mask = b.N > 0 ? 1 : 0;
        r.neutral_mask = b.neutral ? 1 : 0;
        rows[t] = r;
    }
    const int64_t need = (int64_t)b.M + cum;
    const int tiles = (int)((need + WG - 1) / WG);
    if (tiles > b.nt) {   /* does not fit its region: wait for the host */
        b.status = BS_STALLED;
        b.need = tiles;
        return;
    }
    b.Kb = K;
    b.nnew = cum;
    b.tiles = tiles;
}

/* ---- 2. drift + new mutations + tile-local compaction of every simulation's block ---- */
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8))) void k_bt_advance(const BSim *__restrict__ sims,
                                                                                          const BBufs B) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    __shared__ unsigned int sh_wcount[WG / 64];
    __shared__ int sh_pre[WG + 1];
    __shared__ NcdfSeg sh_ncdf[GF_NCDF_SEGMENTS];
    __shared__ ExpSeg sh_eneg[GF_EXPNEG_SEGMENTS];
    load_draw_tables(sh_ncdf, sh_eneg);   /* published by the tile loop's first barrier */
    uint64_t live = 0;          /* this thread's slots' live mutation-generations */
    for (int gt = blockIdx.x;; gt += gridDim.x) {
        __syncthreads();
        if (gt >= B.ntiles) break;
        const int s = *as_const(B.tile_sim + gt);
        const BSim S = load_bsim(sims + s);
        const int j = gt - S.t0;
        if (j >= S.tiles) continue;
        /* the two buffers by selects (indexing the argument's arrays with
         * a run-time value would put them in scratch memory) */
        const bool c1 = S.cur != 0;
        const float *freq_in = c1 ? B.freq[1] : B.freq[0];
        float *freq_out = c1 ? B.freq[0] : B.freq[1];
        const gf_mutid *ids_in = c1 ? B.ids[1] : B.ids[0];
        gf_mutid *ids_out = c1 ? B.ids[0] : B.ids[1];
        const int *pre_in = (c1 ? B.pre[1] : B.pre[0]) + S.p0;
        const int *src0_in = (c1 ? B.src0[1] : B.src0[0]) + S.t0;
        const int64_t base = (int64_t)S.t0 * WG;
        const int M0 = S.M, Mend = S.M + S.nnew;
        const bool dense = S.T < 0;
        const int li = j * WG + (int)threadIdx.x;
        const bool valid = li < Mend;
        const GenRow<1> *rows = B.rows + (size_t)s * BK_MAX;
        int s0 = 0;
        if (!dense && j * WG < M0) s0 = load_frag_window(j, src0_in, pre_in, S.T, sh_pre);
        float x = 0.0f;
        int ogen = 0, opop = 0, born = -1;
        uint32_t tid = 0;
        gf_mutid id;
        id.origin_generation = 0; id.origin_population = 0; id.origin_threadID = 0; id.reserved = 0;
        if (valid) {
            if (li < M0) {
                const int64_t src = base + (dense ? (int64_t)li : frag_slot(li, s0, sh_pre, pre_in, S.T));
                x = freq_in[src];
                id = ids_in[src];
                ogen = id.origin_generation;
                opop = id.origin_population;
                tid = (uint32_t)(id.origin_threadID < 0 ? -id.origin_threadID : id.origin_threadID);
            } else {
                const int rel = li - M0;
                int tb, pb;
                find_birth<1>(rows, S.Kb, rel, tb, pb);
                born = tb;
                ogen = S.g + 1 + tb;
                opop = pb;
                tid = (uint32_t)(rel - rows[tb].newcum[pb] + 1);
                x = rows[tb].newfreq[0];
                id.origin_generation = ogen;
                id.origin_population = opop;
                id.origin_threadID = (int32_t)tid;
                id.reserved = 0;
            }
        }
        const auto *crows = as_const(rows);
        u32x4 rc[1];
        bool keep = false;
        int kfirst = KMAX, klast = -1;
        for (int t = 0; t < S.Kb; t++) {
            const auto &R = crows[t];
            const int g = S.g + 1 + t;
            if (valid && born < t) {
                const bool refresh = ((g & 3) == 0) || (t == born + 1);
                generation_update<1>(&x, R, g, refresh, rc, S.k0, S.k1, tid, ogen, opop, LinearSel(), sh_ncdf, sh_eneg);
            }
            keep = valid && born <= t && keep_rule<1>(&x, R.present_mask, false);
            /* one population: kept over one run of generations (lost and
             * fixed are absorbing) */
            if (keep) { kfirst = t < kfirst ? t : kfirst; klast = t; }
        }
        if (klast >= 0) live += (uint64_t)(klast - kfirst + 1);
        /* tile-local stable packing into the other buffer */
        const uint64_t kb = __ballot(keep);
        if (lane == 0) sh_wcount[w] = (unsigned int)__popcll(kb);
        __syncthreads();
        if (keep) {
            unsigned int pos = 0;
            for (int k = 0; k < w; k++) pos += sh_wcount[k];
            pos += __builtin_amdgcn_mbcnt_hi((uint32_t)(kb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)kb, 0u));
            const int64_t o = base + (int64_t)j * WG + pos;
            __builtin_nontemporal_store(x, &freq_out[o]);
            typedef int i4v __attribute__((ext_vector_type(4)));
            const i4v v = {id.origin_generation, id.origin_population, id.origin_threadID, id.reserved};
            __builtin_nontemporal_store(v, reinterpret_cast<i4v *>(&ids_out[o]));
        }
        if (threadIdx.x == 0) {
            unsigned int c = 0;
            for (int k = 0; k < WG / 64; k++) c += sh_wcount[k];
            B.cnt[gt] = c;
        }
    }
    /* one atomic per wave, spread over 64 counters */
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) live += __shfl_xor(live, off);
    if (lane == 0 && live) atomicAdd(&B.live[(blockIdx.x * (WG / 64) + w) & 63], (unsigned long long)live);
}

/* ---- 3. per simulation: prefix of the block's survivor counts ---- */
static __global__ __launch_bounds__(256) void k_bt_scan(BSim *__restrict__ sims, const BBufs B) {
    const int s = blockIdx.x;
    __shared__ int sh_w[4];
    const int Kb = sims[s].Kb;
    if (Kb == 0) return;   /* the whole workgroup */
    const int T = sims[s].tiles, t0 = sims[s].t0, p0 = sims[s].p0;
    const int out = sims[s].cur ^ 1;
    const unsigned int *cnt = B.cnt + t0;
    int *pre = (out ? B.pre[1] : B.pre[0]) + p0;
    int *src0 = (out ? B.src0[1] : B.src0[0]) + t0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int carry = 0;
    for (int base = 0; base < T; base += 256 * 4) {
        const int b = base + (int)threadIdx.x * 4;
        int v[4], sum = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v[i] = b + i < T ? (int)cnt[b + i] : 0;
            sum += v[i];
        }
        int x = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) sh_w[w] = x;
        __syncthreads();
        int before = 0, round = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            before += k < w ? sh_w[k] : 0;
            round += sh_w[k];
        }
        int run = carry + before + x - sum;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int idx = b + i;
            if (idx < T) {
                pre[idx] = run;
                const int c = v[i];
                for (int jn = (run + WG - 1) / WG; jn * WG <= run + c - 1; jn++) src0[jn] = idx;
                run += c;
            }
        }
        carry += round;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pre[T] = carry;
        BSim &r = sims[s];
        r.M = carry;
        r.T = T;
        r.cur = out;
        r.g += Kb;
        if (r.g >= r.G) r.status = BS_DONE;
    }
}

/* ---- 4. status for the host: [active, stalled], live mutation-generations ---- */
static __global__ __launch_bounds__(1024) void k_bt_status(const BSim *__restrict__ sims, const BBufs B,
                                                           int *__restrict__ host) {
    __shared__ int sh_a[16], sh_st[16];
    int a = 0, st = 0;
    for (int s = threadIdx.x; s < B.S; s += blockDim.x) {
        const int v = sims[s].status;
        a += v == BS_ACTIVE;
        st += v == BS_STALLED;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        st += __shfl_xor(st, off);
    }
    if ((threadIdx.x & 63) == 0) { sh_a[threadIdx.x >> 6] = a; sh_st[threadIdx.x >> 6] = st; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ta = 0, ts = 0;
        for (int k = 0; k < (int)blockDim.x / 64; k++) { ta += sh_a[k]; ts += sh_st[k]; }
        unsigned long long l = 0;
        for (int k = 0; k < 64; k++) l += B.live[k];
        host[0] = ta;
        host[1] = ts;
        *reinterpret_cast<unsigned long long *>(host + 2) = l;   /* pinned host words: read after the replay's event */
    }
}

/* ---- re-layout: dense copy of the listed simulations' current states into new buffers ---- */
struct BMove {
    int64_t src_base;     /* first slot of the old region */
    int64_t dst_base;     /* first slot of the new region */
    const int *pre;       /* old fragmented layout (NULL: dense) */
    const int *src0;
    int T, M, cur, tile_first;   /* tile_first: the move's first tile in the copy grid */
};

static __global__ __launch_bounds__(WG) void k_bt_move(const BMove *__restrict__ mv, int nmv, int ntiles,
                                                       float *const *__restrict__ freq_old,
                                                       gf_mutid *const *__restrict__ ids_old,
                                                       float *__restrict__ freq_new, gf_mutid *__restrict__ ids_new) {
    __shared__ int sh_pre[WG + 1];
    for (int gt = blockIdx.x; gt < ntiles; gt += gridDim.x) {
        __syncthreads();
        int lo = 0, hi = nmv - 1;   /* last move whose first tile <= gt */
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (mv[mid].tile_first <= gt) lo = mid; else hi = mid - 1;
        }
        const BMove m = mv[lo];
        const int j = gt - m.tile_first;
        const int li = j * WG + (int)threadIdx.x;
        int s0 = 0;
        if (m.pre && j * WG < m.M) s0 = load_frag_window(j, m.src0, m.pre, m.T, sh_pre);
        if (li < m.M) {
            const int64_t src = m.src_base + (m.pre ? frag_slot(li, s0, sh_pre, m.pre, m.T) : (int64_t)li);
            freq_new[m.dst_base + li] = freq_old[m.cur][src];
            ids_new[m.dst_base + li] = ids_old[m.cur][src];
        }
    }
}

/* ---- finalisation: SFS (and optionally a dense copy) of finished simulations ---- */
struct BFin {
    int64_t base;          /* first slot of the region */
    const int *pre;        /* fragmented layout (NULL: dense) */
    const unsigned int *cnt;
    int T, M, cur, tile_first;
    int64_t out_base;      /* first slot of the dense copy */
};

template <bool LDS>
static __global__ __launch_bounds__(WG) void k_bt_final(const BFin *__restrict__ fin, int nfin, int ntiles,
                                                        float *const *__restrict__ freq, gf_mutid *const *__restrict__ ids,
                                                        int n, int half_n, const float *__restrict__ lc,
                                                        unsigned long long *__restrict__ acc,
                                                        float *__restrict__ copy_freq, gf_mutid *__restrict__ copy_ids) {
    __shared__ unsigned long long sh[LDS ? SFS_LDS_BINS : 1];
    for (int gt = blockIdx.x; gt < ntiles; gt += gridDim.x) {
        int lo = 0, hi = nfin - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (fin[mid].tile_first <= gt) lo = mid; else hi = mid - 1;
        }
        const BFin f = fin[lo];
        const int j = gt - f.tile_first;
        unsigned long long *out = acc + (size_t)lo * (n + 2);
        __syncthreads();
        if (LDS) {
            for (int k = threadIdx.x; k < n + 2; k += WG) sh[k] = 0;
            __syncthreads();
        }
        /* slot of this lane and its compacted index */
        bool has;
        int ci;
        if (f.pre) {
            has = j < f.T && (unsigned int)threadIdx.x < f.cnt[j];
            ci = has ? f.pre[j] + (int)threadIdx.x : 0;
        } else {
            ci = j * WG + (int)threadIdx.x;
            has = ci < f.M;
        }
        if (has) {
            const int64_t slot = f.base + (int64_t)j * WG + threadIdx.x;
            const float x = freq[f.cur][slot];
            sfs_add_mutation<LDS>(x, n, half_n, lc, sh, out);
            if (copy_freq) {
                copy_freq[f.out_base + ci] = x;
                copy_ids[f.out_base + ci] = ids[f.cur][slot];
            }
        }
        if (LDS) {
            __syncthreads();
            for (int k = threadIdx.x; k < n + 2; k += WG)
                if (sh[k]) atomicAdd(&out[k], sh[k]);
        }
    }
}

/* ---------------- block-length autotuner ----------------
 * Candidate block lengths are tried in turn, two rounds, each on one window
 * of replays of a batch, and scored by the live mutation-generations per
 * second of its windows; after the last window the best is kept.  The
 * state is per batch shape (number of simulations, initial slots and
 * longest run, each to a power of two, and the caller's cap), so short
 * batches of one shape share one tuning and later batches use its result. */
class BlockTuner {
public:
    static BlockTuner &get() {
        static BlockTuner *t = new BlockTuner();
        return *t;
    }
    static constexpr int kNumCand = 5;
    static constexpr int kRounds = 2;
    static int cand(int i) {
        static const int c[kNumCand] = {8, 12, 16, 24, 32};
        return c[i];
    }
    typedef std::tuple<int, int, int, int> Key;
    static Key key(int S, int64_t slots, int maxG, int cap) {
        auto lg = [](int64_t v) { int k = 0; while (v > 1) { v >>= 1; k++; } return k; };
        return Key(lg(S), lg(slots), lg(maxG), cap);
    }
    /* block length of the next window; `measuring`: its time counts */
    int next(const Key &k, int cap, bool &measuring) {
        std::lock_guard<std::mutex> l(mu_);
        State &s = st_[k];
        while (!s.done && s.window < kRounds * kNumCand && cand(s.window % kNumCand) > cap) s.window++;
        if (!s.done && s.window >= kRounds * kNumCand) finish(s, cap);
        measuring = !s.done;
        return s.done ? std::min(s.best, cap) : cand(s.window % kNumCand);
    }
    void record(const Key &k, int K, double work, double ms) {
        std::lock_guard<std::mutex> l(mu_);
        State &s = st_[k];
        if (s.done) return;
        s.work[K] += work;
        s.ms[K] += ms;
        s.window++;
    }

private:
    struct State {
        int window = 0;
        bool done = false;
        int best = 16;
        std::map<int, double> work, ms;
    };
    void finish(State &s, int cap) {
        double best = -1.0;
        s.best = std::min(16, cap);
        for (auto &kv : s.work) {
            const double t = s.ms[kv.first];
            const double rate = t > 0 ? kv.second / t : 0.0;
            if (rate > best) { best = rate; s.best = kv.first; }
        }
        s.done = true;
    }
    std::mutex mu_;
    std::map<Key, State> st_;
};

struct BatchRun {
    const gf_batch_sim *in;
    int S, interval, n;
    bool keep;
    gf_batch_result *R;
    OwnedStream st;   /* before the buffers: they are released while it is alive */
    std::vector<BSim> host;            /* host copy of the records (valid after sync_sims) */
    DevBuf<BSim> d_sims;
    DevBuf<float> freq[2];
    DevBuf<gf_mutid> ids[2];
    DevBuf<int> pre[2], src0[2], tile_sim, d_K;
    DevBuf<unsigned int> cnt;
    DevBuf<GenRow<1>> rows;
    DevBuf<unsigned long long> live;
    DevBuf<float *> d_fptr;
    DevBuf<gf_mutid *> d_iptr;
    PinBuf<int> h_status;              /* [active, stalled, live (2 words)] */
    std::vector<int> layout;           /* simulations in the current layout, in tile order */
    std::vector<char> finalised;
    int64_t ntiles = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int relayouts = 0, blocks = 0;

    ~BatchRun() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }

    /* tiles for a simulation holding M slots that may gain K generations of
     * new mutations (7 standard deviations over their mean), with a quarter
     * of headroom for the survivors */
    static int tiles_for(int64_t M, double lam, int K) {
        const double need = 1.25 * (double)M + lam * K + 8.0 * std::sqrt(lam * K + 1.0) + 256.0;
        return (int)std::min<double>(std::ceil(need / WG), 2e9 / WG);
    }

    int sync_sims() {
        GF_HIP(hipMemcpyAsync(host.data(), d_sims.p, (size_t)S * sizeof(BSim), hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        return GF_OK;
    }

    /* ----- mutation-selection equilibrium of every simulation into buffer 0, dense ----- */
    int init(int Kmax) {
        host.assign(S, BSim{});
        for (int s = 0; s < S; s++) {
            const gf_batch_sim &b = in[s];
            BSim &d = host[s];
            const int N1 = b.N1 > 0 ? b.N1 : b.N;
            d.N = N1; d.F = b.F; d.h = b.h; d.mu = b.mu; d.s = b.s; d.L = b.num_sites;
            d.G = b.num_generations;
            d.k0 = (uint32_t)b.seed1; d.k1 = (uint32_t)b.seed2;
            d.Nc = N1 > 0 ? nchrom(N1, b.F) : 0;
            d.fN = (float)d.Nc;
            d.invN = d.Nc > 0 ? 1.0 / (double)d.fN : 0.0;
            d.c1 = (b.F + b.h) - (b.h * b.F);
            d.c2 = (b.F + (2.0f * b.h)) - ((2.0f * b.h) * b.F);
            d.newfreq = d.Nc > 0 ? 1.0f / (float)d.Nc : 0.0f;
            d.neutral = b.s == 0.0f;
            d.T = -1;
            d.status = d.G > 0 ? BS_ACTIVE : BS_DONE;
            if (b.N < 0 || N1 < 0 || b.F < 0 || b.F > 1 || b.mu < 0 || b.num_generations < 0) {
                set_error("gf_batch_run: simulation " + std::to_string(s) + " has invalid parameters");
                return GF_ERR_PARAM;
            }
        }
        std::vector<MseSeg> segs;
        std::vector<int> seg_sim;
        int64_t coff = 0, aoff = 0, choff = 0;
        bool any_sel = false;
        for (int s = 0; s < S; s++) {
            const int Nc0 = in[s].N > 0 ? nchrom(in[s].N, in[s].F) : 0;
            if (Nc0 <= 1) continue;
            MseSeg m{};
            m.Nc = Nc0; m.Nind = in[s].N; m.F = in[s].F; m.h = in[s].h; m.mu = in[s].mu;
            m.sa = 0.0f; m.sb = in[s].s; m.L = in[s].num_sites;
            m.k0 = host[s].k0; m.k1 = host[s].k1; m.pop = 0;
            m.neutral = host[s].neutral; m.preserve = 0;
            m.class_off = coff; m.area_off = aoff; m.chunk_off = choff;
            coff += m.Nc - 1; aoff += m.Nc; choff += (m.Nc + MSE_CHUNK - 1) / MSE_CHUNK;
            if (!m.neutral) any_sel = true;
            segs.push_back(m);
            seg_sim.push_back(s);
        }
        std::vector<int64_t> Minit(S, 0);
        const int ns = (int)segs.size();
        std::vector<int64_t> scan;
        DevBuf<MseSeg> d_segs;
        DevBuf<int64_t> d_offs, d_scan, d_scratch, d_total;
        const unsigned grid = (unsigned)std::max(1, num_cus() * 8);
        if (ns > 0) {
            std::vector<int64_t> offs(3 * (size_t)ns);
            for (int k = 0; k < ns; k++) { offs[k] = segs[k].class_off; offs[ns + k] = segs[k].area_off; offs[2 * ns + k] = segs[k].chunk_off; }
            GF_TRY(d_segs.alloc(ns));
            GF_TRY(d_offs.alloc(3 * (size_t)ns));
            GF_HIP(hipMemcpyAsync(d_segs.p, segs.data(), ns * sizeof(MseSeg), hipMemcpyHostToDevice, st));
            GF_HIP(hipMemcpyAsync(d_offs.p, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
            DevBuf<float> d_area, d_tot;
            if (any_sel) {
                GF_TRY(d_area.alloc((size_t)aoff));
                GF_TRY(d_tot.alloc((size_t)choff));
                hipLaunchKernelGGL(k_mse_area, dim3(grid), dim3(256), 0, st, d_segs.p, d_offs.p + ns, ns, aoff, d_area.p);
                hipLaunchKernelGGL(k_mse_chunk_local, dim3(grid), dim3(256), 0, st, d_segs.p, d_offs.p + 2 * ns, ns, choff, d_area.p, d_tot.p);
                hipLaunchKernelGGL(k_mse_chunk_prefix, dim3((ns + 63) / 64), dim3(64), 0, st, d_segs.p, ns, d_tot.p);
                hipLaunchKernelGGL(k_mse_chunk_add, dim3(grid), dim3(256), 0, st, d_segs.p, d_offs.p + ns, ns, aoff, d_area.p, d_tot.p);
                GF_HIP(hipGetLastError());
            }
            DevBuf<int> d_counts;
            GF_TRY(d_counts.alloc((size_t)coff));
            GF_TRY(d_scan.alloc((size_t)coff + 1));
            GF_TRY(d_total.alloc(1));
            hipLaunchKernelGGL(k_mse_counts, dim3(grid), dim3(256), 0, st, d_segs.p, d_offs.p, ns, coff,
                               any_sel ? d_area.p : nullptr, d_counts.p);
            GF_HIP(hipGetLastError());
            GF_TRY(scan_counts(d_counts.p, coff, d_scan.p, d_total.p, st, d_scratch));
            GF_HIP(hipMemcpyAsync(d_scan.p + coff, d_total.p, sizeof(int64_t), hipMemcpyDeviceToDevice, st));
            scan.resize((size_t)coff + 1);
            GF_HIP(hipMemcpyAsync(scan.data(), d_scan.p, scan.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st));
            GF_HIP(hipStreamSynchronize(st));
            for (int k = 0; k < ns; k++) {
                const int64_t tot = scan[segs[k].class_off + segs[k].Nc - 1] - scan[segs[k].class_off];
                if (tot > 2000000000LL) { set_error("gf_batch_run: initial mutation count exceeds 2e9"); return GF_ERR_UNSUPPORTED; }
                Minit[seg_sim[k]] = tot;
            }
        }
        /* the layout: every simulation with generations to run, its tiles
         * sized for its equilibrium count and a block of new mutations */
        layout.clear();
        int64_t t = 0;
        int64_t pcount = 0;
        for (int s = 0; s < S; s++) {
            BSim &d = host[s];
            d.M = (int)Minit[s];
            d.t0 = (int)t;
            d.p0 = (int)(t + pcount);
            const double lam = (double)d.mu * (double)d.Nc * (double)d.L;
            /* a population that grows after generation 0 keeps more of its
             * new mutations for a while: room for ~4 ln(G) generations more */
            const int grow = d.N > in[s].N ? (int)std::ceil(4.0 * std::log1p((double)d.G)) : 0;
            const int nt = (int)std::ceil((double)g_region_scale * tiles_for(d.M, lam, Kmax + grow));
            d.nt = std::max(nt, (int)((d.M + WG - 1) / WG));
            t += d.nt;
            pcount++;
            layout.push_back(s);
            if (t > 2000000000LL / WG) { set_error("gf_batch_run: more than 2e9 slots"); return GF_ERR_UNSUPPORTED; }
        }
        GF_TRY(allocate_layout(t));
        /* host arrays given to asynchronous copies live until the
         * synchronisation at the end */
        std::vector<int64_t> out_offs;
        if (ns > 0) {
            out_offs.resize(ns);
            for (int k = 0; k < ns; k++) {
                const int s = seg_sim[k];
                segs[k].slot_base = (int64_t)host[s].t0 * WG;
                segs[k].sim_base = segs[k].slot_base;
                out_offs[k] = scan[segs[k].class_off];
            }
            GF_HIP(hipMemcpyAsync(d_segs.p, segs.data(), ns * sizeof(MseSeg), hipMemcpyHostToDevice, st));
            GF_HIP(hipMemcpyAsync(d_offs.p, out_offs.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, st));
            const int64_t total = scan[coff];
            if (total > 0)
                hipLaunchKernelGGL(k_mse_fill<1>, dim3(grid), dim3(256), 0, st, d_segs.p, d_offs.p, ns, total, d_scan.p,
                                   freq[0].p, (int64_t)freq[0].n, ids[0].p);
            GF_HIP(hipGetLastError());
        }
        GF_HIP(hipMemcpyAsync(d_sims.p, host.data(), (size_t)S * sizeof(BSim), hipMemcpyHostToDevice, st));
        GF_HIP(hipStreamSynchronize(st));
        return GF_OK;
    }

    /* buffers and tables of a layout of `t` tiles (host[] holds t0/nt/p0) */
    int allocate_layout(int64_t t) {
        ntiles = t;
        const size_t slots = (size_t)std::max<int64_t>(t, 1) * WG;
        for (int b = 0; b < 2; b++) {
            GF_TRY(freq[b].alloc(slots));
            GF_TRY(ids[b].alloc(slots));
            GF_TRY(pre[b].alloc((size_t)t + layout.size() + 1));
            GF_TRY(src0[b].alloc((size_t)std::max<int64_t>(t, 1)));
        }
        GF_TRY(cnt.alloc((size_t)std::max<int64_t>(t, 1)));
        std::vector<int> ts((size_t)std::max<int64_t>(t, 1), 0);
        for (int s : layout)
            for (int j = 0; j < host[s].nt; j++) ts[(size_t)host[s].t0 + j] = s;
        GF_TRY(tile_sim.alloc(ts.size()));
        GF_HIP(hipMemcpyAsync(tile_sim.p, ts.data(), ts.size() * sizeof(int), hipMemcpyHostToDevice, st));
        float *fp[2] = {freq[0].p, freq[1].p};
        gf_mutid *ip[2] = {ids[0].p, ids[1].p};
        GF_TRY(d_fptr.alloc(2));
        GF_TRY(d_iptr.alloc(2));
        GF_HIP(hipMemcpyAsync(d_fptr.p, fp, sizeof fp, hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(d_iptr.p, ip, sizeof ip, hipMemcpyHostToDevice, st));
        GF_HIP(hipStreamSynchronize(st));   /* the host arrays above are locals */
        return GF_OK;
    }

    BBufs bufs() const {
        BBufs B;
        for (int b = 0; b < 2; b++) {
            B.freq[b] = freq[b].p; B.ids[b] = ids[b].p; B.pre[b] = pre[b].p; B.src0[b] = src0[b].p;
        }
        B.cnt = cnt.p;
        B.rows = rows.p;
        B.tile_sim = tile_sim.p;
        B.live = live.p;
        B.K = d_K.p;
        B.ntiles = (int)ntiles;
        B.S = S;
        return B;
    }

    /* one block = rows, advance, scan, status: captured once per layout */
    int capture() {
        if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
        const BBufs B = bufs();
        const int grid = std::max(1, std::min<int>((int)ntiles, num_cus() * 8));
        GF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipLaunchKernelGGL(k_bt_rows, dim3((S + 127) / 128), dim3(128), 0, st, d_sims.p, B);
        hipLaunchKernelGGL(k_bt_advance, dim3(grid), dim3(WG), 0, st, (const BSim *)d_sims.p, B);
        hipLaunchKernelGGL(k_bt_scan, dim3(S), dim3(256), 0, st, d_sims.p, B);
        hipLaunchKernelGGL(k_bt_status, dim3(1), dim3(1024), 0, st, (const BSim *)d_sims.p, B, h_status.p);
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess) {
            set_error(std::string("gf_batch_run: graph capture failed: ") + hipGetErrorString(e));
            return GF_ERR_DEVICE;
        }
        GF_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        return GF_OK;
    }

    /* finalise the listed simulations (all finished) from the current buffers */
    int finalise(const std::vector<int> &done) {
        const int nd = (int)done.size();
        if (nd == 0) return GF_OK;
        std::vector<BFin> fin(nd);
        int64_t tiles = 0, total = 0;
        for (int k = 0; k < nd; k++) {
            const BSim &b = host[done[k]];
            BFin &f = fin[k];
            f.base = (int64_t)b.t0 * WG;
            f.pre = b.T >= 0 ? pre[b.cur].p + b.p0 : nullptr;
            f.cnt = cnt.p + b.t0;
            f.T = b.T;
            f.M = b.M;
            f.cur = b.cur;
            f.tile_first = (int)tiles;
            f.out_base = total;
            tiles += b.T >= 0 ? b.T : (b.M + WG - 1) / WG;
            total += b.M;
        }
        std::vector<float> lc;
        int half_n;
        log_binom_table(n, lc, half_n);
        DevBuf<BFin> d_fin;
        DevBuf<float> d_lc;
        DevBuf<unsigned long long> d_acc;
        DevBuf<float> cf;
        DevBuf<gf_mutid> ci;
        GF_TRY(d_fin.alloc(nd));
        GF_TRY(d_lc.alloc(lc.size()));
        GF_TRY(d_acc.alloc((size_t)nd * (n + 2)));
        if (keep && total > 0) {
            GF_TRY(cf.alloc((size_t)total));
            GF_TRY(ci.alloc((size_t)total));
        }
        GF_HIP(hipMemcpyAsync(d_fin.p, fin.data(), nd * sizeof(BFin), hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(d_lc.p, lc.data(), lc.size() * sizeof(float), hipMemcpyHostToDevice, st));
        GF_HIP(hipMemsetAsync(d_acc.p, 0, (size_t)nd * (n + 2) * sizeof(unsigned long long), st));
        if (tiles > 0) {
            const int grid = (int)std::min<int64_t>(tiles, (int64_t)num_cus() * 8);
            float *cfp = keep && total > 0 ? cf.p : nullptr;
            gf_mutid *cip = keep && total > 0 ? ci.p : nullptr;
            if (n + 2 <= SFS_LDS_BINS)
                hipLaunchKernelGGL((k_bt_final<true>), dim3(grid), dim3(WG), 0, st, d_fin.p, nd, (int)tiles, d_fptr.p,
                                   d_iptr.p, n, half_n, d_lc.p, d_acc.p, cfp, cip);
            else
                hipLaunchKernelGGL((k_bt_final<false>), dim3(grid), dim3(WG), 0, st, d_fin.p, nd, (int)tiles, d_fptr.p,
                                   d_iptr.p, n, half_n, d_lc.p, d_acc.p, cfp, cip);
            GF_HIP(hipGetLastError());
        }
        std::vector<unsigned long long> acc((size_t)nd * (n + 2));
        GF_HIP(hipMemcpyAsync(acc.data(), d_acc.p, acc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        std::vector<float> hf;
        std::vector<gf_mutid> hi;
        if (keep && total > 0) {
            hf.resize((size_t)total);
            hi.resize((size_t)total);
            GF_TRY(readback(hf.data(), cf.p, (size_t)total * sizeof(float), st));
            GF_TRY(readback(hi.data(), ci.p, (size_t)total * sizeof(gf_mutid), st));
        }
        GF_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < nd; k++) {
            const int s = done[k];
            const BSim &b = host[s];
            R->num_mutations[s] = b.M;
            finish_sfs(&acc[(size_t)k * (n + 2)], acc[(size_t)k * (n + 2) + n + 1], n, b.M, in[s].num_sites,
                       &R->sfs[(size_t)s * n]);
            if (keep) {
                R->freq[s].assign(hf.begin() + fin[k].out_base, hf.begin() + fin[k].out_base + b.M);
                R->ids[s].assign(hi.begin() + fin[k].out_base, hi.begin() + fin[k].out_base + b.M);
            }
            finalised[s] = 1;
        }
        return GF_OK;
    }

    /* finalise finished simulations, then move the others (dense) into a
     * new layout: stalled ones with at least twice their tiles */
    int relayout(int K) {
        GF_TRY(sync_sims());
        std::vector<int> done, keep_s;
        for (int s : layout) {
            if (finalised[s]) continue;
            if (host[s].status == BS_DONE) done.push_back(s);
            else keep_s.push_back(s);
        }
        GF_TRY(finalise(done));
        /* the new regions */
        std::vector<BSim> old(host);
        int64_t t = 0;
        int k = 0;
        std::vector<BMove> mv;
        int64_t tiles = 0;
        for (int s : keep_s) {
            BSim &b = host[s];
            const double lam = (double)b.mu * (double)b.Nc * (double)b.L;
            int nt = b.nt;
            if (b.status == BS_STALLED) {
                const int sized = (int)std::ceil((double)g_region_scale * tiles_for(b.M, lam, K));
                nt = std::max({2 * b.nt, b.need, sized});
                b.status = BS_ACTIVE;
            }
            b.nt = nt;
            b.t0 = (int)t;
            b.p0 = (int)(t + k);
            BMove m;
            m.src_base = (int64_t)old[s].t0 * WG;
            m.dst_base = (int64_t)b.t0 * WG;
            m.pre = old[s].T >= 0 ? pre[old[s].cur].p + old[s].p0 : nullptr;
            m.src0 = old[s].T >= 0 ? src0[old[s].cur].p + old[s].t0 : nullptr;
            m.T = old[s].T;
            m.M = old[s].M;
            m.cur = old[s].cur;
            m.tile_first = (int)tiles;
            tiles += (b.M + WG - 1) / WG;
            mv.push_back(m);
            b.T = -1;
            b.cur = 0;
            t += nt;
            k++;
            if (t > 2000000000LL / WG) { set_error("gf_batch_run: more than 2e9 slots"); return GF_ERR_UNSUPPORTED; }
        }
        /* copy into fresh buffers, then swap them in */
        DevBuf<float> nf;
        DevBuf<gf_mutid> ni;
        const size_t slots = (size_t)std::max<int64_t>(t, 1) * WG;
        GF_TRY(nf.alloc(slots));
        GF_TRY(ni.alloc(slots));
        if (tiles > 0) {
            DevBuf<BMove> d_mv;
            GF_TRY(d_mv.alloc(mv.size()));
            GF_HIP(hipMemcpyAsync(d_mv.p, mv.data(), mv.size() * sizeof(BMove), hipMemcpyHostToDevice, st));
            const int grid = (int)std::min<int64_t>(tiles, (int64_t)num_cus() * 8);
            hipLaunchKernelGGL(k_bt_move, dim3(grid), dim3(WG), 0, st, d_mv.p, (int)mv.size(), (int)tiles,
                               d_fptr.p, d_iptr.p, nf.p, ni.p);
            GF_HIP(hipGetLastError());
            GF_HIP(hipStreamSynchronize(st));   /* before the old buffers and tables go */
        }
        freq[0].swap(nf);
        ids[0].swap(ni);
        nf.release();
        ni.release();
        freq[1].release();
        ids[1].release();
        for (int b = 0; b < 2; b++) { pre[b].release(); src0[b].release(); }
        layout = keep_s;
        GF_TRY(allocate_layout(t));
        GF_HIP(hipMemcpyAsync(d_sims.p, host.data(), (size_t)S * sizeof(BSim), hipMemcpyHostToDevice, st));
        relayouts++;
        return capture();
    }

    int run() {
        R->S = S;
        R->n = n;
        R->sfs.assign((size_t)S * n, 0.0f);
        R->num_mutations.assign(S, 0);
        if (keep) { R->freq.resize(S); R->ids.resize(S); }
        finalised.assign(S, 0);
        auto t0 = std::chrono::steady_clock::now();
        GF_HIP(st.create());
        StreamScope scope(st);
        hipEvent_t i0, i1, w0, w1;
        GF_HIP(hipEventCreate(&i0));
        GF_HIP(hipEventCreate(&i1));
        GF_HIP(hipEventCreate(&w0));
        GF_HIP(hipEventCreate(&w1));
        struct Ev { hipEvent_t *e[4]; ~Ev() { for (auto p : e) if (*p) (void)hipEventDestroy(*p); } } ev_guard{{&i0, &i1, &w0, &w1}};
        GF_HIP(hipEventRecord(i0, st));

        /* block length: the caller's cap (compaction interval, launch
         * length) or the tuner */
        int cap = std::min(BK_MAX, interval);
        const int forced = generations_per_launch(0);
        if (forced > 0) cap = std::min(cap, forced);
        GF_TRY(d_sims.alloc(S));
        GF_TRY(init(cap));
        GF_HIP(hipEventRecord(i1, st));
        int maxG = 0;
        for (int s = 0; s < S; s++) maxG = std::max(maxG, host[s].G);
        GF_TRY(rows.alloc((size_t)S * BK_MAX));
        GF_TRY(live.alloc(64));
        GF_TRY(d_K.alloc(1));
        GF_TRY(h_status.alloc(4));
        GF_HIP(hipMemsetAsync(live.p, 0, 64 * sizeof(unsigned long long), st));
        const BlockTuner::Key key = BlockTuner::key(S, ntiles * WG, maxG, cap);
        bool measuring = false;
        int K = forced > 0 ? cap : BlockTuner::get().next(key, cap, measuring);
        GF_HIP(hipMemcpyAsync(d_K.p, &K, sizeof(int), hipMemcpyHostToDevice, st));
        GF_HIP(hipStreamSynchronize(st));
        GF_TRY(capture());
        unsigned long long live_prev = 0;
        int per_check = 4;
        /* every replay advances each active simulation by at least one
         * generation unless it stalls, and a stall is resolved at the next
         * check: a bound no progressing batch reaches */
        const int64_t max_blocks = 4 * (int64_t)maxG + 1024;
        for (;;) {
            if ((int64_t)blocks > max_blocks + 64 * (int64_t)(relayouts + 1)) {
                set_error("gf_batch_run: no progress after " + std::to_string(blocks) + " blocks");
                return GF_ERR_DEVICE;
            }
            GF_HIP(hipEventRecord(w0, st));
            for (int r = 0; r < per_check; r++) GF_HIP(hipGraphLaunch(exec, st));
            blocks += per_check;
            GF_HIP(hipEventRecord(w1, st));
            GF_HIP(hipEventSynchronize(w1));
            const int active = h_status.p[0], stalled = h_status.p[1];
            unsigned long long lv;
            memcpy(&lv, h_status.p + 2, sizeof lv);
            if (measuring) {
                float ms = 0;
                GF_HIP(hipEventElapsedTime(&ms, w0, w1));
                BlockTuner::get().record(key, K, (double)(lv - live_prev), ms);
            } else if (per_check < 32) {
                per_check *= 2;
            }
            if (forced <= 0) {
                const int nextK = BlockTuner::get().next(key, cap, measuring);
                if (nextK != K) {
                    K = nextK;
                    GF_HIP(hipMemcpyAsync(d_K.p, &K, sizeof(int), hipMemcpyHostToDevice, st));
                }
            }
            live_prev = lv;
            if (active == 0 && stalled == 0) break;
            if (stalled > 0) {
                GF_TRY(relayout(K));
                continue;
            }
            /* most tiles belong to finished simulations: drop them */
            if (!measuring && ntiles > 4096) {
                GF_TRY(sync_sims());
                int64_t live_tiles = 0;
                for (int s : layout) if (host[s].status != BS_DONE) live_tiles += host[s].nt;
                if (4 * live_tiles < ntiles) GF_TRY(relayout(K));
            }
        }
        GF_TRY(sync_sims());
        std::vector<int> done;
        for (int s = 0; s < S; s++) if (!finalised[s]) done.push_back(s);
        GF_TRY(finalise(done));
        unsigned long long live64[64];
        GF_HIP(hipMemcpyAsync(live64, live.p, sizeof live64, hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        unsigned long long lt = 0;
        for (int k = 0; k < 64; k++) lt += live64[k];
        R->live_mutation_generations = (int64_t)lt;
        float ims = 0;
        (void)hipEventElapsedTime(&ims, i0, i1);
        R->timing.init_ms = ims;
        R->timing.advance_launches = blocks;
        R->timing.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        R->generations_per_block = K;
        R->blocks = blocks;
        R->relayouts = relayouts;
        return GF_OK;
    }
};

} /* namespace gf */

extern "C" {

int32_t gf_batch_run(const gf_batch_sim *sims, int32_t num_sims, int32_t compact_interval, int32_t sfs_n,
                     int32_t keep_states, int32_t device, gf_batch_result **out) {
    using namespace gf;
    if (!out) { set_error("gf_batch_run: out is NULL"); return GF_ERR_ARG; }
    *out = nullptr;
    if (!sims || num_sims <= 0 || sfs_n <= 0) {
        set_error("gf_batch_run: invalid arguments");
        return GF_ERR_ARG;
    }
    if (compact_interval <= 0) {
        set_error("gf_batch_run: compact_interval must be > 0 (the batch compacts at every block of at most "
                  "compact_interval generations)");
        return GF_ERR_ARG;
    }
    GF_TRY(select_device(device));
    std::unique_ptr<gf_batch_result> R(new gf_batch_result);
    BatchRun b;
    b.in = sims;
    b.S = num_sims;
    b.interval = compact_interval;
    b.n = sfs_n;
    b.keep = keep_states != 0;
    b.R = R.get();
    GF_TRY(b.run());
    *out = R.release();
    return GF_OK;
}

int32_t gf_batch_sfs(const gf_batch_result *r, float *out) {
    if (!r || !out) return GF_ERR_ARG;
    memcpy(out, r->sfs.data(), r->sfs.size() * sizeof(float));
    return GF_OK;
}

int32_t gf_batch_num_mutations(const gf_batch_result *r, int32_t *out) {
    if (!r || !out) return GF_ERR_ARG;
    memcpy(out, r->num_mutations.data(), r->num_mutations.size() * sizeof(int32_t));
    return GF_OK;
}

int32_t gf_batch_sim_state(const gf_batch_result *r, int32_t sim, float *freq, gf_mutid *ids) {
    if (!r || sim < 0 || sim >= r->S) { gf::set_error("gf_batch_sim_state: bad simulation index"); return GF_ERR_ARG; }
    if (r->freq.empty()) { gf::set_error("gf_batch_sim_state: states were not kept (keep_states = 0)"); return GF_ERR_ARG; }
    if (freq && !r->freq[sim].empty()) memcpy(freq, r->freq[sim].data(), r->freq[sim].size() * sizeof(float));
    if (ids && !r->ids[sim].empty()) memcpy(ids, r->ids[sim].data(), r->ids[sim].size() * sizeof(gf_mutid));
    return GF_OK;
}

int32_t gf_batch_live_mutation_generations(const gf_batch_result *r, int64_t *out) {
    if (!r || !out) return GF_ERR_ARG;
    *out = r->live_mutation_generations;
    return GF_OK;
}

int32_t gf_batch_timing(const gf_batch_result *r, gf_timing *out) {
    if (!r || !out) return GF_ERR_ARG;
    *out = r->timing;
    return GF_OK;
}

int32_t gf_set_batch_region_scale(float scale) {
    if (!(scale > 0.0f && scale <= 1.0f)) { gf::set_error("gf_set_batch_region_scale: scale must be in (0, 1]"); return GF_ERR_ARG; }
    gf::g_region_scale = scale;
    return GF_OK;
}

int32_t gf_batch_tuning(const gf_batch_result *r, int32_t *generations_per_block, int32_t *blocks, int32_t *relayouts) {
    if (!r) return GF_ERR_ARG;
    if (generations_per_block) *generations_per_block = r->generations_per_block;
    if (blocks) *blocks = r->blocks;
    if (relayouts) *relayouts = r->relayouts;
    return GF_OK;
}

void gf_batch_free(gf_batch_result *r) { delete r; }

} /* extern "C" */
