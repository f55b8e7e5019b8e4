/*
 * gofish.hip -- host driver of one GO_Fish simulation (the plain run_sim
 * path) and the C ABI entry points gf_run / gf_result_*.
 *
 * Reference: GO_Fish::run_sim, 3P/_internal/go_fish_impl.cuh:553-730.
 * The generation loop is planned on the host once (model tables, extinction
 * bookkeeping, parameter checks, new-mutation counts), uploaded in one copy,
 * and then executed as blocks of up to KMAX generations per k_advance launch;
 * a block ends at every generation that needs compaction, a time sample or
 * a preserve flag.  The host synchronises only at compaction (survivor count)
 * and at time samples (device-to-host copy).
 */
#include "gf_kernels.hpp"
#include "gf_runtime.hpp"
#include "gf_spectrum.hpp"

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

/* ---------------- result object ---------------- */
struct gf_sample_rec {
    int32_t num_mutations = 0, generation = 0;
    std::vector<int32_t> Nchrom_e;
    std::vector<uint8_t> extinct;
    std::vector<float> freq;   /* [P][M]; empty when the sample is the device-resident final state */
    bool on_device = false;
};

struct gf_result {
    int32_t P = 0, G = 0;
    std::vector<gf_sample_rec> samples;
    std::vector<gf_mutid> ids;
    int32_t num_ids = 0;
    /* the final state stays in device memory until read: the final
     * sample's frequencies ([P] rows of pitch dev_pitch) and the IDs are
     * copied once, straight into the caller's buffers */
    gf::DevBuf<float> dev_freq;
    gf::DevBuf<gf_mutid> dev_ids;
    int64_t dev_pitch = 0;
    int dev_device = 0;
    std::vector<int64_t> live;
    gf_timing timing{};
    /* stream and scratch of the device-side SFS calls on this result (not
     * its data: mutable through the const accessors; like the reference's
     * API, one result is used by one host thread at a time) */
    mutable gf::OwnedStream sfs_stream;
    mutable gf::SfsScratch sfs_scratch;
};

namespace gf {

/* ---------------- runtime globals ---------------- */
static thread_local std::string t_err;
void set_error(const std::string &m) { t_err = m; }
const std::string &get_error() { return t_err; }
static int g_timing = 0;
int timing_stride() { return g_timing; }
static int g_kpl = 0;    /* generations per advance launch (1..KMAX); 0: the caller's default */
int generations_per_launch(int dflt) { return g_kpl > 0 ? g_kpl : dflt; }

int select_device(int device) {
    if (device >= 0) GF_HIP(hipSetDevice(device));
    return GF_OK;
}

int num_cus() {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) return 256;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, d) != hipSuccess) return 256;
    return pr.multiProcessorCount;
}

int scan_counts(const int *d_in, int64_t n, int64_t *d_out, int64_t *d_total, hipStream_t stream,
                DevBuf<int64_t> &scratch) {
    if (n <= 0) {
        GF_HIP(hipMemsetAsync(d_total, 0, sizeof(int64_t), stream));
        return GF_OK;
    }
    int64_t nb = std::min<int64_t>(1024, std::max<int64_t>(1, (n + 4095) / 4096));
    int64_t per = (n + nb - 1) / nb;
    nb = (n + per - 1) / per;
    GF_TRY(scratch.alloc((size_t)nb));
    hipLaunchKernelGGL(k_scan_partial, dim3((unsigned)nb), dim3(256), 0, stream, d_in, n, per, scratch.p);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, stream, scratch.p, (int)nb, d_total);
    hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(256), 0, stream, d_in, n, per, scratch.p, d_out);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

/* ---------------- parameter checks (reference check_sim_parameters,
 * go_fish_impl.cuh:431-452); messages match the reference text ---------------- */
static int check_params(const gf_sim_desc *d, int t, int g, const std::vector<uint8_t> &extinct) {
    const int P = d->num_populations;
    const gf_pop_gen *pg = d->params + (size_t)t * P;
    const float *mig = d->migration + (size_t)t * P * P;
    char buf[256];
    for (int p = 0; p < P; p++) {
        double migration = 0;
        if (pg[p].mu < 0) {
            snprintf(buf, sizeof buf, "check_sim_parameters: mutation error: mu_rate < 0\tgeneration %d\t population %d", g, p);
            set_error(buf); return GF_ERR_PARAM;
        }
        const int N = pg[p].N;
        if (N > 0 && extinct[p]) {
            snprintf(buf, sizeof buf, "check_sim_parameters: demography error: extinct population with population size > 0\tgeneration %d\t population %d", g, p);
            set_error(buf); return GF_ERR_PARAM;
        }
        if (pg[p].F < 0) {
            snprintf(buf, sizeof buf, "check_sim_parameters: inbreeding error: inbreeding coefficient < 0\tgeneration %d\t population %d", g, p);
            set_error(buf); return GF_ERR_PARAM;
        }
        if (pg[p].F > 1) {
            snprintf(buf, sizeof buf, "check_sim_parameters: inbreeding error: inbreeding coefficient > 1\tgeneration %d\t population %d", g, p);
            set_error(buf); return GF_ERR_PARAM;
        }
        for (int p2 = 0; p2 < P; p2++) {
            const float m_from = mig[p * P + p2];
            if (m_from < 0) {
                snprintf(buf, sizeof buf, "check_sim_parameters: migration error: migration rate < 0\tgeneration %d\t population_from %d\t population_to %d", g, p, p2);
                set_error(buf); return GF_ERR_PARAM;
            }
            if (m_from > 0 && ((N <= 0 || extinct[p]) && !(pg[p2].N <= 0 || extinct[p2]))) {
                snprintf(buf, sizeof buf, "migration error, migration from non-existant population\tgeneration %d\t population_from %d\t population_to %d", g, p, p2);
                set_error(buf); return GF_ERR_PARAM;
            }
            migration += (double)mig[p2 * P + p];
        }
        if ((float)migration != 1.f && !(N <= 0 || extinct[p])) {
            snprintf(buf, sizeof buf, "check_sim_parameters: migration error: migration rate does not sum to 1\tgeneration %d\t population_TO %d\t total_migration_proportion %f", g, p, migration);
            set_error(buf); return GF_ERR_PARAM;
        }
    }
    return GF_OK;
}

/* The checks of check_params for every generation of a run, with the
 * extinction bookkeeping of PlainRun::plan (a population goes extinct when
 * its size drops to 0 after being positive), on the host only. */
static int validate_desc(const gf_sim_desc *d) {
    const int P = d->num_populations;
    const int G = d->num_generations;
    const int start = d->init_mode == GF_INIT_PREV ? d->start_generation : 0;
    std::vector<uint8_t> ext(P, 0);
    if (d->init_mode == GF_INIT_PREV && d->init_num_mutations > 0 && d->init_extinct)
        for (int p = 0; p < P; p++) ext[p] = d->init_extinct[p] ? 1 : 0;
    std::vector<int> prevN(P, 0);
    for (int t = 0; t <= G; t++) {
        GF_TRY(check_params(d, t, start + t, ext));
        const gf_pop_gen *pg = d->params + (size_t)t * P;
        for (int p = 0; p < P; p++) {
            if (t > 0 && !ext[p] && pg[p].N <= 0 && prevN[p] > 0) ext[p] = 1;
            prevN[p] = pg[p].N;
        }
    }
    return GF_OK;
}

/* ---------------- one simulation ---------------- */
template <int P>
struct PlainRun {
    const gf_sim_desc *d;
    const gf_selection_hooks *hooks = nullptr;   /* caller-compiled selection, or NULL: linear form */
    gf_result *R;
    OwnedStream st;   /* before the buffers: they are released while it is alive */
    int G, start, final_g, interval;
    uint32_t k0, k1;
    std::vector<GenRow<P>> rows;          /* rows[t], t = 0..G */
    std::vector<std::vector<uint8_t>> ext_at;
    DevBuf<GenRow<P>> d_rows;
    DevBuf<unsigned long long> d_live;
    DevBuf<unsigned int> d_tickets;     /* one ticket counter per launch */
    DevBuf<float> freqA, freqB;
    DevBuf<gf_mutid> idsA, idsB;
    int64_t cap = 0;
    int M = 0;
    /* compaction scratch */
    DevBuf<unsigned long long> d_masks;
    DevBuf<unsigned int> d_tcount, d_toff;
    DevBuf<int> d_small;       /* tile_first, tile_num, seg_M, seg_present, seg_preserve, seg_total */
    DevBuf<int64_t> d_small64; /* seg_base, seg_nbase */
    PinBuf<int> h_small;
    PinBuf<int64_t> h_small64;
    size_t tile_seg_n = 0;
    KernelTimer t_adv, t_cmp;
    int64_t slot_gens = 0;
    /* fused advance + compaction state */
    DevBuf<int> d_M;                    /* [2] slot counts, alternately in and out */
    int mcur = 0;                       /* d_M[mcur] = current slot count */
    /* fragmented layout of the A buffers after a fused launch (tile i holds
     * its survivors at its start): per-tile counts, and two generations of
     * prefix / first-source-tile arrays and tile counts, selected by par */
    bool frag = false;
    int par = 0;
    int64_t tiles_cap = 0;              /* entries of d_cnt and of each d_src0 half */
    DevBuf<unsigned int> d_cnt;
    DevBuf<int> d_pre, d_src0, d_T;
    DevBuf<int> d_chunk;                /* per-chunk survivor sums (multi-workgroup tile scan) */
    DevBuf<int> d_err;
    PinBuf<int> h_ring;                 /* slot count after each launch */
    std::vector<hipEvent_t> ev_ring;
    int64_t launches = 0;
    int64_t M_bound = 0;                /* upper bound of the current slot count */
    int64_t last_new = 0;               /* new slots of the most recent launch */

    int plan() {
        const int Pn = d->num_populations;
        (void)Pn;
        rows.assign(G + 1, GenRow<P>{});
        ext_at.assign(G + 1, std::vector<uint8_t>(P, 0));
        std::vector<uint8_t> ext(P, 0);
        if (d->init_mode == GF_INIT_PREV && d->init_num_mutations > 0 && d->init_extinct)
            for (int p = 0; p < P; p++) ext[p] = d->init_extinct[p] ? 1 : 0;
        std::vector<int> prevN(P);
        int64_t cum = 0;
        for (int t = 0; t <= G; t++) {
            const int g = start + t;
            GF_TRY(check_params(d, t, g, ext));
            const gf_pop_gen *pg = d->params + (size_t)t * P;
            const float *mig = d->migration + (size_t)t * P * P;
            GenRow<P> &r = rows[t];
            int Nc[P];
            for (int p = 0; p < P; p++) {
                if (t == 0) { Nc[p] = 0; continue; }
                const int N = pg[p].N;
                if (ext[p]) { Nc[p] = 0; continue; }
                if (N <= 0) {
                    if (prevN[p] > 0) ext[p] = 1;
                    Nc[p] = 0;
                    continue;
                }
                Nc[p] = nchrom(N, pg[p].F);
            }
            r.present_mask = 0;
            r.neutral_mask = 0;
            for (int p = 0; p < P; p++) {
                r.Nc[p] = Nc[p];
                r.fN[p] = (float)Nc[p];
                r.invN[p] = Nc[p] > 0 ? 1.0 / (double)(float)Nc[p] : 0.0;
                const float F = pg[p].F, h = pg[p].h;
                r.sa[p] = pg[p].sel_slope;
                r.sb[p] = pg[p].sel_intercept;
                r.c1[p] = (F + h) - (h * F);
                r.c2[p] = (F + (2.0f * h)) - ((2.0f * h) * F);
                for (int k = 0; k < P; k++) r.mig[k][p] = mig[k * P + p];
                r.newfreq[p] = Nc[p] > 0 ? 1.0f / (float)Nc[p] : 0.0f;
                if (pg[p].N > 0) r.present_mask |= 1 << p;
                if (!hooks && pg[p].sel_slope == 0.0f && pg[p].sel_intercept == 0.0f) r.neutral_mask |= 1 << p;
            }
            for (int p = 0; p < P; p++) {
                r.newcum[p] = (int)cum;
                if (t > 0) cum += new_mutation_count(Nc[p], pg[p].mu, d->num_sites, k0, k1, g, p);
                if (cum > 2000000000LL) {
                    set_error("run_sim: more than 2e9 new mutations over the run are not supported");
                    return GF_ERR_UNSUPPORTED;
                }
            }
            r.newcum[P] = (int)cum;
            for (int p = 0; p < P; p++) prevN[p] = pg[p].N;
            ext_at[t] = ext;
        }
        return GF_OK;
    }

    int ensure_capacity(int64_t need) {
        if (need <= cap && freqA.p) return GF_OK;
        int64_t ncap = std::max<int64_t>({need, cap + cap / 2, (int64_t)1 << 16});
        ncap = (ncap + 255) / 256 * 256;
        DevBuf<float> nf;
        DevBuf<gf_mutid> ni;
        GF_TRY(nf.alloc((size_t)P * ncap));
        GF_TRY(ni.alloc((size_t)ncap));
        if (M > 0) {
            GF_HIP(hipMemcpy2DAsync(nf.p, ncap * sizeof(float), freqA.p, cap * sizeof(float),
                                    (size_t)M * sizeof(float), P, hipMemcpyDeviceToDevice, st));
            GF_HIP(hipMemcpyAsync(ni.p, idsA.p, (size_t)M * sizeof(gf_mutid), hipMemcpyDeviceToDevice, st));
        }
        GF_HIP(hipStreamSynchronize(st));
        freqA.swap(nf);
        idsA.swap(ni);
        GF_TRY(freqB.alloc((size_t)P * ncap));
        GF_TRY(idsB.alloc((size_t)ncap));
        cap = ncap;
        /* the A state is dense here: the tile arrays hold nothing live */
        const int64_t nt = ncap / WG;
        if (nt > tiles_cap) {
            GF_TRY(d_cnt.alloc((size_t)nt));
            GF_TRY(d_pre.alloc(2 * (size_t)pre_stride(nt)));
            GF_TRY(d_src0.alloc(2 * (size_t)nt));
            GF_TRY(d_chunk.alloc((size_t)((nt + SCAN_CHUNK - 1) / SCAN_CHUNK)));
            tiles_cap = nt;
        }
        if (!d_T.p) GF_TRY(d_T.alloc(2));
        return GF_OK;
    }

    /* make the A state dense (slot j = compacted index j) */
    int densify() {
        if (!frag) return GF_OK;
        const int64_t nb = std::max<int64_t>(1, std::min<int64_t>((M_bound + WG - 1) / WG, (int64_t)num_cus() * 8));
        hipLaunchKernelGGL((k_densify<P>), dim3((unsigned)nb), dim3(WG), 0, st, freqA.p, idsA.p, freqB.p, idsB.p, cap,
                           d_pre.p + (size_t)par * pre_stride(tiles_cap), d_src0.p + (size_t)par * tiles_cap, d_T.p + par,
                           d_M.p + mcur);
        GF_HIP(hipGetLastError());
        freqA.swap(freqB);
        idsA.swap(idsB);
        frag = false;
        return GF_OK;
    }

    int init_mse() {
        bool pres = d->preserve[0] || d->sample[0];
        std::vector<MseSeg> segs;
        int64_t coff = 0, aoff = 0, choff = 0;
        bool any_sel = false;
        for (int p = 0; p < P; p++) {
            const gf_pop_gen &pg = d->params[p];
            const int Nc = nchrom(pg.N, pg.F);
            if (Nc <= 1) continue;
            MseSeg s{};
            s.Nc = Nc; s.Nind = pg.N; s.F = pg.F; s.h = pg.h; s.mu = pg.mu;
            s.sa = pg.sel_slope; s.sb = pg.sel_intercept; s.L = d->num_sites;
            s.k0 = k0; s.k1 = k1; s.pop = p;
            s.neutral = (!hooks && pg.sel_slope == 0.0f && pg.sel_intercept == 0.0f) ? 1 : 0;
            s.preserve = pres ? 1 : 0;
            s.class_off = coff; s.area_off = aoff; s.chunk_off = choff;
            s.sim_base = 0;
            coff += Nc - 1; aoff += Nc; choff += (Nc + MSE_CHUNK - 1) / MSE_CHUNK;
            if (!s.neutral) any_sel = true;
            segs.push_back(s);
        }
        /* a caller-compiled selection functor: tabulate it at the
         * equilibrium points of every population (generation start) */
        DevBuf<float> d_seltab;
        if (hooks && !segs.empty()) {
            int64_t n = 0;
            for (const MseSeg &m : segs) n += 3 * (int64_t)m.Nc;
            GF_TRY(d_seltab.alloc((size_t)n));
            int64_t o = 0;
            for (MseSeg &m : segs) {
                const int32_t rc = hooks->mse_table(hooks->user, m.pop, start, m.Nc, d_seltab.p + o, (void *)st);
                if (rc != GF_OK) { set_error("run_sim: selection table hook failed"); return rc; }
                m.sel_tab = d_seltab.p + o;
                o += 3 * (int64_t)m.Nc;
            }
        }
        M = 0;
        if (segs.empty()) return ensure_capacity(1);
        const int ns = (int)segs.size();
        std::vector<int64_t> class_offs(ns), area_offs(ns), chunk_offs(ns);
        for (int s = 0; s < ns; s++) { class_offs[s] = segs[s].class_off; area_offs[s] = segs[s].area_off; chunk_offs[s] = segs[s].chunk_off; }
        DevBuf<MseSeg> d_segs;
        DevBuf<int64_t> d_offs;
        GF_TRY(d_segs.alloc(ns));
        GF_TRY(d_offs.alloc(3 * (size_t)ns));
        GF_HIP(hipMemcpyAsync(d_segs.p, segs.data(), ns * sizeof(MseSeg), hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(d_offs.p, class_offs.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(d_offs.p + ns, area_offs.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(d_offs.p + 2 * ns, chunk_offs.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, st));
        DevBuf<float> d_area, d_tot;
        const unsigned grid = (unsigned)std::max(1, num_cus() * 8);
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
        DevBuf<int64_t> d_scan, d_scratch, d_total;
        GF_TRY(d_counts.alloc((size_t)coff));
        GF_TRY(d_scan.alloc((size_t)coff));
        GF_TRY(d_total.alloc(1));
        hipLaunchKernelGGL(k_mse_counts, dim3(grid), dim3(256), 0, st, d_segs.p, d_offs.p, ns, coff,
                           any_sel ? d_area.p : nullptr, d_counts.p);
        GF_HIP(hipGetLastError());
        GF_TRY(scan_counts(d_counts.p, coff, d_scan.p, d_total.p, st, d_scratch));
        /* per-segment first offsets and the total */
        std::vector<int64_t> seg_first(ns);
        for (int s = 0; s < ns; s++)
            GF_HIP(hipMemcpyAsync(&seg_first[s], d_scan.p + segs[s].class_off, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        int64_t total = 0;
        GF_HIP(hipMemcpyAsync(&total, d_total.p, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        if (total > 2000000000LL) { set_error("initial mutation count exceeds 2e9"); return GF_ERR_UNSUPPORTED; }
        for (int s = 0; s < ns; s++) segs[s].slot_base = seg_first[s];
        M = 0;
        GF_TRY(ensure_capacity(total + 1));
        GF_HIP(hipMemcpyAsync(d_segs.p, segs.data(), ns * sizeof(MseSeg), hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(d_offs.p, seg_first.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, st));
        if (total > 0) {
            /* zero every row first: k_mse_fill writes all P rows of each slot */
            hipLaunchKernelGGL(k_mse_fill<P>, dim3(grid), dim3(256), 0, st, d_segs.p, d_offs.p, ns, total,
                               d_scan.p, freqA.p, (int64_t)cap, idsA.p);
            GF_HIP(hipGetLastError());
        }
        GF_HIP(hipStreamSynchronize(st));
        M = (int)total;
        return GF_OK;
    }

    int init_prev() {
        const int n = d->init_num_mutations;
        M = 0;
        GF_TRY(ensure_capacity(std::max(n, 1)));
        if (n > 0) {
            if (!d->init_freq || !d->init_ids) { set_error("init_freq/init_ids missing"); return GF_ERR_ARG; }
            const bool pres = d->preserve[0] || d->sample[0];
            std::vector<gf_mutid> ids(d->init_ids, d->init_ids + n);
            if (pres)
                for (auto &m : ids) { int z = m.origin_threadID; m.origin_threadID = -(z < 0 ? -z : z); }
            GF_HIP(hipMemcpy2DAsync(freqA.p, cap * sizeof(float), d->init_freq, (size_t)n * sizeof(float),
                                    (size_t)n * sizeof(float), P, hipMemcpyHostToDevice, st));
            GF_HIP(hipMemcpyAsync(idsA.p, ids.data(), (size_t)n * sizeof(gf_mutid), hipMemcpyHostToDevice, st));
            GF_HIP(hipStreamSynchronize(st));
        }
        M = n;
        return GF_OK;
    }

    int ensure_tiles(int64_t tiles) {
        if ((size_t)tiles <= tile_seg_n && d_masks.p) return GF_OK;
        size_t n = std::max<size_t>((size_t)tiles, 1024);
        GF_TRY(d_masks.alloc(n * (WG / 64)));
        GF_TRY(d_tcount.alloc(n));
        GF_TRY(d_toff.alloc(n));
        tile_seg_n = n;
        return GF_OK;
    }

    /* count of the slots the compaction rule keeps (initial live count;
     * reference flag_segregating_mutations, go_fish_impl.cuh:137-176) */
    int flag_scan(int present_mask, bool preserve, int &survivors) {
        const int64_t tiles = (M + WG - 1) / WG;
        GF_TRY(ensure_tiles(tiles));
        GF_TRY(h_small.alloc(16));
        GF_TRY(h_small64.alloc(4));
        GF_TRY(d_small.alloc(16));
        GF_TRY(d_small64.alloc(4));
        int *hs = h_small.p;
        hs[0] = 0;                  /* tile_first */
        hs[1] = (int)tiles;         /* tile_num */
        hs[2] = M;                  /* seg_M */
        hs[3] = present_mask;       /* seg_present */
        hs[4] = preserve ? 1 : 0;   /* seg_preserve */
        hs[5] = 0;                  /* seg_total (out) */
        h_small64.p[0] = 0;         /* seg_base */
        h_small64.p[1] = 0;         /* seg_nbase */
        GF_HIP(hipMemcpyAsync(d_small.p, hs, 6 * sizeof(int), hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(d_small64.p, h_small64.p, 2 * sizeof(int64_t), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_flag<P>, dim3((unsigned)tiles), dim3(WG), 0, st, freqA.p, (int64_t)cap, idsA.p,
                           d_small64.p, d_small.p + 2, d_small.p + 3, d_small.p + 0, 1,
                           d_masks.p, d_tcount.p);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, d_tcount.p, d_small.p + 0, d_small.p + 1,
                           d_toff.p, d_small.p + 5);
        GF_HIP(hipGetLastError());
        GF_HIP(hipMemcpyAsync(&hs[8], d_small.p + 5, sizeof(int), hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        survivors = hs[8];
        return GF_OK;
    }

    int take_sample(int t, bool final_state) {
        GF_TRY(densify());
        gf_sample_rec s;
        s.num_mutations = M;
        s.generation = start + t;
        if (final_state) {
            /* the run's last use of the state: hand the buffers to the result */
            GF_HIP(hipStreamSynchronize(st));
            R->dev_freq.swap(freqA);
            R->dev_ids.swap(idsA);
            R->dev_pitch = cap;
            R->num_ids = M;
            GF_HIP(hipGetDevice(&R->dev_device));
            s.on_device = true;
        } else {
            s.freq.resize((size_t)P * M);
            for (int p = 0; p < P && M > 0; p++)
                GF_TRY(readback(s.freq.data() + (size_t)p * M, freqA.p + (size_t)p * cap, (size_t)M * sizeof(float), st));
        }
        GF_HIP(hipStreamSynchronize(st));
        const gf_pop_gen *pg = d->params + (size_t)t * P;
        for (int p = 0; p < P; p++) {
            s.Nchrom_e.push_back(nchrom(pg[p].N, pg[p].F));
            s.extinct.push_back(ext_at[t][p]);
        }
        R->samples.push_back(std::move(s));
        return GF_OK;
    }

    static constexpr int RING = 4;

    /* exact slot count: wait for the stream */
    int exact_M(int &out) {
        GF_HIP(hipMemcpyAsync(&h_ring.p[RING], d_M.p + mcur, sizeof(int), hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        out = h_ring.p[RING];
        M_bound = out;
        return GF_OK;
    }

    /* entries of each half of d_pre: T + 1 prefix values, rounded up to a
     * multiple of 4 so that both halves take 16-byte stores */
    static size_t pre_stride(int64_t tiles) { return ((size_t)tiles + 1 + 3) & ~(size_t)3; }

    /* grow the buffers to hold `need` slots; the current contents (exact
     * count read from the device) move along */
    int grow(int64_t need) {
        if (need <= cap && freqA.p) return GF_OK;
        GF_TRY(densify());
        int cur = 0;
        GF_TRY(exact_M(cur));
        M = cur;
        return ensure_capacity(need);
    }

    /* Advance generations t_first .. t_first+K-1 in one fused launch
     * (reference migration_selection_drift + add_new_mutations per
     * generation, go_fish_impl.cuh:645-684, and, when `compact`,
     * compact :370-429 at the block's last generation). */
    int launch_block(int t_first, int K, bool compact, bool preserve) {
        const int64_t new_in_block = (int64_t)rows[t_first + K - 1].newcum[P] - rows[t_first].newcum[0];
        int64_t bound = M_bound + new_in_block;
        if (bound > 2000000000LL) { set_error("more than 2e9 mutation slots are not supported"); return GF_ERR_UNSUPPORTED; }
        if (bound > cap) {
            GF_TRY(grow(bound + bound / 2));
            bound = M_bound + new_in_block;
        }
        const int ntiles = (int)((bound + WG - 1) / WG);
        const size_t pre_half = pre_stride(tiles_cap), src_half = (size_t)tiles_cap;
        FusedLaunch F{};
        F.freq_in = freqA.p; F.ids_in = idsA.p; F.freq_out = freqB.p; F.ids_out = idsB.p;
        F.pitch = cap;
        F.M_in = d_M.p + mcur;
        F.pre_in = d_pre.p + par * pre_half;
        F.src0_in = d_src0.p + par * src_half;
        F.T_in = d_T.p + par;
        F.cnt_out = d_cnt.p;
        F.ticket = d_tickets.p + launches;
        F.error = d_err.p;
        F.ntiles = ntiles;
        F.new_in_block = (int)new_in_block;
        F.dense = frag ? 0 : 1;
        F.compact = compact ? 1 : 0;
        F.preserve = preserve ? 1 : 0;
        SegLaunch<P> sl{};
        sl.rows = d_rows.p + t_first;
        sl.live = d_live.p;
        sl.base = 0;
        sl.cum_base = rows[t_first].newcum[0];
        sl.K = K;
        sl.g0 = start + t_first - 1;
        sl.t0 = t_first;
        sl.k0 = k0; sl.k1 = k1;
        const int grid = std::max(1, std::min(ntiles, num_cus() * 8));
        hipEvent_t ev = nullptr;
        t_adv.begin(st, ev);
        if (hooks) {
            const int32_t rc = hooks->advance(hooks->user, P, &F, &sl, grid, (void *)st);
            if (rc != GF_OK) { set_error("run_sim: selection advance hook failed"); return rc; }
        } else {
            hipLaunchKernelGGL((k_advance_fused<P, LinearSel>), dim3(grid), dim3(WG), 0, st, F, sl, LinearSel());
        }
        GF_HIP(hipGetLastError());
        t_adv.end(ev);
        /* survivors per tile -> prefix, slot count, next launch's source tiles */
        const int slot = (int)(launches % RING);
        /* more than two chunks of tiles: sum the chunks, then scan them in
         * parallel (one workgroup walking 14 chunks took 110 us) */
        const int nchunks = (int)((ntiles + SCAN_CHUNK - 1) / SCAN_CHUNK);
        const bool chunked = nchunks > 2;
        if (chunked) {
            if ((size_t)nchunks > d_chunk.n) { set_error("tile scan: chunk buffer smaller than the launch"); return GF_ERR_DEVICE; }
            hipLaunchKernelGGL(k_tile_chunk_sums, dim3(nchunks), dim3(1024), 0, st, d_cnt.p, d_M.p + mcur,
                               (int)new_in_block, d_chunk.p);
            GF_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_tile_scan, dim3(chunked ? nchunks : 1), dim3(1024), 0, st, d_cnt.p, d_M.p + mcur,
                           (int)new_in_block, d_pre.p + (par ^ 1) * pre_half, d_src0.p + (par ^ 1) * src_half,
                           d_M.p + (mcur ^ 1), d_T.p + (par ^ 1), &h_ring.p[slot],
                           chunked ? (const int *)d_chunk.p : (const int *)nullptr);
        GF_HIP(hipGetLastError());
        freqA.swap(freqB);
        idsA.swap(idsB);
        mcur ^= 1;
        par ^= 1;
        frag = true;
        /* pipelined exact count: the scan wrote this launch's slot count to
         * pinned host memory; tighten the bound with the previous launch's */
        GF_HIP(hipEventRecord(ev_ring[slot], st));
        if (launches > 0) {
            const int prev = (int)((launches - 1) % RING);
            GF_HIP(hipEventSynchronize(ev_ring[prev]));
            M_bound = (int64_t)h_ring.p[prev] + new_in_block;
        } else {
            M_bound = bound;
        }
        slot_gens += M_bound * K;
        last_new = new_in_block;
        launches++;
        return GF_OK;
    }

    int run() {
        G = d->num_generations;
        start = d->init_mode == GF_INIT_PREV ? d->start_generation : 0;
        final_g = start + G;
        interval = d->compact_interval;
        k0 = (uint32_t)d->seed1;
        k1 = (uint32_t)d->seed2;
        R->P = P;
        R->G = G;
        t_adv.stride = t_cmp.stride = timing_stride();
        auto t_start = std::chrono::steady_clock::now();
        GF_TRY(select_device(d->device));
        GF_HIP(st.create());
        StreamScope scope(st);
        GF_TRY(plan());
        GF_TRY(d_rows.alloc(rows.size()));
        GF_HIP(hipMemcpyAsync(d_rows.p, rows.data(), rows.size() * sizeof(GenRow<P>), hipMemcpyHostToDevice, st));
        GF_TRY(d_live.alloc((size_t)(G + 1) * 64));
        GF_HIP(hipMemsetAsync(d_live.p, 0, (size_t)(G + 1) * 64 * sizeof(unsigned long long), st));

        hipEvent_t i0, i1;
        GF_HIP(hipEventCreate(&i0));
        GF_HIP(hipEventCreate(&i1));
        GF_HIP(hipEventRecord(i0, st));
        if (d->init_mode == GF_INIT_MSE) GF_TRY(init_mse());
        else if (d->init_mode == GF_INIT_PREV) GF_TRY(init_prev());
        else { M = 0; GF_TRY(ensure_capacity(1)); }
        GF_HIP(hipEventRecord(i1, st));

        /* live count of the initial state */
        int64_t live0 = 0;
        if (M > 0) {
            int present = 0;
            for (int p = 0; p < P; p++) if (d->params[p].N > 0) present |= 1 << p;
            int surv = 0;
            GF_TRY(flag_scan(present, false, surv));
            live0 = surv;
        }
        if (d->sample[0] && start != final_g) GF_TRY(take_sample(0, false));

        /* fused-launch state */
        GF_TRY(d_M.alloc(2));
        GF_HIP(hipMemcpyAsync(d_M.p, &M, sizeof(int), hipMemcpyHostToDevice, st));
        GF_TRY(d_tickets.alloc((size_t)G + 2));
        GF_HIP(hipMemsetAsync(d_tickets.p, 0, ((size_t)G + 2) * sizeof(unsigned int), st));
        GF_TRY(d_err.alloc(1));
        GF_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int), st));
        GF_TRY(h_ring.alloc(RING + 1, hipHostMallocCoherent));
        ev_ring.assign(RING, nullptr);
        for (auto &e : ev_ring) GF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        GF_HIP(hipStreamSynchronize(st));
        mcur = 0;
        M_bound = M;

        int next_compact = start + interval;
        int t = 0;   /* generations done */
        /* generations per launch, unless the caller set it.  One population
         * compacts at every launch: a longer launch saves launch overhead
         * but carries the slots lost inside it to its end.  Measured best
         * (N = 1e4, 1,000 generations): 64 at 8.4e3 and 8.4e4 live
         * mutations (-41 % and -38 % against 10), 16 at 8.4e5; 10 at 2.0e6
         * (N = 1e5, theta = 2e5).  Several populations end launches at the
         * compaction interval anyway: up to 2e6 slots a launch runs to it
         * (two populations, N = 1e4, 1,000 generations, interval 35: 20
         * generations per launch 3.73 / 5.82 / 21.9 ms at 1.4e4 / 1.4e5 /
         * 1.4e6 live, up to the interval 3.35 / 4.95 / 19.6 ms), 20
         * generations beyond (three populations of 2e4-2e5 over 2,000
         * generations: 16 per launch 158 ms, 20 146 ms, 24 146 ms). */
        const int k_user = generations_per_launch(0);
        auto launch_length = [&]() {
            int k;
            if (k_user > 0) k = k_user;
            else if (P == 1) k = M_bound <= 200000 ? 64 : (M_bound <= 1200000 ? 16 : 10);
            else k = M_bound <= 2000000 ? 64 : 20;
            return std::max(1, std::min(KMAX, k));
        };
        /* one population: a lost or fixed mutation never changes again, so
         * compacting every launch leaves every reported sample unchanged */
        const bool compact_every_launch = (P == 1 && interval > 0);
        int rc = GF_OK;
        while (t < G) {
            const int kpl = launch_length();
            int K = 0;
            for (;;) {
                K++;
                const int te = t + K;
                const int g = start + te;
                const bool pres = d->preserve[te] || d->sample[te];
                /* with compaction at every launch the reference's compaction
                 * generations need no launch boundary of their own */
                const bool cmp = interval > 0 && ((g == next_compact && !compact_every_launch) || g == final_g || pres);
                if (cmp || d->sample[te] || te == G || K == kpl) break;
            }
            const int te = t + K;
            const int g = start + te;
            const bool pres = d->preserve[te] || d->sample[te];
            const bool user_cmp = interval > 0 && (g == next_compact || g == final_g || pres);
            rc = launch_block(t + 1, K, user_cmp || compact_every_launch, user_cmp && pres);
            if (rc) break;
            t = te;
            if (user_cmp) next_compact = g + interval;
            while (interval > 0 && next_compact <= g) next_compact += interval;
            if (d->sample[t] && g != final_g) {
                int cur = 0;
                if ((rc = exact_M(cur))) break;
                M = cur;
                if ((rc = take_sample(t, false))) break;
            }
        }
        for (auto &e : ev_ring) if (e) (void)hipEventDestroy(e);
        if (rc) return rc;
        {
            int cur = 0;
            GF_TRY(exact_M(cur));
            M = cur;
            int err = 0;
            GF_HIP(hipMemcpy(&err, d_err.p, sizeof(int), hipMemcpyDeviceToHost));
            if (err) { set_error("advance kernel: the launch had fewer tiles than slots (internal bound)"); return GF_ERR_DEVICE; }
        }
        GF_TRY(take_sample(G, true));
        GF_HIP(hipStreamSynchronize(st));

        /* live counts */
        DevBuf<unsigned long long> d_sum;
        GF_TRY(d_sum.alloc((size_t)G + 1));
        hipLaunchKernelGGL(k_sum_live, dim3((G + 1 + 255) / 256), dim3(256), 0, st, d_live.p, d_sum.p, G + 1);
        GF_HIP(hipGetLastError());
        std::vector<unsigned long long> lv((size_t)G + 1);
        GF_HIP(hipMemcpyAsync(lv.data(), d_sum.p, lv.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        R->live.assign(G + 1, 0);
        R->live[0] = live0;
        for (int tt = 1; tt <= G; tt++) R->live[tt] = (int64_t)lv[(size_t)tt];
        float ims = 0;
        (void)hipEventElapsedTime(&ims, i0, i1);
        (void)hipEventDestroy(i0);
        (void)hipEventDestroy(i1);
        R->timing.init_ms = ims;
        R->timing.advance_ms = t_adv.total_ms();
        R->timing.advance_launches = t_adv.count();
        R->timing.advance_timed_launches = t_adv.timed();
        R->timing.compact_ms = t_cmp.total_ms();
        R->timing.compact_launches = t_cmp.count();
        R->timing.compact_timed_launches = t_cmp.timed();
        R->timing.advance_slot_generations = slot_gens;
        R->timing.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        return GF_OK;
    }
};

template <int P>
static int run_p(const gf_sim_desc *d, const gf_selection_hooks *hooks, gf_result *R) {
    PlainRun<P> r;
    r.d = d;
    r.hooks = hooks;
    r.R = R;
    return r.run();
}

} /* namespace gf */

/* ---------------- C ABI ---------------- */
extern "C" {

const char *gf_last_error(void) { return gf::get_error().c_str(); }

int32_t gf_set_kernel_timing(int32_t enable) {
    if (enable < 0) { gf::set_error("kernel timing: enable must be >= 0"); return GF_ERR_ARG; }
    gf::g_timing = enable;
    return GF_OK;
}

int32_t gf_set_generations_per_launch(int32_t k) {
    if (k < 0 || k > gf::KMAX) { gf::set_error("generations per launch must be in [1, 64] (0: default)"); return GF_ERR_ARG; }
    gf::g_kpl = k;
    return GF_OK;
}

int32_t gf_release_device_cache(int32_t device) {
    int cur = 0;
    GF_HIP(hipGetDevice(&cur));
    GF_TRY(gf::select_device(device));
    int dev = 0;
    GF_HIP(hipGetDevice(&dev));
    gf::DevCache::get().flush(dev);
    if (dev != cur) GF_HIP(hipSetDevice(cur));
    return GF_OK;
}

int32_t gf_device_cache_stats(int32_t device, int64_t *out) {
    if (!out) { gf::set_error("gf_device_cache_stats: null output"); return GF_ERR_ARG; }
    int cur = 0;
    GF_HIP(hipGetDevice(&cur));
    GF_TRY(gf::select_device(device));
    int dev = 0;
    GF_HIP(hipGetDevice(&dev));
    gf::DevCache::get().stats(dev, out);
    if (dev != cur) GF_HIP(hipSetDevice(cur));
    return GF_OK;
}

int32_t gf_device_count(int32_t *out) {
    int n = 0;
    GF_HIP(hipGetDeviceCount(&n));
    if (out) *out = n;
    return GF_OK;
}

/* the descriptor checks shared by gf_run and gf_check_sim */
static int32_t check_desc_args(const gf_sim_desc *d) {
    if (!d || !d->params || !d->migration || !d->preserve || !d->sample) {
        gf::set_error("gf_run: descriptor or one of its tables is NULL");
        return GF_ERR_ARG;
    }
    if (d->num_populations < 1 || d->num_populations > GF_MAX_POPULATIONS) {
        gf::set_error("gf_run: num_populations must be in [1, GF_MAX_POPULATIONS]");
        return GF_ERR_ARG;
    }
    if (d->num_generations < 0 || d->init_mode < 0 || d->init_mode > 2 || d->compact_interval < 0 ||
        (d->init_mode == GF_INIT_PREV && d->init_num_mutations < 0)) {
        gf::set_error("gf_run: invalid num_generations, init_mode, compact_interval or init_num_mutations");
        return GF_ERR_ARG;
    }
    return GF_OK;
}

int32_t gf_check_sim(const gf_sim_desc *d) {
    GF_TRY(check_desc_args(d));
    return gf::validate_desc(d);
}

static int32_t run_any(const gf_sim_desc *d, const gf_selection_hooks *hooks, gf_result **out) {
    if (!out) { gf::set_error("gf_run: out is NULL"); return GF_ERR_ARG; }
    *out = nullptr;
    GF_TRY(check_desc_args(d));
    std::unique_ptr<gf_result> R(new gf_result);
    int rc;
    switch (d->num_populations) {
    case 1: rc = gf::run_p<1>(d, hooks, R.get()); break;
    case 2: rc = gf::run_p<2>(d, hooks, R.get()); break;
    case 3: rc = gf::run_p<3>(d, hooks, R.get()); break;
    case 4: rc = gf::run_p<4>(d, hooks, R.get()); break;
    case 5: rc = gf::run_p<5>(d, hooks, R.get()); break;
    case 6: rc = gf::run_p<6>(d, hooks, R.get()); break;
    case 7: rc = gf::run_p<7>(d, hooks, R.get()); break;
    default: rc = gf::run_p<8>(d, hooks, R.get()); break;
    }
    if (rc != GF_OK) return rc;
    *out = R.release();
    return GF_OK;
}

int32_t gf_run(const gf_sim_desc *d, gf_result **out) { return run_any(d, nullptr, out); }

int32_t gf_run_with_selection(const gf_sim_desc *d, const gf_selection_hooks *hooks, gf_result **out) {
    if (!hooks || !hooks->mse_table || !hooks->advance) {
        gf::set_error("gf_run_with_selection: hooks or one of its functions is NULL");
        if (out) *out = nullptr;
        return GF_ERR_ARG;
    }
    return run_any(d, hooks, out);
}

int32_t gf_result_num_samples(const gf_result *r) { return r ? (int32_t)r->samples.size() : GF_ERR_ARG; }

int32_t gf_result_sample_info(const gf_result *r, int32_t s, int32_t *M, int32_t *gen, int32_t *Nc, uint8_t *ext) {
    if (!r || s < 0 || s >= (int32_t)r->samples.size()) { gf::set_error("sample index out of range"); return GF_ERR_ARG; }
    const gf_sample_rec &x = r->samples[s];
    if (M) *M = x.num_mutations;
    if (gen) *gen = x.generation;
    for (int p = 0; p < r->P; p++) {
        if (Nc) Nc[p] = x.Nchrom_e[p];
        if (ext) ext[p] = x.extinct[p];
    }
    return GF_OK;
}

/* device -> host copy of a result's device-resident final state */
static int32_t result_readback(const gf_result *r, void *dst, const void *src, size_t bytes) {
    int cur = 0;
    GF_HIP(hipGetDevice(&cur));
    if (cur != r->dev_device) GF_HIP(hipSetDevice(r->dev_device));
    const int rc = gf::readback(dst, src, bytes, nullptr);
    if (cur != r->dev_device) (void)hipSetDevice(cur);
    return rc;
}

int32_t gf_result_sample_freq(const gf_result *r, int32_t s, float *out) {
    if (!r || !out || s < 0 || s >= (int32_t)r->samples.size()) { gf::set_error("sample index out of range"); return GF_ERR_ARG; }
    const gf_sample_rec &x = r->samples[s];
    if (x.on_device) {
        const size_t M = (size_t)x.num_mutations;
        for (int p = 0; p < r->P && M > 0; p++)
            GF_TRY(result_readback(r, out + (size_t)p * M, r->dev_freq.p + (size_t)p * r->dev_pitch, M * sizeof(float)));
    } else if (!x.freq.empty()) {
        memcpy(out, x.freq.data(), x.freq.size() * sizeof(float));
    }
    return GF_OK;
}

int32_t gf_result_site_frequency_spectrum(const gf_result *r, int32_t s, int32_t pop, int32_t n, float num_sites,
                                          float *out) {
    if (!r || !out || s < 0 || s >= (int32_t)r->samples.size() || pop < 0 || pop >= r->P || n <= 0) {
        gf::set_error("site_frequency_spectrum: invalid arguments");
        return GF_ERR_ARG;
    }
    const gf_sample_rec &x = r->samples[s];
    if (n >= x.Nchrom_e[pop]) {
        gf::set_error("site_frequency_spectrum error: requested sample_size out of range [1,population_size): sample_size " + std::to_string(n) + " [1," + std::to_string(x.Nchrom_e[pop]) + ")");
        return GF_ERR_ARG;
    }
    if (!x.on_device)
        return gf_site_frequency_spectrum(x.freq.data() + (size_t)pop * x.num_mutations, x.num_mutations, n, num_sites,
                                          out, r->dev_device);
    int cur = 0;
    GF_HIP(hipGetDevice(&cur));
    if (cur != r->dev_device) GF_HIP(hipSetDevice(r->dev_device));
    GF_HIP(r->sfs_stream.create());
    gf::StreamScope scope(r->sfs_stream);
    const int rc = gf::sfs_of_device_row(r->dev_freq.p + (size_t)pop * r->dev_pitch, x.num_mutations, n, num_sites, out,
                                         r->sfs_stream, &r->sfs_scratch);
    if (cur != r->dev_device) (void)hipSetDevice(cur);
    return rc;
}

int32_t gf_result_joint_site_frequency_spectrum(const gf_result *r, int32_t s, int32_t pa, int32_t pb, int32_t na,
                                                int32_t nb, float num_sites, float *out) {
    if (!r || !out || s < 0 || s >= (int32_t)r->samples.size() || pa < 0 || pa >= r->P || pb < 0 || pb >= r->P ||
        na <= 0 || nb <= 0) {
        gf::set_error("joint_site_frequency_spectrum: invalid arguments");
        return GF_ERR_ARG;
    }
    const gf_sample_rec &x = r->samples[s];
    const int M = x.num_mutations;
    if (!x.on_device)
        return gf_joint_site_frequency_spectrum(x.freq.data() + (size_t)pa * M, x.freq.data() + (size_t)pb * M, M, na, nb,
                                                num_sites, out, r->dev_device);
    int cur = 0;
    GF_HIP(hipGetDevice(&cur));
    if (cur != r->dev_device) GF_HIP(hipSetDevice(r->dev_device));
    int rc;
    {
        gf::OwnedStream st;
        GF_HIP(st.create());
        gf::StreamScope scope(st);
        rc = gf::joint_sfs_of_device_rows(r->dev_freq.p + (size_t)pa * r->dev_pitch,
                                          r->dev_freq.p + (size_t)pb * r->dev_pitch, M, na, nb, num_sites, out, st);
    }
    if (cur != r->dev_device) (void)hipSetDevice(cur);
    return rc;
}

int32_t gf_result_num_ids(const gf_result *r) { return r ? r->num_ids : GF_ERR_ARG; }

int32_t gf_result_ids(const gf_result *r, gf_mutid *out) {
    if (!r || !out) return GF_ERR_ARG;
    if (r->num_ids > 0 && r->dev_ids.p)
        return result_readback(r, out, r->dev_ids.p, (size_t)r->num_ids * sizeof(gf_mutid));
    if (!r->ids.empty()) memcpy(out, r->ids.data(), r->ids.size() * sizeof(gf_mutid));
    return GF_OK;
}

int32_t gf_result_live_counts(const gf_result *r, int64_t *out) {
    if (!r || !out) return GF_ERR_ARG;
    memcpy(out, r->live.data(), r->live.size() * sizeof(int64_t));
    return GF_OK;
}

int32_t gf_result_timing(const gf_result *r, gf_timing *out) {
    if (!r || !out) return GF_ERR_ARG;
    *out = r->timing;
    return GF_OK;
}

void gf_result_free(gf_result *r) { delete r; }

} /* extern "C" */
