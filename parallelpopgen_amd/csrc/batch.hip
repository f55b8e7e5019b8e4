/*
 * batch.hip -- batched simulation: thousands of independent single-
 * population simulations advanced together, each launch covering every
 * live simulation (no reference counterpart; it replaces the back-to-back
 * run_sim + site_frequency_spectrum loop of examples/example_dadi/run.cu:
 * 83-94, which recreates streams and buffers and synchronises the host
 * several times per simulation).
 *
 * All simulations share one structure-of-arrays buffer; simulation s owns
 * the slot range [base_s, base_s + cap_s), 256-slot aligned, so every
 * 256-slot tile belongs to exactly one simulation and per-simulation
 * parameters are wave-uniform scalar loads.  The per-generation kernels are
 * the plain path's (k_advance, k_flag, k_scan_tiles, k_scatter, k_mse_*,
 * k_sfs) driven through per-simulation offset tables, and the RNG is keyed
 * by (seed, generation, mutation ID), so every simulation's result is bit
 * for bit the result gf_run gives for it alone.
 *
 * Per block of up to min(compact_interval, 64) generations the host:
 *   1. flags survivors of the current state (segmented) and draws the next
 *      block's new-mutation counts on the device,
 *   2. reads back the per-simulation survivor and new-mutation totals (the
 *      only synchronisation of the block),
 *   3. scatters survivors into a fresh layout sized for the block,
 *   4. finalises simulations that reached their last generation (SFS),
 *   5. launches one k_advance over all live simulations.
 */
#include "gf_kernels.hpp"
#include "gf_runtime.hpp"
#include "gf_spectrum.hpp"

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

struct gf_batch_result {
    int32_t S = 0, n = 0;
    std::vector<float> sfs;                  /* [S][n] */
    std::vector<int32_t> num_mutations;      /* final compacted count */
    std::vector<std::vector<float>> freq;    /* optional final state */
    std::vector<std::vector<gf_mutid>> ids;
    int64_t live_mutation_generations = 0;
    gf_timing timing{};
};

namespace gf {

struct BatchSimDev {
    int Nc, N, G, neutral;
    float F, h, mu, s, L, c1, c2, newfreq, fN;
    uint32_t k0, k1;
    int pad;
};

/* new-mutation counts and model rows of the next block for every active
 * simulation (one thread per simulation; sequential cumulative counts) */
static __global__ void k_batch_rows(const BatchSimDev *__restrict__ sims, const int *__restrict__ active, int nact,
                                    int g0, int K, GenRow<1> *__restrict__ rows, int *__restrict__ tot) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nact) return;
    const BatchSimDev S = sims[active[a]];
    const int Ks = min(K, S.G - g0);
    int cum = 0;
    for (int t = 0; t < Ks; t++) {
        const int g = g0 + 1 + t;
        GenRow<1> r;
        r.Nc[0] = S.Nc;
        r.fN[0] = S.fN;
        r.invN[0] = S.Nc > 0 ? 1.0 / (double)S.fN : 0.0;
        r.sa[0] = 0.0f;
        r.sb[0] = S.s;
        r.c1[0] = S.c1;
        r.c2[0] = S.c2;
        r.mig[0][0] = 1.0f;
        r.newfreq[0] = S.newfreq;
        r.newcum[0] = cum;
        cum += new_mutation_count(S.Nc, S.mu, S.L, S.k0, S.k1, g, 0);
        r.newcum[1] = cum;
        r.present_mask = S.N > 0 ? 1 : 0;
        r.neutral_mask = S.neutral ? 1 : 0;
        rows[(size_t)a * K + t] = r;
    }
    tot[a] = cum;
}

struct BatchRun {
    const gf_batch_sim *in;
    int S, interval, n;
    bool keep;
    gf_batch_result *R;
    OwnedStream st;   /* before the buffers: they are released while it is alive */
    std::vector<BatchSimDev> sims;
    DevBuf<BatchSimDev> d_sims;
    DevBuf<float> freqA, freqB;
    DevBuf<gf_mutid> idsA, idsB;
    int64_t cap = 0;
    /* layout */
    std::vector<int64_t> base, ncap;
    std::vector<int> M;
    std::vector<int> G;
    KernelTimer t_adv, t_cmp;

    int ensure_total(int64_t need) {
        if (need <= cap && freqA.p) return GF_OK;
        /* only called when the A buffers hold no live data */
        int64_t c = std::max<int64_t>({need, cap + cap / 4, (int64_t)1 << 16});
        c = (c + 255) / 256 * 256;
        GF_TRY(freqA.alloc(c));
        GF_TRY(freqB.alloc(c));
        GF_TRY(idsA.alloc(c));
        GF_TRY(idsB.alloc(c));
        cap = c;
        return GF_OK;
    }

    int ensure_B(int64_t need) {
        if (need <= (int64_t)freqB.n && freqB.p) return GF_OK;
        int64_t c = std::max<int64_t>(need, (int64_t)freqB.n + (int64_t)freqB.n / 4);
        c = (c + 255) / 256 * 256;
        freqB.release();
        idsB.release();
        GF_TRY(freqB.alloc(c));
        GF_TRY(idsB.alloc(c));
        return GF_OK;
    }

    int init() {
        sims.resize(S);
        for (int s = 0; s < S; s++) {
            const gf_batch_sim &b = in[s];
            BatchSimDev &d = sims[s];
            memset(&d, 0, sizeof d);
            const int N1 = b.N1 > 0 ? b.N1 : b.N;
            d.N = N1; d.F = b.F; d.h = b.h; d.mu = b.mu; d.s = b.s; d.L = b.num_sites;
            d.G = b.num_generations;
            d.k0 = (uint32_t)b.seed1; d.k1 = (uint32_t)b.seed2;
            d.Nc = N1 > 0 ? nchrom(N1, b.F) : 0;
            d.fN = (float)d.Nc;
            d.c1 = (b.F + b.h) - (b.h * b.F);
            d.c2 = (b.F + (2.0f * b.h)) - ((2.0f * b.h) * b.F);
            d.newfreq = d.Nc > 0 ? 1.0f / (float)d.Nc : 0.0f;
            d.neutral = b.s == 0.0f;
            if (b.N < 0 || N1 < 0 || b.F < 0 || b.F > 1 || b.mu < 0 || b.num_generations < 0) {
                set_error("gf_batch_run: simulation " + std::to_string(s) + " has invalid parameters");
                return GF_ERR_PARAM;
            }
        }
        GF_TRY(d_sims.alloc(S));
        GF_HIP(hipMemcpyAsync(d_sims.p, sims.data(), S * sizeof(BatchSimDev), hipMemcpyHostToDevice, st));
        G.resize(S);
        for (int s = 0; s < S; s++) G[s] = in[s].num_generations;

        /* ----- mutation-selection equilibrium of every simulation ----- */
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
            m.k0 = sims[s].k0; m.k1 = sims[s].k1; m.pop = 0;
            m.neutral = sims[s].neutral; m.preserve = 0;
            m.class_off = coff; m.area_off = aoff; m.chunk_off = choff;
            coff += m.Nc - 1; aoff += m.Nc; choff += (m.Nc + MSE_CHUNK - 1) / MSE_CHUNK;
            if (!m.neutral) any_sel = true;
            segs.push_back(m);
            seg_sim.push_back(s);
        }
        base.assign(S, 0);
        ncap.assign(S, 0);
        M.assign(S, 0);
        const int ns = (int)segs.size();
        if (ns == 0) return ensure_total(256);
        std::vector<int64_t> offs(3 * (size_t)ns);
        for (int k = 0; k < ns; k++) { offs[k] = segs[k].class_off; offs[ns + k] = segs[k].area_off; offs[2 * ns + k] = segs[k].chunk_off; }
        DevBuf<MseSeg> d_segs;
        DevBuf<int64_t> d_offs;
        GF_TRY(d_segs.alloc(ns));
        GF_TRY(d_offs.alloc(3 * (size_t)ns));
        GF_HIP(hipMemcpyAsync(d_segs.p, segs.data(), ns * sizeof(MseSeg), hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(d_offs.p, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
        const unsigned grid = (unsigned)std::max(1, num_cus() * 8);
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
        DevBuf<int64_t> d_scan, d_scratch, d_total;
        GF_TRY(d_counts.alloc((size_t)coff));
        GF_TRY(d_scan.alloc((size_t)coff + 1));
        GF_TRY(d_total.alloc(1));
        hipLaunchKernelGGL(k_mse_counts, dim3(grid), dim3(256), 0, st, d_segs.p, d_offs.p, ns, coff,
                           any_sel ? d_area.p : nullptr, d_counts.p);
        GF_HIP(hipGetLastError());
        GF_TRY(scan_counts(d_counts.p, coff, d_scan.p, d_total.p, st, d_scratch));
        GF_HIP(hipMemcpyAsync(d_scan.p + coff, d_total.p, sizeof(int64_t), hipMemcpyDeviceToDevice, st));
        /* per-segment first offsets: gather with a copy per segment would be
         * S round trips; copy the whole scan instead */
        std::vector<int64_t> scan((size_t)coff + 1);
        GF_HIP(hipMemcpyAsync(scan.data(), d_scan.p, scan.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        std::vector<int64_t> out_offs(ns);
        int64_t b = 0;
        for (int k = 0; k < ns; k++) {
            const int s = seg_sim[k];
            const int64_t first = scan[segs[k].class_off];
            const int64_t last = scan[segs[k].class_off + segs[k].Nc - 1];
            const int64_t tot = last - first;
            if (tot > 2000000000LL) { set_error("gf_batch_run: initial mutation count exceeds 2e9"); return GF_ERR_UNSUPPORTED; }
            M[s] = (int)tot;
            base[s] = b;
            ncap[s] = (tot + 255) / 256 * 256;
            segs[k].slot_base = b;
            segs[k].sim_base = b;
            out_offs[k] = first;   /* position in the global scan */
            b += ncap[s];
        }
        GF_TRY(ensure_total(b + 256));
        /* the fill kernel maps output index j (global scan space) to segment
         * k by out_offs and writes slot slot_base + (j - out_offs[k]) */
        GF_HIP(hipMemcpyAsync(d_segs.p, segs.data(), ns * sizeof(MseSeg), hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(d_offs.p, out_offs.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, st));
        const int64_t total = scan[coff];
        if (total > 0)
            hipLaunchKernelGGL(k_mse_fill<1>, dim3(grid), dim3(256), 0, st, d_segs.p, d_offs.p, ns, total, d_scan.p,
                               freqA.p, (int64_t)cap, idsA.p);
        GF_HIP(hipGetLastError());
        GF_HIP(hipStreamSynchronize(st));
        return GF_OK;
    }

    int run() {
        R->S = S;
        R->n = n;
        R->sfs.assign((size_t)S * n, 0.0f);
        R->num_mutations.assign(S, 0);
        if (keep) { R->freq.resize(S); R->ids.resize(S); }
        t_adv.stride = t_cmp.stride = timing_stride();
        auto t0 = std::chrono::steady_clock::now();
        GF_HIP(st.create());
        StreamScope scope(st);
        hipEvent_t i0, i1;
        GF_HIP(hipEventCreate(&i0));
        GF_HIP(hipEventCreate(&i1));
        GF_HIP(hipEventRecord(i0, st));
        GF_TRY(init());
        GF_HIP(hipEventRecord(i1, st));

        const int kb = std::max(1, std::min({KMAX, interval > 0 ? interval : KMAX, generations_per_launch(16)}));
        DevBuf<unsigned long long> d_live;   /* 64 partial counters, summed at the end */
        GF_TRY(d_live.alloc(64));
        DevBuf<unsigned int> d_ticket;
        GF_TRY(d_ticket.alloc(1));
        GF_HIP(hipMemsetAsync(d_live.p, 0, 64 * sizeof(unsigned long long), st));

        /* device scratch */
        DevBuf<int> d_tile_seg, d_meta, d_active, d_tot;
        DevBuf<int64_t> d_base64;
        DevBuf<unsigned long long> d_masks, d_acc;
        DevBuf<unsigned int> d_tcount, d_toff;
        DevBuf<GenRow<1>> d_rows;
        DevBuf<SegLaunch<1>> d_sl;
        DevBuf<float> d_lc;
        PinBuf<int> h_i;
        PinBuf<int64_t> h_64;
        PinBuf<SegLaunch<1>> h_sl;

        std::vector<int> layout;   /* simulations currently in the A buffers */
        for (int s = 0; s < S; s++) if (G[s] >= 0) layout.push_back(s);
        std::vector<int> Mend(M);
        bool flags_ready = false;   /* d_masks / d_tcount hold the current layout's flags */
        int g = 0;
        std::vector<char> finished(S, 0);
        for (;;) {
            /* simulations that still have generations to run */
            std::vector<int> active;
            for (int s : layout) if (G[s] > g) active.push_back(s);
            const int nl = (int)layout.size();
            const int na = (int)active.size();
            const int K = kb;
            /* ---- 1. segmented flag over the current layout (the previous
             * block's advance already wrote the flags; only the first block
             * runs the flag pass) ---- */
            std::vector<int> tile_first(nl), tile_num(nl);
            int ntiles = 0;
            for (int k = 0; k < nl; k++) {
                const int s = layout[k];
                tile_first[k] = ntiles;
                const int nt = (Mend[s] + WG - 1) / WG;
                tile_num[k] = nt;
                ntiles += nt;
            }
            /* pinned host staging: [tile_first | tile_num | seg_M | present | preserve | total | active | tot] */
            const size_t nmeta = 6 * (size_t)nl + 2 * (size_t)na + 8;
            GF_TRY(h_i.alloc(nmeta));
            GF_TRY(d_meta.alloc(nmeta));
            int *hm = h_i.p;
            int *h_tile_first = hm, *h_tile_num = h_tile_first + nl, *h_M = h_tile_num + nl,
                *h_present = h_M + nl, *h_pres = h_present + nl, *h_total = h_pres + nl, *h_active = h_total + nl,
                *h_tot = h_active + na;
            for (int k = 0; k < nl; k++) {
                const int s = layout[k];
                h_tile_first[k] = tile_first[k];
                h_tile_num[k] = tile_num[k];
                h_M[k] = Mend[s];
                h_present[k] = sims[s].N > 0 ? 1 : 0;
                h_pres[k] = 0;
                h_total[k] = 0;
            }
            for (int a = 0; a < na; a++) { h_active[a] = active[a]; h_tot[a] = 0; }
            const ptrdiff_t o_tile_first = 0, o_tile_num = o_tile_first + nl, o_M = o_tile_num + nl,
                            o_present = o_M + nl, o_pres = o_present + nl, o_total = o_pres + nl,
                            o_active = o_total + nl, o_tot = o_active + na;
            GF_TRY(h_64.alloc(2 * (size_t)nl + 2));
            GF_TRY(d_base64.alloc(2 * (size_t)nl + 2));
            for (int k = 0; k < nl; k++) h_64.p[k] = base[layout[k]];
            GF_HIP(hipMemcpyAsync(d_meta.p, hm, nmeta * sizeof(int), hipMemcpyHostToDevice, st));
            GF_HIP(hipMemcpyAsync(d_base64.p, h_64.p, (size_t)nl * sizeof(int64_t), hipMemcpyHostToDevice, st));
            /* the advance sized d_masks / d_tcount for exactly these tiles
             * and filled them: no reallocation in between */
            if (!flags_ready) {
                GF_TRY(d_masks.alloc((size_t)std::max(ntiles, 1) * (WG / 64)));
                GF_TRY(d_tcount.alloc((size_t)std::max(ntiles, 1)));
            }
            GF_TRY(d_toff.alloc((size_t)std::max(ntiles, 1)));
            hipEvent_t ce = nullptr;
            t_cmp.begin(st, ce);
            if (ntiles > 0) {
                if (!flags_ready)
                    hipLaunchKernelGGL(k_flag<1>, dim3(ntiles), dim3(WG), 0, st, freqA.p, (int64_t)cap, idsA.p,
                                       d_base64.p, d_meta.p + o_M, d_meta.p + o_present, d_meta.p + o_tile_first, nl,
                                       d_masks.p, d_tcount.p);
                hipLaunchKernelGGL(k_scan_tiles, dim3(nl), dim3(1024), 0, st, d_tcount.p, d_meta.p + o_tile_first,
                                   d_meta.p + o_tile_num, d_toff.p, d_meta.p + o_total);
                GF_HIP(hipGetLastError());
            }
            /* ---- 1b. next block's rows and new-mutation totals ---- */
            if (na > 0) {
                GF_TRY(d_rows.alloc((size_t)na * K));
                hipLaunchKernelGGL(k_batch_rows, dim3((na + 127) / 128), dim3(128), 0, st, d_sims.p,
                                   d_meta.p + o_active, na, g, K, d_rows.p, d_meta.p + o_tot);
                GF_HIP(hipGetLastError());
            }
            /* ---- 2. read back survivors and new-mutation totals ---- */
            /* [total (nl) | active (na) | tot (na)] */
            GF_HIP(hipMemcpyAsync(h_total, d_meta.p + o_total, ((size_t)nl + 2 * (size_t)na) * sizeof(int), hipMemcpyDeviceToHost, st));
            GF_HIP(hipStreamSynchronize(st));
            /* ---- 3. new layout: survivors + the block's new mutations ---- */
            std::vector<int64_t> nbase(nl);
            std::vector<int> tot_of(S, 0);
            for (int a = 0; a < na; a++) tot_of[active[a]] = h_tot[a];
            int64_t b = 0;
            for (int k = 0; k < nl; k++) {
                const int s = layout[k];
                const int64_t need = (int64_t)h_total[k] + tot_of[s];
                if (need > 2000000000LL) { set_error("gf_batch_run: simulation exceeds 2e9 slots"); return GF_ERR_UNSUPPORTED; }
                nbase[k] = b;
                b += (need + 255) / 256 * 256;
            }
            GF_TRY(ensure_B(b + 256));
            for (int k = 0; k < nl; k++) h_64.p[nl + k] = nbase[k];
            GF_HIP(hipMemcpyAsync(d_base64.p + nl, h_64.p + nl, (size_t)nl * sizeof(int64_t), hipMemcpyHostToDevice, st));
            if (ntiles > 0) {
                hipLaunchKernelGGL(k_scatter<1>, dim3((ntiles + SCATTER_TILES - 1) / SCATTER_TILES), dim3(WG), 0, st,
                                   freqA.p, (int64_t)cap, idsA.p, freqB.p, (int64_t)freqB.n, idsB.p, d_base64.p,
                                   d_base64.p + nl, d_meta.p + o_pres, d_meta.p + o_tile_first, nl, d_masks.p, d_toff.p,
                                   ntiles);
                GF_HIP(hipGetLastError());
            }
            t_cmp.end(ce);
            freqA.swap(freqB);
            idsA.swap(idsB);
            cap = (int64_t)freqA.n;
            for (int k = 0; k < nl; k++) {
                const int s = layout[k];
                base[s] = nbase[k];
                M[s] = h_total[k];
            }
            /* ---- 4. finalise simulations whose last generation is g ---- */
            std::vector<int> done;
            for (int s : layout) if (G[s] <= g && !finished[s]) done.push_back(s);
            if (!done.empty()) GF_TRY(finalise(done, d_acc, d_lc));
            for (int s : done) finished[s] = 1;
            if (na == 0) break;
            /* ---- 5. one advance launch over every live simulation ---- */
            GF_TRY(h_sl.alloc(na));
            GF_TRY(d_sl.alloc(na));
            std::vector<int> atile_first(na);
            int ntl = 0;
            for (int a = 0; a < na; a++) {
                const int s = active[a];
                SegLaunch<1> &L = h_sl.p[a];
                memset(&L, 0, sizeof L);
                L.rows = d_rows.p + (size_t)a * K;
                L.live = d_live.p;
                L.base = base[s];
                L.M0 = M[s];
                L.Mend = M[s] + tot_of[s];
                L.cum_base = 0;
                L.K = std::min(K, G[s] - g);
                L.g0 = g;
                L.t0 = 0;
                L.k0 = sims[s].k0;
                L.k1 = sims[s].k1;
                atile_first[a] = ntl;
                ntl += (L.Mend + WG - 1) / WG;
                Mend[s] = L.Mend;
            }
            GF_TRY(d_tile_seg.alloc((size_t)na + 1));
            GF_HIP(hipMemcpyAsync(d_sl.p, h_sl.p, na * sizeof(SegLaunch<1>), hipMemcpyHostToDevice, st));
            GF_HIP(hipMemcpyAsync(d_tile_seg.p, atile_first.data(), (size_t)na * sizeof(int), hipMemcpyHostToDevice, st));
            /* the advance writes the next block's compaction flags */
            GF_TRY(d_masks.alloc((size_t)std::max(ntl, 1) * (WG / 64)));
            GF_TRY(d_tcount.alloc((size_t)std::max(ntl, 1)));
            flags_ready = true;
            if (ntl > 0) {
                const int grid = std::min(ntl, num_cus() * 8);
                GF_HIP(hipMemsetAsync(d_ticket.p, 0, sizeof(unsigned int), st));
                hipEvent_t ae = nullptr;
                t_adv.begin(st, ae);
                hipLaunchKernelGGL((k_advance<1, true>), dim3(grid), dim3(WG), 0, st, freqA.p, (int64_t)cap, idsA.p,
                                   h_sl.p[0], (const SegLaunch<1> *)d_sl.p, (const int *)d_tile_seg.p, na, ntl,
                                   d_ticket.p, d_masks.p, d_tcount.p);
                GF_HIP(hipGetLastError());
                t_adv.end(ae);
            }
            /* the host copy of tmeta must outlive the async copy */
            GF_HIP(hipStreamSynchronize(st));
            /* finished simulations leave the layout; a simulation whose last
             * generation falls inside the block ran L.K < K generations and
             * is finalised at the top of the next iteration (G[s] <= g) */
            layout = active;
            g += K;
        }
        GF_HIP(hipStreamSynchronize(st));
        unsigned long long live64[64];
        GF_HIP(hipMemcpy(live64, d_live.p, sizeof live64, hipMemcpyDeviceToHost));
        unsigned long long live = 0;
        for (int k = 0; k < 64; k++) live += live64[k];
        R->live_mutation_generations = (int64_t)live;
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
        R->timing.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return GF_OK;
    }

    /* SFS (and optionally the state) of finished simulations */
    int finalise(const std::vector<int> &done, DevBuf<unsigned long long> &d_acc, DevBuf<float> &d_lc) {
        const int nd = (int)done.size();
        std::vector<int> tseg, tfirst(nd), segM(nd);
        std::vector<int64_t> sbase(nd);
        for (int k = 0; k < nd; k++) {
            const int s = done[k];
            tfirst[k] = (int)tseg.size();
            segM[k] = M[s];
            sbase[k] = base[s];
            for (int q = 0; q < (M[s] + WG - 1) / WG; q++) tseg.push_back(k);
        }
        const int nt = (int)tseg.size();
        std::vector<int> meta(tseg);
        meta.insert(meta.end(), tfirst.begin(), tfirst.end());
        meta.insert(meta.end(), segM.begin(), segM.end());
        DevBuf<int> d_m;
        DevBuf<int64_t> d_b;
        GF_TRY(d_m.alloc(meta.size() + 1));
        GF_TRY(d_b.alloc(nd));
        GF_TRY(d_acc.alloc((size_t)nd * (n + 2)));
        GF_HIP(hipMemcpyAsync(d_m.p, meta.data(), meta.size() * sizeof(int), hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(d_b.p, sbase.data(), nd * sizeof(int64_t), hipMemcpyHostToDevice, st));
        GF_TRY(sfs_device(freqA.p, d_b.p, d_m.p + nt + nd, d_m.p, d_m.p + nt, nt, nd, n, d_acc.p, st, d_lc));
        std::vector<unsigned long long> acc((size_t)nd * (n + 2));
        GF_HIP(hipMemcpyAsync(acc.data(), d_acc.p, acc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        if (keep) {
            for (int k = 0; k < nd; k++) {
                const int s = done[k];
                R->freq[s].resize(M[s]);
                R->ids[s].resize(M[s]);
                if (M[s] > 0) {
                    GF_HIP(hipMemcpyAsync(R->freq[s].data(), freqA.p + base[s], (size_t)M[s] * sizeof(float), hipMemcpyDeviceToHost, st));
                    GF_HIP(hipMemcpyAsync(R->ids[s].data(), idsA.p + base[s], (size_t)M[s] * sizeof(gf_mutid), hipMemcpyDeviceToHost, st));
                }
            }
        }
        GF_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < nd; k++) {
            const int s = done[k];
            R->num_mutations[s] = M[s];
            finish_sfs(&acc[(size_t)k * (n + 2)], acc[(size_t)k * (n + 2) + n + 1], n, M[s], in[s].num_sites,
                       &R->sfs[(size_t)s * n]);
        }
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
    if (!sims || num_sims <= 0 || sfs_n <= 0 || compact_interval < 0) {
        set_error("gf_batch_run: invalid arguments");
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

void gf_batch_free(gf_batch_result *r) { delete r; }

} /* extern "C" */
