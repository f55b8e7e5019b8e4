/*
 * gofish_capi.h -- C ABI of the MI355X (gfx950) GO_Fish / Spectrum library.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, int status codes.
 * No C++ or GPU-runtime types appear in any signature, so the same header is
 * consumed by the C++ API layer (go_fish.cuh, spectrum.h), by ctypes, and by
 * any other FFI.  Every entry point returns GF_OK (0) or a negative GF_ERR_*
 * code; gf_last_error() returns the message of the last failure on the
 * calling thread.
 *
 * The reference (ParallelPopGen 0.3.2) has no C ABI: its surface is the C++
 * template API that go_fish.cuh/spectrum.h re-declare.  Each entry point
 * below names the reference function(s) whose work it takes over.
 */
#ifndef GOFISH_CAPI_H_
#define GOFISH_CAPI_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----- status codes ----- */
#define GF_OK              0
#define GF_ERR_ARG        -1   /* bad argument (null pointer, size out of range) */
#define GF_ERR_PARAM      -2   /* simulation parameter check failed (check_sim_parameters) */
#define GF_ERR_DEVICE     -3   /* HIP runtime error */
#define GF_ERR_ALLOC      -4   /* device or host allocation failed */
#define GF_ERR_UNSUPPORTED -5  /* valid request this build does not implement */

/* Maximum number of populations one simulation may hold (register-resident
 * per-mutation state in the drift kernel). */
#define GF_MAX_POPULATIONS 8

/* Mutation identifier; same layout and meaning as GO_Fish::mutID
 * (reference 3P/go_fish_data_struct.h:35-51).  A negative origin_threadID
 * marks the mutation as preserved. */
typedef struct gf_mutid {
    int32_t origin_generation;
    int32_t origin_population;
    int32_t origin_threadID;
    int32_t reserved;
} gf_mutid;

/* Value of every host-side model function for one population at one
 * generation.  The selection model is carried in its linear form
 * s(x) = sel_slope * x + sel_intercept, which every Sim_Model selection
 * functor reduces to for a fixed (population, generation)
 * (reference 3P/_internal/template_inline_simulation_functors.cuh:46-158). */
typedef struct gf_pop_gen {
    int32_t N;              /* demography(pop, gen): number of individuals   */
    float   F;              /* inbreeding coefficient FI(pop, gen)           */
    float   h;              /* dominance(pop, gen)                           */
    float   mu;             /* per-site mutation rate mu_rate(pop, gen)      */
    float   sel_slope;      /* selection: s(pop, gen, x) = slope*x + intercept */
    float   sel_intercept;
} gf_pop_gen;

/* Initialisation modes (reference run_sim, go_fish_impl.cuh:593-623). */
#define GF_INIT_MSE       0   /* mutation-selection equilibrium            */
#define GF_INIT_BLANK     1   /* empty mutation array                      */
#define GF_INIT_PREV      2   /* arrays of a previous run's time sample    */

/* Full description of one run_sim call.  Tables are indexed by
 * t = generation - start_generation for t in [0, num_generations]. */
typedef struct gf_sim_desc {
    int32_t num_populations;      /* P, 1..GF_MAX_POPULATIONS */
    float   num_sites;            /* L */
    int32_t seed1, seed2;         /* Philox key */
    int32_t num_generations;      /* generations to simulate after the start */
    int32_t compact_interval;     /* 0 disables compaction */
    int32_t init_mode;            /* GF_INIT_* */
    int32_t start_generation;     /* 0 unless GF_INIT_PREV (sampled generation of prev run) */
    /* GF_INIT_PREV only: */
    int32_t        init_num_mutations;
    const float   *init_freq;     /* [P][init_num_mutations] row-major */
    const gf_mutid*init_ids;      /* [init_num_mutations] */
    const uint8_t *init_extinct;  /* [P] */
    /* model tables */
    const gf_pop_gen *params;     /* [(num_generations+1)][P] */
    const float      *migration;  /* [(num_generations+1)][P from][P to] */
    const uint8_t    *preserve;   /* [(num_generations+1)] preserve_mutations(g) */
    const uint8_t    *sample;     /* [(num_generations+1)] take_sample(g) */
    int32_t device;               /* HIP device ordinal, -1 = current */
} gf_sim_desc;

typedef struct gf_result gf_result;

/* Kernel timing collected while gf_run executes (only when enabled with
 * gf_set_kernel_timing); durations in milliseconds from HIP events recorded
 * on the library's stream around the timed launches (every launch, or every
 * s-th); the average duration of a launch is advance_ms /
 * advance_timed_launches. */
typedef struct gf_timing {
    double  advance_ms;        /* migration_selection_drift + add_new_mutations, timed launches */
    int64_t advance_launches;  /* all launches */
    double  compact_ms;        /* flag + scan + scatter, timed launches */
    int64_t compact_launches;
    double  init_ms;           /* mutation-selection-equilibrium initialisation */
    double  total_ms;          /* whole gf_run, host clock */
    int64_t advance_slot_generations;  /* slot x generation updates issued by the advance kernel */
    int64_t advance_timed_launches;
    int64_t compact_timed_launches;
} gf_timing;

/* Runs one simulation (replaces GO_Fish::run_sim, go_fish_impl.cuh:553-730,
 * and its helpers initialize_mse :242-332, init_blank_prev_run :334-368,
 * compact :370-429, calc_new_mutations_Index :454-467,
 * store_time_sample :477-501; kernels migration_selection_drift :93-124,
 * add_new_mutations go_fish_impl.cu:138-145, flag_segregating_mutations
 * :137-176, scatter_arrays go_fish_impl.cu:147-189). */
int32_t gf_run(const gf_sim_desc *desc, gf_result **out);

/* The parameter checks gf_run makes, for every generation of `desc`, on
 * the host only (no device is touched): GF_OK, or GF_ERR_ARG / GF_ERR_PARAM
 * with the reference's message in gf_last_error() (replaces the
 * per-generation check_sim_parameters, go_fish_impl.cuh:431-452, called at
 * :596 and :647; run_sim calls this before it selects a device, so a bad
 * model fails the same way with or without one). */
int32_t gf_check_sim(const gf_sim_desc *desc);

/* Selection evaluated by code compiled in the caller's program.  The
 * reference takes any __device__ float(int population, int generation,
 * float freq) selection functor and compiles migration_selection_drift
 * with it in the user's translation unit (go_fish_impl.cuh:93-124, called
 * from run_sim :645-714; the functor contract is
 * template_inline_simulation_functors.cuh:35-37).  gf_run lowers selection
 * to a per-(population, generation) linear form; for a functor without
 * one, include/go_fish.cuh instantiates the library's drift kernel with the
 * functor in the caller's program and passes these launchers:
 *   mse_table : fill `table` (device memory, 3*Nc floats) with the
 *               functor at the equilibrium-start points of population
 *               `population` (k_mse_sel_table layout, gf_drift.hpp), on
 *               `stream` (a hipStream_t);
 *   advance   : launch k_advance_fused<num_populations, functor> with the
 *               launch descriptors `fused` (gf::FusedLaunch) and `segment`
 *               (gf::SegLaunch<num_populations>) on `grid` workgroups of
 *               256 threads, on `stream`.
 * Each returns GF_OK or a GF_ERR_* code. */
typedef struct gf_selection_hooks {
    void *user;
    int32_t (*mse_table)(void *user, int32_t population, int32_t generation, int32_t Nc, float *table,
                         void *stream);
    int32_t (*advance)(void *user, int32_t num_populations, const void *fused, const void *segment,
                       int32_t grid, void *stream);
} gf_selection_hooks;

/* gf_run with selection from `hooks` instead of the sel_slope /
 * sel_intercept columns of desc->params (which are then ignored). */
int32_t gf_run_with_selection(const gf_sim_desc *desc, const gf_selection_hooks *hooks, gf_result **out);

/* Result accessors (reference allele_trajectories accessors,
 * inline_go_fish_data_struct.hpp:263-323). */
int32_t gf_result_num_samples(const gf_result *r);
int32_t gf_result_sample_info(const gf_result *r, int32_t sample,
                              int32_t *num_mutations, int32_t *generation,
                              int32_t *Nchrom_e /*[P]*/, uint8_t *extinct /*[P]*/);
int32_t gf_result_sample_freq(const gf_result *r, int32_t sample, float *out /*[P][M]*/);
int32_t gf_result_num_ids(const gf_result *r);
int32_t gf_result_ids(const gf_result *r, gf_mutid *out);
/* live (segregating or preserved) mutation count after each generation's
 * update, t = 0..num_generations (t = 0: the initial state) */
int32_t gf_result_live_counts(const gf_result *r, int64_t *out /*[num_generations+1]*/);
int32_t gf_result_timing(const gf_result *r, gf_timing *out);
void    gf_result_free(gf_result *r);

/* Population frequency histogram of one population row
 * (replaces Spectrum::population_frequency_histogram, spectrum.cu:153-210,
 * kernels population_hist :74-83 and uint_to_float :85-88).
 * freq: host pointer, M values.  out: host pointer, Nchrom_e floats. */
int32_t gf_population_histogram(const float *freq, int32_t num_mutations,
                                int32_t Nchrom_e, float num_sites,
                                float *out, int32_t device);

/* Expected site frequency spectrum of a sample of n chromosomes
 * (replaces Spectrum::site_frequency_spectrum, spectrum.cu:213-294,
 * kernels binom_coeff :90-95 and binom :111-139).  Accumulation is in
 * 2^-32 fixed point, so the result does not depend on thread order.
 * freq: host pointer, M values.  out: host pointer, n floats. */
int32_t gf_site_frequency_spectrum(const float *freq, int32_t num_mutations,
                                   int32_t sample_size, float num_sites,
                                   float *out, int32_t device);

/* gf_site_frequency_spectrum on a row already in device memory, with the
 * spectrum written to device memory, that allocates nothing (no
 * counterpart in the reference: it is for callers that cannot allocate --
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  d_freq: num_mutations floats and d_out: n floats, both on
 * the current device; the kernels run on `stream` (a hipStream_t; NULL is
 * the null stream) and the call returns without waiting for them.  Its
 * scratch is one fixed device array per device (656 KiB), shared by every
 * call on that device: calls that overlap on one device race.  n is at most
 * 65,536 (GF_ERR_ARG beyond).  Same result, bit for bit, as
 * gf_site_frequency_spectrum. */
int32_t gf_site_frequency_spectrum_noalloc(const float *d_freq, int32_t num_mutations, int32_t n,
                                           float num_sites, float *d_out, void *stream);

/* The same spectrum for time sample `sample`, population `population` of a
 * gf_run result, computed where the sample lives: the final sample stays
 * in device memory after gf_run, so it is binned there without a copy to
 * the host and back (the reference copies every sample to the host,
 * go_fish_impl.cuh:477-501, and back to the device for the spectrum,
 * spectrum.cu:213-294).  Requires 1 <= sample_size < the population's
 * chromosome count in that sample. */
int32_t gf_result_site_frequency_spectrum(const gf_result *result, int32_t sample, int32_t population,
                                          int32_t sample_size, float num_sites, float *out);

/* Joint expected SFS of samples of n_a and n_b chromosomes from two
 * populations (no reference counterpart: the reference's spectrum is per
 * population, spectrum.h:23).  out[k_a * (n_b + 1) + k_b], k_a in
 * [0, n_a], k_b in [0, n_b]: the sum over mutations of
 * P(k_a | n_a, x_a) P(k_b | n_b, x_b) (a frequency of 0 or 1 puts all its
 * mass at k = 0 or n); num_sites - num_mutations monomorphic sites are
 * added at [0][0].  Each product is summed in 2^-32 fixed point, so the
 * result does not depend on thread order.  freq_a, freq_b: host pointers,
 * M values each (the two populations' rows of one time sample). */
int32_t gf_joint_site_frequency_spectrum(const float *freq_a, const float *freq_b, int32_t num_mutations,
                                         int32_t n_a, int32_t n_b, float num_sites, float *out, int32_t device);
/* the same for populations pop_a, pop_b of time sample `sample` of a
 * gf_run result, binned where the sample lives (see
 * gf_result_site_frequency_spectrum) */
int32_t gf_result_joint_site_frequency_spectrum(const gf_result *result, int32_t sample, int32_t pop_a, int32_t pop_b,
                                                int32_t n_a, int32_t n_b, float num_sites, float *out);

/* ----- batched simulation (no reference counterpart: replaces the
 * back-to-back run_sim + site_frequency_spectrum loop of
 * examples/example_dadi/run.cu:83-94) ----- */
typedef struct gf_batch_sim {
    int32_t N;                 /* individuals at generation 0 (equilibrium start) */
    int32_t N1;                /* individuals from generation 1 on (<= 0: N); one instantaneous size change */
    float   F, h, mu;
    float   s;                 /* constant selection coefficient */
    float   num_sites;
    int32_t seed1, seed2;
    int32_t num_generations;
} gf_batch_sim;

typedef struct gf_batch_result gf_batch_result;

/* Runs every simulation from mutation-selection equilibrium to its own final
 * generation, one population each, and reports the final SFS of a sample of
 * sfs_n chromosomes per simulation.  Per-simulation results are identical
 * to gf_run on the equivalent gf_sim_desc with any compaction interval > 0.
 * Generations advance in blocks (one captured graph replay each) of at most
 * min(compact_interval, 32, gf_set_generations_per_launch) generations, the
 * length chosen by the library's autotuner; every block compacts, so
 * compact_interval must be > 0 (GF_ERR_ARG otherwise). */
int32_t gf_batch_run(const gf_batch_sim *sims, int32_t num_sims,
                     int32_t compact_interval, int32_t sfs_n,
                     int32_t keep_states, int32_t device, gf_batch_result **out);
int32_t gf_batch_sfs(const gf_batch_result *r, float *out /*[num_sims][sfs_n]*/);
int32_t gf_batch_num_mutations(const gf_batch_result *r, int32_t *out /*[num_sims]*/);
/* final frequencies and IDs of one simulation, compacted (only when
 * gf_batch_run was called with keep_states != 0); out arrays sized by
 * gf_batch_num_mutations */
int32_t gf_batch_sim_state(const gf_batch_result *r, int32_t sim, float *freq, gf_mutid *ids);
int32_t gf_batch_live_mutation_generations(const gf_batch_result *r, int64_t *out);
int32_t gf_batch_timing(const gf_batch_result *r, gf_timing *out);
/* the block length the autotuner settled on, the number of graph replays
 * and the number of re-layouts (simulations that outgrew their tiles) */
int32_t gf_batch_tuning(const gf_batch_result *r, int32_t *generations_per_block, int32_t *blocks,
                        int32_t *relayouts);
/* Scales the tiles each simulation of a batch is first given (default 1;
 * (0, 1]): smaller regions make simulations outgrow them, stall and be laid
 * out again sooner; results are identical with any scale. */
int32_t gf_set_batch_region_scale(float scale);
void    gf_batch_free(gf_batch_result *r);

/* ----- learned surrogate of the sample SFS (no reference counterpart) -----
 * An ensemble of `ensemble` MLPs, 6 inputs -> 4 hidden layers of 512 (GELU)
 * -> 99 softmax bins (normalised SFS of an n = 100 sample, bins 1..99) + a
 * log-total head; BF16 weights, FP32 accumulation, trained by the library's
 * own MFMA forward / backward and Adam kernels.  Inputs per example:
 * log(N1/N0), growth time, log|gamma|, sign(gamma), h, F. */
typedef struct gf_sg_model gf_sg_model;
int32_t gf_sg_create(int32_t ensemble, uint64_t seed, int32_t device, gf_sg_model **out);
void    gf_sg_free(gf_sg_model *model);
/* parameters per ensemble member (floats); get/set copy [ensemble][n] */
int32_t gf_sg_num_params(const gf_sg_model *model);
int32_t gf_sg_get_params(gf_sg_model *model, float *out);
int32_t gf_sg_set_params(gf_sg_model *model, const float *in);
/* one forward + backward pass (no update), batch a multiple of 64:
 * gradients [ensemble][n_params] and per-member loss */
int32_t gf_sg_gradients(gf_sg_model *model, const float *x /*[B][6]*/, const float *y /*[B][99]*/,
                        const float *log_total /*[B]*/, int32_t batch, float total_weight,
                        float *grad_out, float *loss_out);
/* `epochs` passes of Adam over n examples in mini-batches (a multiple of
 * 64) visited in `order` ([epochs][n] row indices, or NULL); mean loss per
 * epoch */
int32_t gf_sg_train(gf_sg_model *model, const float *x, const float *y, const float *log_total, int32_t n,
                    int32_t epochs, int32_t batch, float lr, float total_weight, const int32_t *order,
                    float *epoch_loss);
/* ensemble mean SFS [n][99], mean log-total [n], disagreement [n] */
int32_t gf_sg_predict(gf_sg_model *model, const float *x, int32_t n, float *sfs, float *log_total,
                      float *disagreement);
/* gf_sg_predict with x, sfs, log_total and disagreement in device memory
 * of the current device (log_total and disagreement may be NULL); returns
 * when the outputs are written.  The caller's writes to x must be complete
 * (the model's stream does not wait for other streams). */
int32_t gf_sg_predict_device(gf_sg_model *model, const float *x, int32_t n, float *sfs, float *log_total,
                             float *disagreement);
/* raw logits of every member, n a multiple of 64: [ensemble][n][128] */
int32_t gf_sg_forward(gf_sg_model *model, const float *x, int32_t n, float *logits);

/* ----- diagnostics (no reference counterpart) -----
 * Evaluates the device implementation of one random-number / elementary
 * function on n inputs (host arrays in, host array out), for checking it
 * bit for bit against the host restatement.
 *   PHILOX     in0: uint32 [n][6] = key0 key1 c0 c1 c2 c3   out: uint32 [n][4]
 *   U01        in0: uint32                                  out: float
 *   EXP, LOG, NORMCDFINV, SQRT, ROUND_AWAY: in0 float       out: float
 *   DIV        in0 / in1 (float)                             out: float
 *   BINOM_DRAW in0 uint32 bits, in1 mean, in2 var, in3 N     out: int32
 *   POIS_DRAW  in0 uint32 bits, in1 mean, in2 N              out: int32
 *   COPY       in0 float                                     out: float */
#define GF_MATH_PHILOX      0
#define GF_MATH_U01         1
#define GF_MATH_EXP         2
#define GF_MATH_LOG         3
#define GF_MATH_NORMCDFINV  4
#define GF_MATH_BINOM_DRAW  5
#define GF_MATH_POIS_DRAW   6
#define GF_MATH_SQRT        7
#define GF_MATH_DIV         8
#define GF_MATH_ROUND_AWAY  9
#define GF_MATH_COPY       10   /* out = in0 (float): a plain copy kernel for counter calibration */
#define GF_MATH_F2I        11   /* out (int32) = saturating float -> int conversion of in0 */
int32_t gf_device_math(int32_t op, const void *in0, const void *in1, const void *in2, const void *in3,
                       void *out, int64_t n, int32_t device);

/* Frees the device memory the library keeps for reuse on `device` (-1:
 * the current device).  Released device buffers are kept in a per-device
 * cache of at most 32 GiB instead of being returned with hipFree, which
 * waits for the whole device (no counterpart in the reference, which
 * allocates and frees on every compaction, go_fish_impl.cuh:377-428). */
int32_t gf_release_device_cache(int32_t device);
/* The cache's state on `device` (-1: the current device): out[0] bytes
 * and out[1] blocks kept for reuse, out[2] bytes held from the device (in
 * use + kept), out[3] hipMalloc and out[4] hipFree calls so far (no
 * counterpart in the reference). */
int32_t gf_device_cache_stats(int32_t device, int64_t *out);

/* ----- runtime ----- */
const char *gf_last_error(void);
/* 0: off; 1: every kernel launch is timed; s > 1: every s-th launch (a
 * pair of timing events costs a few microseconds of the stream's time) */
int32_t gf_set_kernel_timing(int32_t enable);
/* Number of generations the drift kernel advances per launch at most
 * (temporal blocking depth); 1 gives one generation per launch, 0 restores
 * the default (one population: 64 while at most 2e5 slots are in use, 16 up to
 * 1.2e6, 10 beyond; several: 64 up to 2e6 slots, 20 beyond, a launch
 * ending at each compaction generation either way; batches: 16). */
int32_t gf_set_generations_per_launch(int32_t k);
int32_t gf_device_count(int32_t *out);

#ifdef __cplusplus
}
#endif

#endif /* GOFISH_CAPI_H_ */
