/*
 * sim_model.hip -- host-side checks of the C++ API (include/go_fish.cuh,
 * spectrum.h, go_fish_data_struct.h), run by tests/test_sim_model.py.
 *
 *   sim_model tables <dir>   writes, as raw files in <dir>:
 *       three_pop_params.bin  gf_pop_gen[(G+1) * 3]   the three-population
 *       three_pop_mig.bin     float[(G+1) * 3 * 3]    scenario's model tables,
 *       three_pop_pres.bin    uint8[G+1]              built from Sim_Model
 *       three_pop_samp.bin    uint8[G+1]              functors by run_sim's
 *                                                     own table builder
 *       functors.txt          "name generation value" lines of further
 *                             functors (logistic growth, sine-wave
 *                             demography, F/mu/h sine wave and piecewise,
 *                             bool_piecewise, sine-wave selection)
 *   sim_model die <case>     takes one of the API's error exits
 *                            (message on stderr, exit status 1)
 *   sim_model run_then_die <case>
 *                            runs a small simulation on the GPU, then takes
 *                            an error exit that needs a non-empty result
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "go_fish.cuh"
#include "spectrum.h"

using namespace Sim_Model;

namespace {

/* preserve every k-th generation (a caller's own functor) */
struct every_k {
    int k;
    every_k() : k(1) {}
    explicit every_k(int k) : k(k) {}
    bool operator()(int generation) const { return generation % k == 0; }
};

template <typename T>
bool write_raw(const std::string &path, const std::vector<T> &v) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const size_t n = std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
    return n == v.size();
}

/* The three-population scenario: population 0 grows exponentially from
 * 2e4 to 2e5 over generations 0..500 and stays at 2e5; population 1 has a
 * bottleneck of 2e3 in generations [500, 600); population 2 keeps 2e4 with
 * F = 0.1; migration_constant_equal(1e-4, 3); h = 0.3; mu = 1e-9;
 * sine-wave selection of gamma amplitude 5, period 200; preserve every 100
 * generations after generation 0. */
int tables(const std::string &dir) {
    const int N0 = 20000, G = 2000, P = 3;
    typedef demography_piecewise<demography_exponential_growth, demography_constant> growth_t;
    typedef demography_piecewise<demography_piecewise<demography_constant, demography_constant>, demography_constant>
        bottleneck_t;
    typedef demography_population_specific<demography_population_specific<demography_constant, growth_t>, bottleneck_t>
        demography_t;
    const growth_t growth(demography_exponential_growth((float)(std::log(10.0) / 500.0), N0), demography_constant(10 * N0), 500);
    const bottleneck_t bottleneck(
        demography_piecewise<demography_constant, demography_constant>(demography_constant(N0), demography_constant(N0 / 10), 500),
        demography_constant(N0), 600);
    const demography_t demography(
        demography_population_specific<demography_constant, growth_t>(demography_constant(N0), growth, 0), bottleneck, 1);
    const F_mu_h_population_specific<F_mu_h_constant, F_mu_h_constant> inbreeding(F_mu_h_constant(0.0f),
                                                                                  F_mu_h_constant(0.1f), 2);
    const F_mu_h_constant dominance(0.3f), mutation(1e-9f);
    const selection_sine_wave selection(5.0f, (float)(2.0 * M_PI / 200.0), 0.0f, demography, inbreeding);
    const migration_constant_equal migration(1e-4f, P);
    const bool_piecewise<bool_off, every_k> preserve(bool_off(), every_k(100), 1);
    const bool_off sample;
    go_fish_details::sim_tables T;
    go_fish_details::build_tables(T, P, G, 0, mutation, demography, migration, selection, inbreeding, dominance,
                                  preserve, sample);
    if (!write_raw(dir + "/three_pop_params.bin", T.params) || !write_raw(dir + "/three_pop_mig.bin", T.mig) ||
        !write_raw(dir + "/three_pop_pres.bin", T.pres) || !write_raw(dir + "/three_pop_samp.bin", T.samp)) {
        std::fprintf(stderr, "cannot write the tables to %s\n", dir.c_str());
        return 1;
    }

    FILE *f = std::fopen((dir + "/functors.txt").c_str(), "w");
    if (!f) return 1;
    const demography_logistic_growth logistic(0.01f, 1000, 100000, 7);
    const demography_sine_wave dsine(2.9f, 0.1f, 1000, 0.25f, 3);
    const F_mu_h_sine_wave fsine(0.05f, (float)(2.0 * M_PI / 50.0), 0.1f, 0.3f, 5);
    const F_mu_h_piecewise<F_mu_h_constant, F_mu_h_sine_wave> fpiece(F_mu_h_constant(0.1f), fsine, 100, 3);
    /* the reference's documented sampling example: generation 0 and
     * generations [100, 110] (template_inline_simulation_functors.cuh:620-628) */
    typedef bool_pulse<bool_off, bool_on> pulse_on;
    typedef bool_piecewise<pulse_on, bool_on> switch_on;
    typedef bool_piecewise<switch_on, bool_off> switch_off;
    const pulse_on sample_generation_0;
    const switch_on sample_gen_0_100_XX(sample_generation_0, bool_on(), 100);
    const switch_off sample_gen_0_100_110(sample_gen_0_100_XX, bool_off(), 111);
    const selection_sine_wave ssine(0.002f, 0.05f, -0.001f, 0.7f, 11);
    for (int g = 0; g <= 2000; g++) {
        std::fprintf(f, "logistic %d %d\n", g, logistic(0, g));
        std::fprintf(f, "demography_sine %d %d\n", g, dsine(0, g));
        std::fprintf(f, "F_sine %d %.9g\n", g, fsine(0, g));
        std::fprintf(f, "F_piecewise %d %.9g\n", g, fpiece(0, g));
        std::fprintf(f, "bool_piecewise %d %d\n", g, sample_gen_0_100_110(g) ? 1 : 0);
        float sl, ic;
        go_fish_details::lower_selection(ssine, 0, g, sl, ic);
        std::fprintf(f, "selection_sine %d %.9g %.9g\n", g, sl, ic);
    }
    std::fclose(f);
    return 0;
}

/* a small two-population blank-start run whose parts the death cases
 * replace */
template <typename Mu, typename Dem, typename Mig, typename F>
void run_two(GO_Fish::allele_trajectories &a, Mu mu, Dem dem, Mig mig, F inbreeding) {
    a.sim_input_constants.num_sites = 1e6;
    a.sim_input_constants.num_populations = 2;
    a.sim_input_constants.num_generations = 20;
    a.sim_input_constants.init_mse = false;
    GO_Fish::run_sim(a, mu, dem, mig, selection_constant(0.0f), inbreeding, F_mu_h_constant(0.5f), bool_off(), bool_off());
}

/* population 1 at N = 0 (not founded) migrating into population 0 */
struct mig_from_unfounded {
    __host__ __device__ float operator()(int from, int to, int) const {
        if (to == 0) return from == 0 ? 0.9f : 0.1f;
        return from == to ? 1.0f : 0.0f;
    }
};

/* the migration into population 0 sums to 1.1 */
struct mig_sum_wrong {
    __host__ __device__ float operator()(int from, int to, int) const { return to == 0 ? (from == 0 ? 1.0f : 0.1f) : (from == to ? 1.0f : 0.0f); }
};

struct mig_negative {
    __host__ __device__ float operator()(int from, int to, int) const { return to == 0 ? (from == 0 ? 1.1f : -0.1f) : (from == to ? 1.0f : 0.0f); }
};

/* population 1 exists for generations < 10, is extinct at 10 and revives at 15 */
struct dem_revive {
    __host__ __device__ int operator()(int pop, int g) const { return pop == 0 ? 1000 : ((g < 10 || g >= 15) ? 500 : 0); }
};

int die(const std::string &c) {
    GO_Fish::allele_trajectories a;
    const F_mu_h_constant mu(1e-9f), F0(0.0f);
    const demography_constant dem(1000);
    const migration_constant_equal mig(0.01f, 2);
    if (c == "mu_negative") run_two(a, F_mu_h_constant(-1e-9f), dem, mig, F0);
    else if (c == "inbreeding_above_1") run_two(a, mu, dem, mig, F_mu_h_constant(1.5f));
    else if (c == "inbreeding_below_0") run_two(a, mu, dem, mig, F_mu_h_constant(-0.5f));
    else if (c == "migration_negative") run_two(a, mu, dem, mig_negative(), F0);
    else if (c == "migration_sum") run_two(a, mu, dem, mig_sum_wrong(), F0);
    else if (c == "migration_from_unfounded") {
        run_two(a, mu, demography_population_specific<demography_constant, demography_constant>(dem, demography_constant(0), 1),
                mig_from_unfounded(), F0);
    } else if (c == "revived_extinct") {
        run_two(a, mu, dem_revive(), migration_constant_equal(0.0f, 2), F0);
    } else if (c == "num_populations") {
        a.sim_input_constants.num_populations = 9;
        GO_Fish::run_sim(a, mu, dem, migration_constant_equal(0.0f, 9), selection_constant(0.0f), F0,
                         F_mu_h_constant(0.5f), bool_off(), bool_off());
    } else if (c == "prev_sim_empty") {
        a.sim_input_constants.init_mse = false;
        a.sim_input_constants.prev_sim_sample = 0;
        GO_Fish::allele_trajectories empty;
        GO_Fish::run_sim(a, mu, dem, migration_constant_equal(0.0f, 1), selection_constant(0.0f), F0,
                         F_mu_h_constant(0.5f), bool_off(), bool_off(), empty);
    } else if (c == "num_mutations_empty") (void)a.num_mutations_time_sample(0);
    else if (c == "sampled_generation_empty") (void)a.sampled_generation(0);
    else if (c == "extinct_empty") (void)a.extinct(0, 0);
    else if (c == "Nchrom_empty") (void)a.effective_number_of_chromosomes(0, 0);
    else if (c == "frequency_empty") (void)a.frequency(0, 0, 0);
    else if (c == "mutation_ID_empty") (void)a.mutation_ID(0);
    else if (c == "delete_time_sample_empty") a.delete_time_sample(0);
    else if (c == "sfs_empty") {
        Spectrum::SFS s;
        Spectrum::site_frequency_spectrum(s, a, 0, 0, 10);
    } else if (c == "histogram_empty") {
        Spectrum::SFS s;
        Spectrum::population_frequency_histogram(s, a, 0, 0);
    } else {
        std::fprintf(stderr, "unknown case %s\n", c.c_str());
        return 2;
    }
    std::printf("no error exit\n");
    return 0;
}

int run_then_die(const std::string &c) {
    GO_Fish::allele_trajectories a;
    a.sim_input_constants.num_sites = 1e7;
    a.sim_input_constants.num_populations = 1;
    a.sim_input_constants.num_generations = 10;
    GO_Fish::run_sim(a, F_mu_h_constant(1e-9f), demography_constant(500), migration_constant_equal(0.0f, 1),
                     selection_constant(0.0f), F_mu_h_constant(0.0f), F_mu_h_constant(0.5f), bool_off(), bool_off());
    if (a.num_mutations_time_sample(0) <= 0) {
        std::fprintf(stderr, "the run left no mutations\n");
        return 2;
    }
    Spectrum::SFS s;
    if (c == "sfs_sample_size") Spectrum::site_frequency_spectrum(s, a, 0, 0, 1000);   /* Nc = 1000 */
    else if (c == "sfs_indices") Spectrum::site_frequency_spectrum(s, a, 0, 1, 10);
    else if (c == "frequency_bounds") (void)a.frequency(0, 0, a.num_mutations_time_sample(0));
    else if (c == "mutation_ID_bounds") (void)a.mutation_ID(a.maximal_num_mutations());
    else if (c == "extinct_bounds") (void)a.extinct(0, 1);
    else if (c == "prev_sim_mismatch") {
        GO_Fish::allele_trajectories b;
        b.sim_input_constants.num_sites = 2e7;
        b.sim_input_constants.num_generations = 10;
        b.sim_input_constants.init_mse = false;
        b.sim_input_constants.prev_sim_sample = 0;
        GO_Fish::run_sim(b, F_mu_h_constant(1e-9f), demography_constant(500), migration_constant_equal(0.0f, 1),
                         selection_constant(0.0f), F_mu_h_constant(0.0f), F_mu_h_constant(0.5f), bool_off(), bool_off(), a);
    } else if (c == "prev_sim_index") {
        GO_Fish::allele_trajectories b;
        b.sim_input_constants.num_sites = 1e7;
        b.sim_input_constants.num_generations = 10;
        b.sim_input_constants.init_mse = false;
        b.sim_input_constants.prev_sim_sample = 3;
        GO_Fish::run_sim(b, F_mu_h_constant(1e-9f), demography_constant(500), migration_constant_equal(0.0f, 1),
                         selection_constant(0.0f), F_mu_h_constant(0.0f), F_mu_h_constant(0.5f), bool_off(), bool_off(), a);
    } else {
        std::fprintf(stderr, "unknown case %s\n", c.c_str());
        return 2;
    }
    std::printf("no error exit\n");
    return 0;
}

} /* namespace */

int main(int argc, char **argv) {
    if (argc == 3 && std::string(argv[1]) == "tables") return tables(argv[2]);
    if (argc == 3 && std::string(argv[1]) == "die") return die(argv[2]);
    if (argc == 3 && std::string(argv[1]) == "run_then_die") return run_then_die(argv[2]);
    std::fprintf(stderr, "usage: sim_model tables <dir> | die <case> | run_then_die <case>\n");
    return 2;
}
