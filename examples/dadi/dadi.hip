/*
 * dadi.hip -- the reference's dadi validation scenario
 * (examples/example_dadi/run.cu:23-117) on the MI355X build:
 * population 0 starts at equilibrium with N_ind = 1e4, doubles at
 * generation 1; population 1 splits off at generation 100 with 0.05*N_ind
 * and grows exponentially to 5*N_ind over the next 900 generations;
 * migration 1/(2 N_ind) from generation 101; gamma = -4, h = 0.5, F = 0;
 * 1,001 generations.  Prints, for the n = 1001 SFS of population 1 over
 * `iters` runs, the mean, standard deviation and coefficient of variation
 * of every normalised bin, in the reference's output layout.
 *   ./dadi [num_sites=2e9] [iters=50]
 */
#include "go_fish.cuh"
#include "spectrum.h"
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <vector>

int main(int argc, char **argv) {
    typedef Sim_Model::demography_constant dem_const;
    typedef Sim_Model::demography_population_specific<dem_const, dem_const> dem_pop_const;
    typedef Sim_Model::demography_piecewise<dem_pop_const, dem_pop_const> init_expansion;
    typedef Sim_Model::demography_exponential_growth exp_growth;
    typedef Sim_Model::demography_population_specific<dem_const, exp_growth> dem_pop_const_exp;
    typedef Sim_Model::migration_constant_equal mig_const;
    typedef Sim_Model::migration_constant_directional<mig_const> mig_dir;
    typedef Sim_Model::migration_constant_directional<mig_dir> mig_split;
    typedef Sim_Model::migration_piecewise<mig_const, mig_split> split_pop0;

    const float num_sites = argc > 1 ? (float)std::atof(argv[1]) : 2e9f;
    const int num_iter = argc > 2 ? std::atoi(argv[2]) : 50;
    const float scale_factor = 1.0f;

    GO_Fish::allele_trajectories b;
    b.sim_input_constants.num_populations = 2;
    b.sim_input_constants.num_generations = scale_factor * std::pow(10.f, 3) + 1;
    Sim_Model::F_mu_h_constant codominant(0.5f);
    Sim_Model::F_mu_h_constant outbred(0.f);
    Sim_Model::F_mu_h_constant mutation(std::pow(10.f, -9) / scale_factor);

    const int N_ind = scale_factor * std::pow(10.f, 4);
    dem_const pop0(N_ind);
    dem_const pop1(0);
    dem_pop_const gen0(pop0, pop1, 1);
    dem_const pop0_final(2 * N_ind);
    dem_pop_const gen1(pop0_final, pop1, 1);
    init_expansion gen_0_1(gen0, gen1, 1);
    exp_growth pop1_gen100((std::log(100.f) / (scale_factor * 900.f)), 0.05 * N_ind, scale_factor * 100);
    dem_pop_const_exp gen100(pop0_final, pop1_gen100, 1);
    Sim_Model::demography_piecewise<init_expansion, dem_pop_const_exp> demography_model(gen_0_1, gen100, scale_factor * 100);

    mig_const no_mig_pop0;
    mig_dir no_pop1_gen0(0.f, 1, 1, no_mig_pop0);
    mig_split create_pop1(1.f, 0, 1, no_pop1_gen0);
    split_pop0 migration_split(no_mig_pop0, create_pop1, scale_factor * 100);
    const float mig = 1.f / (2.f * N_ind);
    mig_const mig_prop(mig, b.sim_input_constants.num_populations);
    Sim_Model::migration_piecewise<split_pop0, mig_const> mig_model(migration_split, mig_prop, scale_factor * 100 + 1);

    const float gamma = -4;
    Sim_Model::selection_constant weak_del(gamma, demography_model, outbred);

    b.sim_input_constants.compact_interval = 30;
    b.sim_input_constants.num_sites = num_sites;
    const int sample_size = 1001;
    Spectrum::SFS my_spectra;

    double avg_num_mutations = 0, avg_num_mutations_sim = 0;
    std::vector<std::vector<float> > results(num_iter, std::vector<float>(sample_size));
    std::chrono::steady_clock::time_point t0;
    for (int j = 0; j < num_iter; j++) {
        if (j == num_iter / 2) t0 = std::chrono::steady_clock::now();
        b.sim_input_constants.seed1 = 0xbeeff00d + 2 * j;
        b.sim_input_constants.seed2 = 0xdecafbad - 2 * j;
        GO_Fish::run_sim(b, mutation, demography_model, mig_model, weak_del, outbred, codominant, Sim_Model::bool_off(), Sim_Model::bool_off());
        Spectrum::site_frequency_spectrum(my_spectra, b, 0, 1, sample_size);
        avg_num_mutations += ((double)my_spectra.num_mutations) / num_iter;
        avg_num_mutations_sim += (double)b.maximal_num_mutations() / num_iter;
        for (int i = 0; i < sample_size; i++) results[j][i] = my_spectra.frequency_spectrum[i];
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    std::cout << "SFS :" << std::endl << "allele count\tavg# mutations\tstandard dev\tcoeff of variation (aka relative standard deviation)" << std::endl;
    for (int i = 1; i < sample_size; i++) {
        double avg = 0, sd = 0;
        for (int j = 0; j < num_iter; j++) {
            const double nm = b.num_sites() - results[j][0];
            avg += results[j][i] / (num_iter * nm);
        }
        for (int j = 0; j < num_iter; j++) {
            const double nm = b.num_sites() - results[j][0];
            sd += 1.0 / (num_iter - 1) * std::pow(results[j][i] / nm - avg, 2);
        }
        sd = std::sqrt(sd);
        std::cout << i << "\t" << avg << "\t" << sd << "\t" << (sd / avg) << std::endl;
    }
    std::cout << "\nnumber of sites in simulation: " << b.num_sites() << "\ncompact interval: " << b.last_run_constants().compact_interval;
    std::cout << "\naverage number of mutations in simulation: " << avg_num_mutations_sim
              << "\naverage number of mutations in SFS: " << avg_num_mutations
              << "\ntime elapsed (ms): " << 2 * ms / num_iter << std::endl;
    return 0;
}
