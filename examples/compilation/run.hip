/*
 * run.hip -- compiled by hipcc.  Two populations of 1,000 haploid
 * individuals start blank, burn in for 20,000 generations to migration
 * equilibrium with weakly deleterious mutations, then mutations become
 * beneficial in population 1 for 1,000 generations; the second run starts
 * from the burn-in's final sample (prev_sim restart) and samples its first
 * and last generations (scenario of the reference's
 * examples/example_compilation/run.cu:12-42).
 */
#include "go_fish.cuh"
#include "run.h"

void run_migration_equilibrium_simulation(GO_Fish::allele_trajectories &a, float migration_rate) {
    using namespace Sim_Model;
    using namespace GO_Fish;
    a.sim_input_constants.num_sites = 20 * std::pow(10.f, 7);
    a.sim_input_constants.num_populations = 2;
    a.sim_input_constants.init_mse = false;
    F_mu_h_constant mutation(std::pow(10.f, -9));
    F_mu_h_constant inbreeding(1.f);
    demography_constant demography(1000);
    migration_constant_equal migration(migration_rate, a.sim_input_constants.num_populations);
    selection_constant deleterious(-1.f / 1000.f);
    F_mu_h_constant dominance(0.f);
    bool_off dont_preserve;
    bool_off dont_sample;
    a.sim_input_constants.compact_interval = 100;
    a.sim_input_constants.num_generations = 2 * std::pow(10.f, 4);
    run_sim(a, mutation, demography, migration, deleterious, inbreeding, dominance, dont_preserve, dont_sample);

    allele_trajectories c(a);
    bool_on sample;
    bool_pulse<bool_off, bool_on> sample_strategy(dont_sample, sample, 0, a.sim_input_constants.num_generations);
    a.sim_input_constants.num_generations = std::pow(10.f, 3);
    a.sim_input_constants.prev_sim_sample = 0;
    selection_constant beneficial(20.f / 1000.f);
    selection_population_specific<selection_constant, selection_constant> selection_model(deleterious, beneficial, 1);
    run_sim(a, mutation, demography, migration, selection_model, inbreeding, dominance, dont_preserve, sample_strategy, a);
}
