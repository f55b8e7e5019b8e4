/*
 * speed.hip -- GO_Fish speed benchmark scenario on the MI355X build.
 * Same model as the reference's examples/example_speed/run.cu:12-25: one
 * population of 2e5 haploid individuals (F = 1), mu = 1e-9, neutral,
 * 1,000 generations from equilibrium, L = xL * 2e7 sites; 10 warm-up runs,
 * then the mean of 10 timed runs (std::chrono instead of CUDA events).
 *   ./speed [xL] [compact_interval]
 */
#include "go_fish.cuh"
#include <chrono>
#include <cstdlib>
#include <iostream>

int main(int argc, char **argv) {
    const float xL = argc > 1 ? (float)std::atof(argv[1]) : 20.f;
    const int compact = argc > 2 ? std::atoi(argv[2]) : 15;
    GO_Fish::allele_trajectories a;
    a.sim_input_constants.num_populations = 1;
    Sim_Model::F_mu_h_constant mutation(std::pow(10.f, -9));
    Sim_Model::F_mu_h_constant inbreeding(1.f);
    Sim_Model::demography_constant demography(std::pow(10.f, 5) * (1 + inbreeding(0, 0)));
    Sim_Model::migration_constant_equal migration;
    Sim_Model::selection_constant selection(0);
    Sim_Model::F_mu_h_constant dominance(0.f);
    a.sim_input_constants.num_generations = 1000;
    a.sim_input_constants.seed1 = 0xbeeff00d;
    a.sim_input_constants.seed2 = 0xdecafbad;
    a.sim_input_constants.num_sites = xL * 2 * std::pow(10.f, 7);
    a.sim_input_constants.compact_interval = compact;
    const int iters = 20;
    std::chrono::steady_clock::time_point t0;
    for (int i = 0; i < iters; i++) {
        if (i == iters / 2) t0 = std::chrono::steady_clock::now();
        GO_Fish::run_sim(a, mutation, demography, migration, selection, inbreeding, dominance, Sim_Model::bool_off(), Sim_Model::bool_off());
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "number of sites:\t" << a.sim_input_constants.num_sites << std::endl
              << "compact interval:\t" << a.sim_input_constants.compact_interval << std::endl
              << "number of mutations:\t" << a.maximal_num_mutations() << std::endl
              << "time elapsed (ms):\t" << 2 * ms / iters << std::endl;
    return 0;
}
