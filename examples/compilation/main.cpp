/*
 * main.cpp -- compiled by g++ (no GPU headers): drives the simulation in
 * run.hip through run.h, copies and prints the allele_trajectories, and
 * writes the full text dump to bfile.dat.
 */
#include "run.h"
#include <fstream>
#include <iostream>

int main(int argc, char **argv) {
    GO_Fish::allele_trajectories a;
    a.sim_input_constants.seed1 = 0xbeeff00d;
    a.sim_input_constants.seed2 = 0xdecafbad;
    const float migration_rate = 0.005f;
    run_migration_equilibrium_simulation(a, migration_rate);
    GO_Fish::allele_trajectories b = a;
    a.reset();
    std::ofstream outfile(argc > 1 ? argv[1] : "bfile.dat");
    outfile << b;
    outfile.close();
    std::cout << std::endl << "number of time samples: " << b.num_time_samples();
    std::cout << std::endl << "mutations in first time sample: " << b.num_mutations_time_sample(0) << std::endl
              << "mutations in final time sample: " << b.maximal_num_mutations() << std::endl;
    std::cout << std::endl << "mutations from burn-in simulation\n";
    std::cout << "ID\tstart frequency pop 1\tstart frequency pop 2\tfinal frequency pop 1\tfinal frequency pop 2" << std::endl;
    for (int i = 0; i < 10 && i < b.maximal_num_mutations(); i++)
        std::cout << b.mutation_ID(i) << "\t" << b.frequency(0, 0, i) << "\t" << b.frequency(0, 1, i) << "\t"
                  << b.frequency(1, 0, i) << "\t" << b.frequency(1, 1, i) << std::endl;
    std::cout << std::endl << "mutations from scenario simulation\n";
    std::cout << "ID\tstart frequency pop 1\tstart frequency pop 2\tfinal frequency pop 1\tfinal frequency pop 2" << std::endl;
    const int b0 = b.num_mutations_time_sample(0);
    for (int i = b0; i < b0 + 10 && i < b.maximal_num_mutations(); i++)
        std::cout << b.mutation_ID(i) << "\t" << b.frequency(0, 0, i) << "\t" << b.frequency(0, 1, i) << "\t"
                  << b.frequency(1, 0, i) << "\t" << b.frequency(1, 1, i) << std::endl;
    return 0;
}
