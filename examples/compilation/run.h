/* run.h -- plain C++ interface between the hipcc-compiled simulation
 * (run.hip) and the g++-compiled driver (main.cpp) */
#ifndef RUN_SIM_H_
#define RUN_SIM_H_
#include "go_fish_data_struct.h"
void run_migration_equilibrium_simulation(GO_Fish::allele_trajectories &a, float migration_rate);
#endif
