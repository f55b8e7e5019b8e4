/*
 * custom.hip -- GO_Fish::run_sim with selection functors that have no
 * linear form: the reference's functor contract
 *   __device__ float operator()(int population, int generation, float freq) const
 * (template_inline_simulation_functors.cuh:35-37).  Such a functor is
 * compiled into the drift kernel in this program and reaches the library
 * through gf_run_with_selection.
 *
 * Prints one line per check:
 *   constant  : a functor returning s  == Sim_Model::selection_constant(s)      (bit for bit)
 *   linear    : slope*freq + intercept == selection_linear_frequency_dependent  (bit for bit;
 *               this file is compiled with -ffp-contract=off like the library)
 *   two_pops  : the constant functor over two populations with migration      (bit for bit)
 *   halving   : s = -1 above frequency 1/2 (a mutation there is halved each
 *               generation): largest final frequency, and the neutral run's
 *   joint     : Spectrum::joint_site_frequency_spectrum of the two-population
 *               run (n = 30, 20): total mass over num_sites, and the largest
 *               difference between its population-0 marginal and
 *               Spectrum::site_frequency_spectrum (bins 1..29), relative to
 *               max(1, bin)
 *   ./custom
 */
#include "go_fish.cuh"
#include "spectrum.h"
#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <cstdio>

struct constant_functor {
    float s;
    __host__ __device__ float operator()(int, int, float) const { return s; }
};

struct linear_functor {
    float slope, intercept;
    __host__ __device__ float operator()(int, int, float freq) const { return slope * freq + intercept; }
};

struct halving_functor {
    __host__ __device__ float operator()(int, int, float freq) const { return freq > 0.5f ? -1.0f : 0.0f; }
};

static uint64_t digest(GO_Fish::allele_trajectories &a) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](uint32_t v) { h = (h ^ v) * 1099511628211ull; };
    const int P = a.num_populations();
    for (int s = 0; s < a.num_time_samples(); s++) {
        const int M = a.num_mutations_time_sample(s);
        mix((uint32_t)M);
        for (int p = 0; p < P; p++)
            for (int m = 0; m < M; m++) {
                float f = a.frequency(s, p, m);
                uint32_t u;
                std::memcpy(&u, &f, 4);
                mix(u);
            }
    }
    for (int m = 0; m < a.maximal_num_mutations(); m++) {
        GO_Fish::mutID id = a.mutation_ID(m);
        mix((uint32_t)id.origin_generation);
        mix((uint32_t)id.origin_population);
        mix((uint32_t)id.origin_threadID);
    }
    return h;
}

static float max_frequency(GO_Fish::allele_trajectories &a) {
    float mx = 0.0f;
    const int s = a.num_time_samples() - 1;
    for (int m = 0; m < a.num_mutations_time_sample(s); m++) mx = std::max(mx, a.frequency(s, 0, m));
    return mx;
}

template <typename Sel>
static GO_Fish::allele_trajectories one(int P, Sel sel, int G, float L) {
    GO_Fish::allele_trajectories a;
    a.sim_input_constants.num_populations = P;
    a.sim_input_constants.num_generations = G;
    a.sim_input_constants.num_sites = L;
    a.sim_input_constants.seed1 = 0x1234;
    a.sim_input_constants.seed2 = 0x5678;
    a.sim_input_constants.compact_interval = 20;
    GO_Fish::run_sim(a, Sim_Model::F_mu_h_constant(2e-9f), Sim_Model::demography_constant(5000),
                     Sim_Model::migration_constant_equal(P > 1 ? 0.01f : 0.0f, P), sel,
                     Sim_Model::F_mu_h_constant(0.0f), Sim_Model::F_mu_h_constant(0.3f), Sim_Model::bool_off(),
                     Sim_Model::bool_off());
    return a;
}

int main() {
    {
        auto a = one(1, Sim_Model::selection_constant(-0.004f), 300, 2e7f);
        auto b = one(1, constant_functor{-0.004f}, 300, 2e7f);
        std::printf("constant\t%016llx\t%016llx\t%d\n", (unsigned long long)digest(a), (unsigned long long)digest(b),
                    a.maximal_num_mutations());
    }
    {
        auto a = one(1, Sim_Model::selection_linear_frequency_dependent(0.01f, -0.003f), 300, 2e7f);
        auto b = one(1, linear_functor{0.01f, -0.003f}, 300, 2e7f);
        std::printf("linear\t%016llx\t%016llx\t%d\n", (unsigned long long)digest(a), (unsigned long long)digest(b),
                    a.maximal_num_mutations());
    }
    {
        auto a = one(2, Sim_Model::selection_constant(0.002f), 200, 1e7f);
        auto b = one(2, constant_functor{0.002f}, 200, 1e7f);
        std::printf("two_pops\t%016llx\t%016llx\t%d\n", (unsigned long long)digest(a), (unsigned long long)digest(b),
                    a.maximal_num_mutations());
        const int s_last = a.num_time_samples() - 1;
        Spectrum::SFS joint, marg;
        Spectrum::joint_site_frequency_spectrum(joint, a, s_last, 0, 1, 30, 20);
        Spectrum::site_frequency_spectrum(marg, a, s_last, 0, 30);
        double total = 0.0, worst = 0.0;
        for (int k = 0; k < 31 * 21; k++) total += joint.frequency_spectrum[k];
        for (int ka = 1; ka < 30; ka++) {
            double m = 0.0;
            for (int kb = 0; kb <= 20; kb++) m += joint.frequency_spectrum[ka * 21 + kb];
            /* relative for bins above one site, absolute below: products under 2^-32 truncate to 0 */
            const double r = std::fabs(m - marg.frequency_spectrum[ka]) / std::max(1.0, (double)marg.frequency_spectrum[ka]);
            worst = std::max(worst, r);
        }
        std::printf("joint\t%.7f\t%.3g\t%d\n", total / joint.num_sites, worst, joint.num_populations);
    }
    {
        auto a = one(1, halving_functor{}, 400, 2e8f);
        auto b = one(1, Sim_Model::selection_constant(0.0f), 400, 2e8f);
        std::printf("halving\t%.6f\t%.6f\t%d\n", max_frequency(a), max_frequency(b), a.maximal_num_mutations());
    }
    return 0;
}
