/*
 * spectrum_api.cpp -- C++ Spectrum API (include/spectrum.h) on top of the
 * C ABI.  Reference: 3P/_internal/spectrum.cu:147-294 (argument checks and
 * their messages, SFS field conventions).  Reusing an SFS frees the arrays
 * of its previous contents (the reference overwrote the pointers).
 */
#include "../../include/spectrum.h"
#include "../../include/gofish_capi.h"

#include <cstdio>
#include <cstdlib>

namespace Spectrum_details {

class transfer_allele_trajectories {
public:
    struct view {
        const float *row;
        int num_mutations, Nchrom_e, sampled_generation;
        float num_sites;
    };
    static view get(const GO_Fish::allele_trajectories &a, int sample_index, int population_index) {
        if (!a.time_samples || a.num_samples == 0) {
            std::fprintf(stderr, "error transferring allele_trajectories to spectrum: empty allele_trajectories\n");
            std::exit(1);
        }
        const int P = a.sim_run_constants.num_populations;
        if (!(sample_index >= 0 && sample_index < a.num_samples) || !(population_index >= 0 && population_index < P)) {
            std::fprintf(stderr, "site_frequency_spectrum error: requested indices out of bounds: sample %d [0 %d), population %d [0 %d)\n",
                         sample_index, a.num_samples, population_index, P);
            std::exit(1);
        }
        const auto *t = a.time_samples[sample_index];
        view v;
        v.num_mutations = t->num_mutations;
        v.row = t->mutations_freq + (size_t)population_index * t->num_mutations;
        v.Nchrom_e = t->Nchrom_e[population_index];
        v.sampled_generation = t->sampled_generation;
        v.num_sites = a.sim_run_constants.num_sites;
        return v;
    }
};

static void fill(Spectrum::SFS &s, float *spectrum, int n, int population_index, const transfer_allele_trajectories::view &v) {
    delete[] s.frequency_spectrum;
    delete[] s.sample_size;
    delete[] s.populations;
    s.frequency_spectrum = spectrum;
    s.num_populations = 1;
    s.sample_size = new int[1];
    s.sample_size[0] = n;
    s.num_sites = v.num_sites;
    s.num_mutations = s.num_sites - s.frequency_spectrum[0];
    s.populations = new int[1];
    s.populations[0] = population_index;
    s.sampled_generation = v.sampled_generation;
}

} /* namespace Spectrum_details */

namespace Spectrum {

SFS::SFS() : frequency_spectrum(NULL), populations(NULL), sample_size(NULL), num_populations(0), num_sites(0),
             num_mutations(0), sampled_generation(0) {}

SFS::~SFS() {
    delete[] frequency_spectrum;
    delete[] populations;
    delete[] sample_size;
    frequency_spectrum = NULL;
    populations = NULL;
    sample_size = NULL;
}

void population_frequency_histogram(SFS &mySFS, const GO_Fish::allele_trajectories &all_results, const int sample_index,
                                    const int population_index, int cuda_device) {
    using namespace Spectrum_details;
    const transfer_allele_trajectories::view v = transfer_allele_trajectories::get(all_results, sample_index, population_index);
    const int Nc = v.Nchrom_e;
    if (Nc <= 0) {
        std::fprintf(stderr, "population_frequency_histogram error: population %d has no chromosomes in sample %d\n", population_index, sample_index);
        std::exit(1);
    }
    float *h = new float[Nc];
    if (gf_population_histogram(v.row, v.num_mutations, Nc, v.num_sites, h, cuda_device) != GF_OK) {
        std::fprintf(stderr, "%s\n", gf_last_error());
        std::exit(1);
    }
    fill(mySFS, h, Nc, population_index, v);
}

void site_frequency_spectrum(SFS &mySFS, const GO_Fish::allele_trajectories &all_results, const int sample_index,
                             const int population_index, const int sample_size, int cuda_device) {
    using namespace Spectrum_details;
    const transfer_allele_trajectories::view v = transfer_allele_trajectories::get(all_results, sample_index, population_index);
    if (sample_size <= 0 || sample_size >= v.Nchrom_e) {
        std::fprintf(stderr, "site_frequency_spectrum error: requested sample_size out of range [1,population_size): sample_size %d [1,%d)",
                     sample_size, v.Nchrom_e);
        std::exit(1);
    }
    float *s = new float[sample_size];
    if (gf_site_frequency_spectrum(v.row, v.num_mutations, sample_size, v.num_sites, s, cuda_device) != GF_OK) {
        std::fprintf(stderr, "%s\n", gf_last_error());
        std::exit(1);
    }
    fill(mySFS, s, sample_size, population_index, v);
}

void joint_site_frequency_spectrum(SFS &mySFS, const GO_Fish::allele_trajectories &all_results, const int sample_index,
                                   const int population_a, const int population_b, const int sample_size_a,
                                   const int sample_size_b, int cuda_device) {
    using namespace Spectrum_details;
    const transfer_allele_trajectories::view va = transfer_allele_trajectories::get(all_results, sample_index, population_a);
    const transfer_allele_trajectories::view vb = transfer_allele_trajectories::get(all_results, sample_index, population_b);
    if (sample_size_a <= 0 || sample_size_a >= va.Nchrom_e || sample_size_b <= 0 || sample_size_b >= vb.Nchrom_e) {
        std::fprintf(stderr, "joint_site_frequency_spectrum error: requested sample sizes out of range: sample_size %d [1,%d), sample_size %d [1,%d)\n",
                     sample_size_a, va.Nchrom_e, sample_size_b, vb.Nchrom_e);
        std::exit(1);
    }
    const size_t B = (size_t)(sample_size_a + 1) * (size_t)(sample_size_b + 1);
    float *j = new float[B];
    if (gf_joint_site_frequency_spectrum(va.row, vb.row, va.num_mutations, sample_size_a, sample_size_b, va.num_sites, j,
                                         cuda_device) != GF_OK) {
        std::fprintf(stderr, "%s\n", gf_last_error());
        std::exit(1);
    }
    fill(mySFS, j, sample_size_a, population_a, va);
    mySFS.num_populations = 2;
    delete[] mySFS.sample_size;
    delete[] mySFS.populations;
    mySFS.sample_size = new int[2];
    mySFS.sample_size[0] = sample_size_a;
    mySFS.sample_size[1] = sample_size_b;
    mySFS.populations = new int[2];
    mySFS.populations[0] = population_a;
    mySFS.populations[1] = population_b;
}

} /* namespace Spectrum */
