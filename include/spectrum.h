/*
 * spectrum.h -- Spectrum API (site frequency spectra of GO_Fish results).
 *
 * Source-compatible with the reference's 3P/spectrum.h:24-43.  Like the
 * reference it has no GPU types in it, so it can be included from files
 * compiled by g++; the functions are defined in libgofish.so (they call the
 * gfx950 kernels through gf_population_histogram / gf_site_frequency_spectrum).
 */
#ifndef SPECTRUM_H_
#define SPECTRUM_H_

#include "go_fish_data_struct.h"

namespace Spectrum {

/** site frequency spectrum of one population at one time point */
struct SFS {
    float *frequency_spectrum; /**< bins 0..sample_size-1; bin 0 holds monomorphic sites */
    int *populations;          /**< populations in the SFS */
    int *sample_size;          /**< sample size per population */
    int num_populations;       /**< number of populations in the SFS */
    float num_sites;           /**< number of sites */
    float num_mutations;       /**< number of segregating mutations (num_sites - bin 0) */
    int sampled_generation;    /**< generation of the time sample */

    SFS();
    ~SFS();
};

/** frequency histogram (Nchrom_e bins) of time sample sample_index in population population_index */
void population_frequency_histogram(SFS &mySFS, const GO_Fish::allele_trajectories &all_results, const int sample_index,
                                    const int population_index, int cuda_device = -1);

/** expected SFS of a sample of sample_size chromosomes, 1 <= sample_size < Nchrom_e */
void site_frequency_spectrum(SFS &mySFS, const GO_Fish::allele_trajectories &all_results, const int sample_index,
                             const int population_index, const int sample_size, int cuda_device = -1);

/** joint expected SFS of samples of sample_size_a and sample_size_b chromosomes from populations population_a and
 *  population_b of time sample sample_index (no reference counterpart: the reference's spectrum is per
 *  population).  num_populations = 2; frequency_spectrum has (sample_size_a + 1) * (sample_size_b + 1) bins,
 *  bin [k_a][k_b] at k_a * (sample_size_b + 1) + k_b; bin [0][0] holds the sites monomorphic in both samples. */
void joint_site_frequency_spectrum(SFS &mySFS, const GO_Fish::allele_trajectories &all_results, const int sample_index,
                                   const int population_a, const int population_b, const int sample_size_a,
                                   const int sample_size_b, int cuda_device = -1);

} /* namespace Spectrum */

#endif /* SPECTRUM_H_ */
