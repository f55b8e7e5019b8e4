/*
 * go_fish_data_struct.h -- GO_Fish output / control data structures.
 *
 * Source-compatible with the reference's 3P/go_fish_data_struct.h
 * (GO_Fish::mutID, GO_Fish::allele_trajectories and its accessors,
 * operator<<, swap; declarations at go_fish_data_struct.h:35-170).  Plain
 * C++: it can be included from files compiled by g++ as well as hipcc, so
 * simulation results can be handed to code built by another compiler.
 *
 * Error convention of the reference: an out-of-range request prints a
 * message to stderr and calls exit(1).
 */
#ifndef GO_FISH_DATA_H_
#define GO_FISH_DATA_H_

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

namespace Spectrum_details { class transfer_allele_trajectories; }

namespace GO_Fish {

/* ----- sim result output ----- */
/** identifier of a mutation: generation and population of origin, index of
 * the mutation among those born in that generation and population (threadID,
 * never 0; negative while the mutation is flagged preserved), reserved */
struct mutID {
    int origin_generation;
    int origin_population;
    int origin_threadID;
    int reserved;

    inline mutID();
    inline mutID(int origin_generation, int origin_population, int origin_threadID, int reserved);
    inline std::string toString();
    inline std::string toString() const;
};

/** control (sim_input_constants) and output (time samples, mutation IDs) of
 * a GO_Fish simulation run */
struct allele_trajectories {
    struct sim_constants {
        int seed1;            /**< default 0xbeeff00d */
        int seed2;            /**< default 0xdecafbad */
        int num_generations;  /**< default 0 */
        float num_sites;      /**< default 1000 */
        int num_populations;  /**< default 1 */
        bool init_mse;        /**< default true */
        int prev_sim_sample;  /**< default -1 */
        int compact_interval; /**< default 35; 0 turns compaction off */
        int device;           /**< default -1 */
        inline sim_constants();
    };

    sim_constants sim_input_constants;

    inline allele_trajectories();
    inline allele_trajectories(const allele_trajectories &in);
    inline allele_trajectories &operator=(allele_trajectories in);
    inline sim_constants last_run_constants();
    inline int num_sites();
    inline int num_populations();
    inline int num_time_samples();
    inline int maximal_num_mutations();
    inline int num_mutations_time_sample(int sample_index);
    inline int final_generation();
    inline int sampled_generation(int sample_index);
    inline bool extinct(int sample_index, int population_index);
    inline int effective_number_of_chromosomes(int sample_index, int population_index);
    inline float frequency(int sample_index, int population_index, int mutation_index);
    inline mutID mutation_ID(int mutation_index);
    inline void delete_time_sample(int sample_index);
    inline void reset();
    inline ~allele_trajectories();

    friend void swap(allele_trajectories &a, allele_trajectories &b);
    friend std::ostream &operator<<(std::ostream &stream, allele_trajectories &A);
    template <typename Functor_mutation, typename Functor_demography, typename Functor_migration,
              typename Functor_selection, typename Functor_inbreeding, typename Functor_dominance,
              typename Functor_preserve, typename Functor_timesample>
    friend void run_sim(allele_trajectories &all_results, const Functor_mutation mu_rate,
                        const Functor_demography demography, const Functor_migration mig_prop,
                        const Functor_selection sel_coeff, const Functor_inbreeding FI,
                        const Functor_dominance dominance, const Functor_preserve preserve_mutations,
                        const Functor_timesample take_sample, const allele_trajectories &prev_sim);
    friend class Spectrum_details::transfer_allele_trajectories;

private:
    struct time_sample {
        float *mutations_freq;  /* [population][mutation] */
        bool *extinct;
        int *Nchrom_e;
        int num_mutations;
        int sampled_generation;
        time_sample() : mutations_freq(NULL), extinct(NULL), Nchrom_e(NULL), num_mutations(0), sampled_generation(0) {}
        ~time_sample() {
            delete[] mutations_freq;
            delete[] extinct;
            delete[] Nchrom_e;
        }
    };

    inline void initialize_sim_result_vector(int new_length);

    sim_constants sim_run_constants;
    time_sample **time_samples;
    int num_samples;
    mutID *mutations_ID;
    int all_mutations;
};

inline std::ostream &operator<<(std::ostream &stream, const mutID &id);
inline std::ostream &operator<<(std::ostream &stream, allele_trajectories &A);
inline void swap(allele_trajectories &a, allele_trajectories &b);

/* ===================== inline definitions ===================== */

inline mutID::mutID() : origin_generation(0), origin_population(0), origin_threadID(0), reserved(0) {}
inline mutID::mutID(int g, int p, int t, int r) : origin_generation(g), origin_population(p), origin_threadID(t), reserved(r) {}

namespace detail {
inline std::string mutid_string(const mutID &m) {
    std::ostringstream o;
    o << "(" << m.origin_generation << "," << m.origin_population << "," << std::abs(m.origin_threadID) << ","
      << m.reserved << ")";
    return o.str();
}
inline void fail(const char *msg) {
    std::fputs(msg, stderr);
    std::exit(1);
}
} /* namespace detail */

/** "(origin_generation,origin_population,origin_threadID,reserved)", threadID
 * printed without the preservation sign */
inline std::string mutID::toString() { return detail::mutid_string(*this); }
inline std::string mutID::toString() const { return detail::mutid_string(*this); }

inline allele_trajectories::sim_constants::sim_constants()
    : seed1((int)0xbeeff00d), seed2((int)0xdecafbad), num_generations(0), num_sites(1000), num_populations(1),
      init_mse(true), prev_sim_sample(-1), compact_interval(35), device(-1) {}

inline allele_trajectories::allele_trajectories() : time_samples(NULL), num_samples(0), mutations_ID(NULL), all_mutations(0) {}

inline allele_trajectories::allele_trajectories(const allele_trajectories &in)
    : sim_input_constants(in.sim_input_constants), sim_run_constants(in.sim_run_constants), time_samples(NULL),
      num_samples(in.num_samples), mutations_ID(NULL), all_mutations(in.all_mutations) {
    if (all_mutations > 0) {
        mutations_ID = (mutID *)std::malloc(sizeof(mutID) * all_mutations);
        std::memcpy(mutations_ID, in.mutations_ID, sizeof(mutID) * all_mutations);
    }
    if (num_samples > 0) {
        const int P = sim_run_constants.num_populations;
        time_samples = new time_sample *[num_samples];
        for (int s = 0; s < num_samples; s++) {
            const time_sample &src = *in.time_samples[s];
            time_sample *t = new time_sample;
            t->num_mutations = src.num_mutations;
            t->sampled_generation = src.sampled_generation;
            t->Nchrom_e = new int[P];
            t->extinct = new bool[P];
            for (int p = 0; p < P; p++) { t->Nchrom_e[p] = src.Nchrom_e[p]; t->extinct[p] = src.extinct[p]; }
            t->mutations_freq = new float[(size_t)P * src.num_mutations];
            std::memcpy(t->mutations_freq, src.mutations_freq, sizeof(float) * (size_t)P * src.num_mutations);
            time_samples[s] = t;
        }
    }
}

inline allele_trajectories &allele_trajectories::operator=(allele_trajectories in) {
    swap(*this, in);
    return *this;
}

inline allele_trajectories::sim_constants allele_trajectories::last_run_constants() { return sim_run_constants; }
inline int allele_trajectories::num_sites() { return (int)sim_run_constants.num_sites; }
inline int allele_trajectories::num_populations() { return sim_run_constants.num_populations; }
inline int allele_trajectories::num_time_samples() { return num_samples; }
inline int allele_trajectories::maximal_num_mutations() { return all_mutations; }

inline int allele_trajectories::num_mutations_time_sample(int sample_index) {
    if (!time_samples || num_samples == 0) detail::fail("num_mutations error: empty allele_trajectories\n");
    if (sample_index < 0 || sample_index >= num_samples) {
        std::fprintf(stderr, "num_mutations error: requested sample index out of bounds: sample %d [0 %d)\n", sample_index, num_samples);
        std::exit(1);
    }
    return time_samples[sample_index]->num_mutations;
}

inline int allele_trajectories::final_generation() { return sampled_generation(num_samples - 1); }

inline int allele_trajectories::sampled_generation(int sample_index) {
    if (!time_samples || num_samples == 0) detail::fail("sampled_generation error: empty allele_trajectories\n");
    if (sample_index < 0 || sample_index >= num_samples) {
        std::fprintf(stderr, "sampled_generation error: requested sample index out of bounds: sample %d [0 %d)\n", sample_index, num_samples);
        std::exit(1);
    }
    return time_samples[sample_index]->sampled_generation;
}

inline bool allele_trajectories::extinct(int sample_index, int population_index) {
    if (!time_samples || num_samples == 0) detail::fail("extinct error: empty allele_trajectories\n");
    const int P = sim_run_constants.num_populations;
    if (sample_index < 0 || sample_index >= num_samples || population_index < 0 || population_index >= P) {
        std::fprintf(stderr, "extinct error: index out of bounds: sample %d [0 %d), population %d [0 %d)\n", sample_index, num_samples, population_index, P);
        std::exit(1);
    }
    return time_samples[sample_index]->extinct[population_index];
}

inline int allele_trajectories::effective_number_of_chromosomes(int sample_index, int population_index) {
    if (!time_samples || num_samples == 0) detail::fail("effective_number_of_chromosomes error: empty allele_trajectories\n");
    const int P = sim_run_constants.num_populations;
    if (sample_index < 0 || sample_index >= num_samples || population_index < 0 || population_index >= P) {
        std::fprintf(stderr, "effective_number_of_chromosomes error: index out of bounds: sample %d [0 %d), population %d [0 %d)\n", sample_index, num_samples, population_index, P);
        std::exit(1);
    }
    return time_samples[sample_index]->Nchrom_e[population_index];
}

/** frequency of mutation mutation_index in population population_index at
 * time sample sample_index; 0 if the mutation had not arisen by then */
inline float allele_trajectories::frequency(int sample_index, int population_index, int mutation_index) {
    const int P = sim_run_constants.num_populations;
    if (!time_samples || num_samples == 0) detail::fail("frequency error: empty allele_trajectories\n");
    const int M_final = time_samples[num_samples - 1]->num_mutations;
    if (sample_index < 0 || sample_index >= num_samples || population_index < 0 || population_index >= P ||
        mutation_index < 0 || mutation_index >= M_final) {
        std::fprintf(stderr, "frequency error: index out of bounds: sample %d [0 %d), population %d [0 %d), mutation %d [0 %d)\n",
                     sample_index, num_samples, population_index, P, mutation_index, M_final);
        std::exit(1);
    }
    const time_sample &t = *time_samples[sample_index];
    if (mutation_index >= t.num_mutations) return 0;
    return t.mutations_freq[(size_t)population_index * t.num_mutations + mutation_index];
}

/** ID of mutation mutation_index, threadID without the preservation sign */
inline mutID allele_trajectories::mutation_ID(int mutation_index) {
    if (num_samples <= 0 || !time_samples) detail::fail("mutation_ID error: empty allele_trajectories\n");
    if (mutation_index < 0 || mutation_index >= all_mutations) {
        std::fprintf(stderr, "mutation_ID error: requested mutation index out of bounds: mutation %d [0 %d)\n", mutation_index, all_mutations);
        std::exit(1);
    }
    const mutID &m = mutations_ID[mutation_index];
    return mutID(m.origin_generation, m.origin_population, std::abs(m.origin_threadID), m.reserved);
}

/** deletes one time sample; deleting the last remaining one resets the
 * object; deleting the final one shortens maximal_num_mutations() to the
 * new last sample without freeing the ID array */
inline void allele_trajectories::delete_time_sample(int sample_index) {
    if (sample_index < 0 || sample_index >= num_samples) {
        if (!time_samples || num_samples == 0) detail::fail("delete_time_sample error: empty allele_trajectories\n");
        std::fprintf(stderr, "delete_time_sample error: requested sample index out of bounds: sample %d [0 %d)\n", sample_index, num_samples);
        std::exit(1);
    }
    if (num_samples == 1) { reset(); return; }
    delete time_samples[sample_index];
    time_sample **kept = new time_sample *[num_samples - 1];
    for (int s = 0, k = 0; s < num_samples; s++)
        if (s != sample_index) kept[k++] = time_samples[s];
    delete[] time_samples;
    time_samples = kept;
    num_samples--;
    all_mutations = time_samples[num_samples - 1]->num_mutations;
}

inline void allele_trajectories::reset() {
    if (time_samples) {
        for (int s = 0; s < num_samples; s++) delete time_samples[s];
        delete[] time_samples;
    }
    std::free(mutations_ID);
    time_samples = NULL;
    mutations_ID = NULL;
    num_samples = 0;
    all_mutations = 0;
    sim_run_constants = sim_constants();
    sim_input_constants = sim_constants();
}

inline allele_trajectories::~allele_trajectories() { reset(); }

inline void allele_trajectories::initialize_sim_result_vector(int new_length) {
    const sim_constants keep = sim_input_constants;
    reset();
    num_samples = new_length;
    time_samples = new time_sample *[num_samples];
    for (int s = 0; s < num_samples; s++) time_samples[s] = new time_sample;
    sim_input_constants = keep;
    sim_run_constants = keep;
}

inline std::ostream &operator<<(std::ostream &stream, const mutID &id) { return stream << id.toString(); }

/** text dump: run constants, per-sample rows (generation, mutation count,
 * population, Nchrom_e, extinction) and one row of frequencies per mutation
 * (same layout as the reference's bfile.dat) */
inline std::ostream &operator<<(std::ostream &stream, allele_trajectories &A) {
    const allele_trajectories::sim_constants &c = A.sim_run_constants;
    stream << "seed1" << "\t" << c.seed1 << std::endl;
    stream << "seed2" << "\t" << c.seed2 << std::endl;
    stream << "num_generations" << "\t" << c.num_generations << std::endl;
    stream << "num_sites" << "\t" << c.num_sites << std::endl;
    stream << "num_populations" << "\t" << c.num_populations << std::endl;
    stream << "init_mse" << "\t" << c.init_mse << std::endl;
    stream << "prev_sim_sample" << "\t" << c.prev_sim_sample << std::endl;
    stream << "compact_interval" << "\t" << c.compact_interval << std::endl;
    stream << "device" << "\t" << c.device << std::endl << std::endl;
    if (A.num_samples == 0) { stream << "no simulation stored" << std::endl; return stream; }
    const int P = c.num_populations;
    const int S = A.num_samples;
    stream << "time sample:";
    for (int s = 0; s < S; s++) { stream << "\t" << s; for (int k = 1; k < P; k++) stream << "\t"; }
    stream << std::endl << "generation:";
    for (int s = 0; s < S; s++) { stream << "\t" << A.sampled_generation(s); for (int k = 1; k < P; k++) stream << "\t"; }
    stream << std::endl << "number of mutations reported:";
    for (int s = 0; s < S; s++) { stream << "\t" << A.num_mutations_time_sample(s); for (int k = 1; k < P; k++) stream << "\t"; }
    stream << std::endl << "population:";
    for (int s = 0; s < S; s++) for (int p = 0; p < P; p++) stream << "\t" << p;
    stream << std::endl << "effective population size (chromosomes):";
    for (int s = 0; s < S; s++) for (int p = 0; p < P; p++) stream << "\t" << A.effective_number_of_chromosomes(s, p);
    stream << std::endl << "population extinct:";
    for (int s = 0; s < S; s++) for (int p = 0; p < P; p++) stream << "\t" << A.extinct(s, p);
    stream << std::endl;
    if (A.all_mutations == 0) { stream << std::endl << "no mutations stored" << std::endl; return stream; }
    stream << "mutation ID (origin_generation,origin_population,origin_threadID,reserved)";
    for (int s = 0; s < S; s++) for (int p = 0; p < P; p++) stream << "\t" << "frequency";
    stream << std::endl;
    for (int i = 0; i < A.all_mutations; i++) {
        stream << A.mutation_ID(i) << ":";
        for (int s = 0; s < S; s++) for (int p = 0; p < P; p++) stream << "\t" << A.frequency(s, p, i);
        stream << std::endl;
    }
    return stream;
}

inline void swap(allele_trajectories &a, allele_trajectories &b) {
    std::swap(a.sim_input_constants, b.sim_input_constants);
    std::swap(a.sim_run_constants, b.sim_run_constants);
    std::swap(a.all_mutations, b.all_mutations);
    std::swap(a.num_samples, b.num_samples);
    std::swap(a.mutations_ID, b.mutations_ID);
    std::swap(a.time_samples, b.time_samples);
}

} /* namespace GO_Fish */

#endif /* GO_FISH_DATA_H_ */
