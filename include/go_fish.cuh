/*
 * go_fish.cuh -- GO_Fish simulation API for MI355X (HIP, gfx950).
 *
 * Source-compatible with the reference's 3P/go_fish.cuh:
 *   namespace Sim_Model   model functors (reference go_fish.cuh:26-323,
 *                         implementations template_inline_simulation_functors.cuh)
 *   GO_Fish::run_sim      (reference go_fish.cuh:330-334, go_fish_impl.cuh:524-730)
 * A program written against the reference recompiles unchanged with hipcc
 * (-I include, link -lgofish): only the examples' own direct CUDA event calls
 * need their HIP (or std::chrono) spelling.
 *
 * run_sim evaluates the host-side model functors (demography, migration,
 * F, mu, h, preserve, sample) for every generation of the run into tables,
 * lowers the selection functor to its per-(population, generation) linear
 * form s(x) = slope*x + intercept, and hands the tables to the gfx950
 * library through the C ABI (gf_run, include/gofish_capi.h); the results
 * are copied into allele_trajectories.  Every Sim_Model selection functor has
 * that linear form; a user-written selection functor may provide it through a
 *   void linear_form(int population, int generation, float &slope, float &intercept) const
 * member.  Any other selection functor -- the reference's contract,
 *   __device__ float operator()(int population, int generation, float freq) const
 * -- is compiled into the library's drift kernel in the calling program
 * (parallelpopgen_amd/csrc/gf_drift.hpp) and passed to the library through
 * gf_run_with_selection, as the reference compiles migration_selection_drift
 * with the user's functor (go_fish_impl.cuh:93-124).
 */
#ifndef GO_FISH_API_H_
#define GO_FISH_API_H_

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <utility>
#include <vector>

#include "go_fish_data_struct.h"
#include "gofish_capi.h"
#include "../parallelpopgen_amd/csrc/gf_drift.hpp"

namespace Sim_Model {

/* ================= selection ================= */

/** constant selection coefficient s */
struct selection_constant {
    float s;
    inline selection_constant() : s(0) {}
    inline selection_constant(float s) : s(s) {}
    /** effective selection: s = gamma / (2*N(0,g)/(1+F(0,g))) with g = forward_generation_shift */
    template <typename Functor_demography, typename Functor_inbreeding>
    inline selection_constant(float gamma, Functor_demography demography, Functor_inbreeding F, int forward_generation_shift = 0) {
        s = gamma / (2 * demography(0, forward_generation_shift) / (1 + F(0, forward_generation_shift)));
    }
    __host__ __device__ __forceinline__ float operator()(const int population, const int generation, const float freq) const { return s; }
    inline void linear_form(int, int, float &slope, float &intercept) const { slope = 0.0f; intercept = s; }
};

/** s = slope*freq + intercept */
struct selection_linear_frequency_dependent {
    float slope;
    float intercept;
    inline selection_linear_frequency_dependent() : slope(0), intercept(0) {}
    inline selection_linear_frequency_dependent(float slope, float intercept) : slope(slope), intercept(intercept) {}
    template <typename Functor_demography, typename Functor_inbreeding>
    inline selection_linear_frequency_dependent(float gamma_slope, float gamma_intercept, Functor_demography demography, Functor_inbreeding F, int forward_generation_shift = 0) {
        slope = gamma_slope / (2 * demography(0, forward_generation_shift) / (1 + F(0, forward_generation_shift)));
        intercept = gamma_intercept / (2 * demography(0, forward_generation_shift) / (1 + F(0, forward_generation_shift)));
    }
    __host__ __device__ __forceinline__ float operator()(const int population, const int generation, const float freq) const { return slope * freq + intercept; }
    inline void linear_form(int, int, float &sl, float &ic) const { sl = slope; ic = intercept; }
};

/** s = A*sin(pi*(generation - generation_shift) + rho) + D */
struct selection_sine_wave {
    float A;
    float pi;
    float rho;
    float D;
    int generation_shift;
    inline selection_sine_wave() : A(0), pi(0), rho(0), D(0), generation_shift(0) {}
    inline selection_sine_wave(float A, float pi, float D, float rho = 0, int generation_shift = 0)
        : A(A), pi(pi), rho(rho), D(D), generation_shift(generation_shift) {}
    template <typename Functor_demography, typename Functor_inbreeding>
    inline selection_sine_wave(float gamma_A, float pi, float gamma_D, Functor_demography demography, Functor_inbreeding F,
                               float rho = 0, int generation_shift = 0, int forward_generation_shift = 0)
        : pi(pi), rho(rho), generation_shift(generation_shift) {
        A = gamma_A / (2 * demography(0, forward_generation_shift) / (1 + F(0, forward_generation_shift)));
        D = gamma_D / (2 * demography(0, forward_generation_shift) / (1 + F(0, forward_generation_shift)));
    }
    __host__ __device__ __forceinline__ float operator()(const int population, const int generation, const float freq) const {
        return A * sinf(pi * (generation - generation_shift) + rho) + D;
    }
    inline void linear_form(int population, int generation, float &sl, float &ic) const {
        sl = 0.0f;
        ic = A * std::sin(pi * (float)(generation - generation_shift) + rho) + D;
    }
};

/** population pop uses s_pop, all others s */
template <typename Functor_sel, typename Functor_sel_pop>
struct selection_population_specific {
    int pop;
    int generation_shift;
    Functor_sel s;
    Functor_sel_pop s_pop;
    inline selection_population_specific() : pop(0), generation_shift(0), s(), s_pop() {}
    inline selection_population_specific(Functor_sel s_in, Functor_sel_pop s_pop_in, int pop, int generation_shift = 0)
        : pop(pop), generation_shift(generation_shift), s(s_in), s_pop(s_pop_in) {}
    __host__ __device__ __forceinline__ float operator()(const int population, const int generation, const float freq) const {
        return pop == population ? s_pop(population, generation - generation_shift, freq) : s(population, generation - generation_shift, freq);
    }
    inline void linear_form(int population, int generation, float &sl, float &ic) const;
};

/** s1 before generation inflection_point (+shift), s2 from then on */
template <typename Functor_sel1, typename Functor_sel2>
struct selection_piecewise {
    int inflection_point;
    int generation_shift;
    Functor_sel1 s1;
    Functor_sel2 s2;
    inline selection_piecewise() : inflection_point(0), generation_shift(0), s1(), s2() {}
    inline selection_piecewise(Functor_sel1 s1_in, Functor_sel2 s2_in, int inflection_point, int generation_shift = 0)
        : inflection_point(inflection_point), generation_shift(generation_shift), s1(s1_in), s2(s2_in) {}
    __host__ __device__ __forceinline__ float operator()(const int population, const int generation, const float freq) const {
        return generation >= inflection_point + generation_shift ? s2(population, generation - generation_shift, freq)
                                                                 : s1(population, generation - generation_shift, freq);
    }
    inline void linear_form(int population, int generation, float &sl, float &ic) const;
};

/* ================= inbreeding, mutation, dominance ================= */

struct F_mu_h_constant {
    float p;
    inline F_mu_h_constant() : p(0) {}
    inline F_mu_h_constant(float p) : p(p) {}
    __host__ __forceinline__ float operator()(const int population, const int generation) const { return p; }
};

struct F_mu_h_sine_wave {
    float A;
    float pi;
    float rho;
    float D;
    int generation_shift;
    inline F_mu_h_sine_wave() : A(0), pi(0), rho(0), D(0), generation_shift(0) {}
    inline F_mu_h_sine_wave(float A, float pi, float D, float rho = 0, int generation_shift = 0)
        : A(A), pi(pi), rho(rho), D(D), generation_shift(generation_shift) {}
    __host__ __forceinline__ float operator()(const int population, const int generation) const {
        return A * std::sin(pi * (float)(generation - generation_shift) + rho) + D;
    }
};

template <typename Functor_p, typename Functor_p_pop>
struct F_mu_h_population_specific {
    int pop;
    int generation_shift;
    Functor_p p;
    Functor_p_pop p_pop;
    inline F_mu_h_population_specific() : pop(0), generation_shift(0), p(), p_pop() {}
    inline F_mu_h_population_specific(Functor_p p_in, Functor_p_pop p_pop_in, int pop, int generation_shift = 0)
        : pop(pop), generation_shift(generation_shift), p(p_in), p_pop(p_pop_in) {}
    __host__ __forceinline__ float operator()(const int population, const int generation) const {
        return pop == population ? p_pop(population, generation - generation_shift) : p(population, generation - generation_shift);
    }
};

template <typename Functor_p1, typename Functor_p2>
struct F_mu_h_piecewise {
    int inflection_point;
    int generation_shift;
    Functor_p1 p1;
    Functor_p2 p2;
    inline F_mu_h_piecewise() : inflection_point(0), generation_shift(0), p1(), p2() {}
    inline F_mu_h_piecewise(Functor_p1 p1_in, Functor_p2 p2_in, int inflection_point, int generation_shift = 0)
        : inflection_point(inflection_point), generation_shift(generation_shift), p1(p1_in), p2(p2_in) {}
    __host__ __forceinline__ float operator()(const int population, const int generation) const {
        return generation >= inflection_point + generation_shift ? p2(population, generation - generation_shift)
                                                                 : p1(population, generation - generation_shift);
    }
};

/* ================= demography ================= */

struct demography_constant {
    int N;
    inline demography_constant() : N(0) {}
    inline demography_constant(int p) : N(p) {}
    __host__ __device__ __forceinline__ int operator()(const int population, const int generation) const { return N; }
};

/** N = (int)A * sin(pi*(generation-shift) + rho) + D: the cast applies to A
 * alone, as in the reference (template_inline_simulation_functors.cuh:291) */
struct demography_sine_wave {
    float A;
    float pi;
    float rho;
    float D;
    int generation_shift;
    inline demography_sine_wave() : A(0), pi(0), rho(0), D(0), generation_shift(0) {}
    inline demography_sine_wave(float A, float pi, int D, float rho = 0, int generation_shift = 0)
        : A(A), pi(pi), rho(rho), D(D), generation_shift(generation_shift) {}
    __host__ __device__ __forceinline__ int operator()(const int population, const int generation) const {
        return (int)((int)A * sinf(pi * (float)(generation - generation_shift) + rho) + D);
    }
};

/** N = round(initial_population_size * exp(rate*(generation-shift))) */
struct demography_exponential_growth {
    float rate;
    int initial_population_size;
    int generation_shift;
    inline demography_exponential_growth() : rate(0), initial_population_size(0), generation_shift(0) {}
    inline demography_exponential_growth(float rate, int initial_population_size, int generation_shift = 0)
        : rate(rate), initial_population_size(initial_population_size), generation_shift(generation_shift) {}
    __host__ __device__ __forceinline__ int operator()(const int population, const int generation) const {
        return (int)roundf(initial_population_size * expf(rate * (float)(generation - generation_shift)));
    }
};

/** logistic growth to carrying_capacity */
struct demography_logistic_growth {
    float rate;
    int initial_population_size;
    int carrying_capacity;
    int generation_shift;
    inline demography_logistic_growth() : rate(0), initial_population_size(0), carrying_capacity(0), generation_shift(0) {}
    inline demography_logistic_growth(float rate, int initial_population_size, int carrying_capacity, int generation_shift = 0)
        : rate(rate), initial_population_size(initial_population_size), carrying_capacity(carrying_capacity), generation_shift(generation_shift) {}
    __host__ __device__ __forceinline__ int operator()(const int population, const int generation) const {
        const float term = expf(rate * (float)(generation - generation_shift));
        return (int)roundf(carrying_capacity * initial_population_size * term / (carrying_capacity + initial_population_size * (term - 1)));
    }
};

template <typename Functor_d, typename Functor_d_pop>
struct demography_population_specific {
    int pop;
    int generation_shift;
    Functor_d d;
    Functor_d_pop d_pop;
    inline demography_population_specific() : pop(0), generation_shift(0), d(), d_pop() {}
    inline demography_population_specific(Functor_d d_in, Functor_d_pop d_pop_in, int pop, int generation_shift = 0)
        : pop(pop), generation_shift(generation_shift), d(d_in), d_pop(d_pop_in) {}
    __host__ __device__ __forceinline__ int operator()(const int population, const int generation) const {
        return pop == population ? d_pop(population, generation - generation_shift) : d(population, generation - generation_shift);
    }
};

template <typename Functor_d1, typename Functor_d2>
struct demography_piecewise {
    int inflection_point;
    int generation_shift;
    Functor_d1 d1;
    Functor_d2 d2;
    inline demography_piecewise() : inflection_point(0), generation_shift(0), d1(), d2() {}
    inline demography_piecewise(Functor_d1 d1_in, Functor_d2 d2_in, int inflection_point, int generation_shift = 0)
        : inflection_point(inflection_point), generation_shift(generation_shift), d1(d1_in), d2(d2_in) {}
    __host__ __device__ __forceinline__ int operator()(const int population, const int generation) const {
        return generation >= inflection_point + generation_shift ? d2(population, generation - generation_shift)
                                                                 : d1(population, generation - generation_shift);
    }
};

/* ================= migration (conservative model) ================= */

/** m from every population to every other, 1-(num_pop-1)*m to itself */
struct migration_constant_equal {
    float m;
    int num_pop;
    inline migration_constant_equal() : m(0), num_pop(1) {}
    inline migration_constant_equal(float m, int num_pop) : m(m), num_pop(num_pop > 1 ? num_pop : 1) {}
    __host__ __device__ __forceinline__ float operator()(const int pop_FROM, const int pop_TO, const int generation) const {
        if (pop_FROM == pop_TO) return 1 - (num_pop - 1) * m;
        return (num_pop > 1) * m;
    }
};

/** m from pop1 to pop2, rest(...) for every other direction */
template <typename Functor_m1>
struct migration_constant_directional {
    float m;
    int pop1;
    int pop2;
    Functor_m1 rest;
    inline migration_constant_directional() : m(0), pop1(0), pop2(0), rest() {}
    inline migration_constant_directional(float m, int pop1, int pop2, Functor_m1 rest_in) : m(m), pop1(pop1), pop2(pop2), rest(rest_in) {}
    __host__ __device__ __forceinline__ float operator()(const int pop_FROM, const int pop_TO, const int generation) const {
        if (pop_FROM == pop1 && pop_TO == pop2) return m;
        return rest(pop_FROM, pop_TO, generation);
    }
};

template <typename Functor_m1, typename Functor_m2>
struct migration_piecewise {
    int inflection_point;
    int generation_shift;
    Functor_m1 m1;
    Functor_m2 m2;
    inline migration_piecewise() : inflection_point(0), generation_shift(0), m1(), m2() {}
    inline migration_piecewise(Functor_m1 m1_in, Functor_m2 m2_in, int inflection_point, int generation_shift = 0)
        : inflection_point(inflection_point), generation_shift(generation_shift), m1(m1_in), m2(m2_in) {}
    __host__ __device__ __forceinline__ float operator()(const int pop_FROM, const int pop_TO, const int generation) const {
        return generation >= inflection_point + generation_shift ? m2(pop_FROM, pop_TO, generation - generation_shift)
                                                                 : m1(pop_FROM, pop_TO, generation - generation_shift);
    }
};

/* ================= preserving and sampling ================= */

struct bool_off {
    __host__ __forceinline__ bool operator()(const int generation) const { return false; }
};

struct bool_on {
    __host__ __forceinline__ bool operator()(const int generation) const { return true; }
};

/** f_action at generation pulse (+shift), f_default otherwise */
template <typename Functor_default, typename Functor_action>
struct bool_pulse {
    int pulse;
    int generation_shift;
    Functor_default f_default;
    Functor_action f_action;
    inline bool_pulse() : pulse(0), generation_shift(0), f_default(), f_action() {}
    inline bool_pulse(int pulse, int generation_shift = 0) : pulse(pulse), generation_shift(generation_shift), f_default(), f_action() {}
    inline bool_pulse(Functor_default f_default_in, Functor_action f_action_in, int pulse, int generation_shift = 0)
        : pulse(pulse), generation_shift(generation_shift), f_default(f_default_in), f_action(f_action_in) {}
    __host__ __forceinline__ bool operator()(const int generation) const {
        return generation == pulse + generation_shift ? f_action(generation) : f_default(generation);
    }
};

/** f1 before inflection_point (+shift), f2 from then on */
template <typename Functor_first, typename Functor_second>
struct bool_piecewise {
    int inflection_point;
    int generation_shift;
    Functor_first f1;
    Functor_second f2;
    inline bool_piecewise() : inflection_point(0), generation_shift(0), f1(), f2() {}
    inline bool_piecewise(int inflection_point, int generation_shift = 0)
        : inflection_point(inflection_point), generation_shift(generation_shift), f1(), f2() {}
    inline bool_piecewise(Functor_first f1_in, Functor_second f2_in, int inflection_point, int generation_shift = 0)
        : inflection_point(inflection_point), generation_shift(generation_shift), f1(f1_in), f2(f2_in) {}
    __host__ __forceinline__ bool operator()(const int generation) const {
        return generation >= inflection_point + generation_shift ? f2(generation) : f1(generation);
    }
};

} /* namespace Sim_Model */

/* ================= lowering of selection functors ================= */
namespace go_fish_details {

/* true if a selection functor has a per-(population, generation) linear
 * form: a linear_form member, and for the composite Sim_Model types, parts
 * that all have one */
template <typename T>
struct has_linear_form {
    template <typename U>
    static auto test(int) -> decltype(std::declval<const U &>().linear_form(0, 0, std::declval<float &>(), std::declval<float &>()), std::true_type());
    template <typename>
    static std::false_type test(...);
    static constexpr bool value = decltype(test<T>(0))::value;
};
template <typename A, typename B>
struct has_linear_form<Sim_Model::selection_population_specific<A, B>>
    : std::integral_constant<bool, has_linear_form<A>::value && has_linear_form<B>::value> {};
template <typename A, typename B>
struct has_linear_form<Sim_Model::selection_piecewise<A, B>>
    : std::integral_constant<bool, has_linear_form<A>::value && has_linear_form<B>::value> {};

template <typename Sel>
inline void lower_selection(const Sel &sel, int population, int generation, float &slope, float &intercept) {
    if constexpr (has_linear_form<Sel>::value) {
        sel.linear_form(population, generation, slope, intercept);
    } else {
        slope = 0.0f;
        intercept = 0.0f;   /* not used: selection comes from the functor itself */
    }
}

/* Launchers handed to gf_run_with_selection for a functor without a linear
 * form: the library's drift kernel and equilibrium selection table,
 * instantiated here with the functor (user = the functor object). */
template <typename Sel>
struct selection_hooks {
    static int32_t mse_table(void *user, int32_t pop, int32_t gen, int32_t Nc, float *table, void *stream) {
        const Sel &f = *static_cast<const Sel *>(user);
        const int grid = (Nc + 255) / 256 < 1024 ? (Nc + 255) / 256 : 1024;
        hipLaunchKernelGGL((gf::k_mse_sel_table<Sel>), dim3(grid > 0 ? grid : 1), dim3(256), 0,
                           static_cast<hipStream_t>(stream), f, pop, gen, Nc, table);
        return hipGetLastError() == hipSuccess ? GF_OK : GF_ERR_DEVICE;
    }
    template <int P>
    static void launch(const Sel &f, const void *fused, const void *segment, int grid, hipStream_t st) {
        hipLaunchKernelGGL((gf::k_advance_fused<P, gf::FunctorSel<Sel>>), dim3(grid), dim3(gf::WG), 0, st,
                           *static_cast<const gf::FusedLaunch *>(fused),
                           *static_cast<const gf::SegLaunch<P> *>(segment), gf::FunctorSel<Sel>{f});
    }
    static int32_t advance(void *user, int32_t P, const void *fused, const void *segment, int32_t grid, void *stream) {
        const Sel &f = *static_cast<const Sel *>(user);
        hipStream_t st = static_cast<hipStream_t>(stream);
        switch (P) {
        case 1: launch<1>(f, fused, segment, grid, st); break;
        case 2: launch<2>(f, fused, segment, grid, st); break;
        case 3: launch<3>(f, fused, segment, grid, st); break;
        case 4: launch<4>(f, fused, segment, grid, st); break;
        case 5: launch<5>(f, fused, segment, grid, st); break;
        case 6: launch<6>(f, fused, segment, grid, st); break;
        case 7: launch<7>(f, fused, segment, grid, st); break;
        case 8: launch<8>(f, fused, segment, grid, st); break;
        default: return GF_ERR_ARG;
        }
        return hipGetLastError() == hipSuccess ? GF_OK : GF_ERR_DEVICE;
    }
};

/* The model tables of a run (the tables of gf_sim_desc): one gf_pop_gen
 * row per (generation, population), the P x P migration matrix of every
 * generation ([from][to]) and the preserve / sample flags. */
struct sim_tables {
    std::vector<gf_pop_gen> params;
    std::vector<float> mig;
    std::vector<uint8_t> pres, samp;
};

/* Evaluates the host-callable model functors at generations start ..
 * start + G, as the reference evaluates them on the host every generation
 * (go_fish_impl.cuh:647-661, :454-467, :705-712), and lowers selection to
 * its per-(population, generation) linear form where it has one. */
template <typename Functor_mutation, typename Functor_demography, typename Functor_migration, typename Functor_selection,
          typename Functor_inbreeding, typename Functor_dominance, typename Functor_preserve, typename Functor_timesample>
inline void build_tables(sim_tables &T, int P, int G, int start, const Functor_mutation &mu_rate,
                         const Functor_demography &demography, const Functor_migration &mig_prop,
                         const Functor_selection &sel_coeff, const Functor_inbreeding &FI, const Functor_dominance &dominance,
                         const Functor_preserve &preserve_mutations, const Functor_timesample &take_sample) {
    T.params.assign((size_t)(G + 1) * P, gf_pop_gen());
    T.mig.assign((size_t)(G + 1) * P * P, 0.0f);
    T.pres.assign(G + 1, 0);
    T.samp.assign(G + 1, 0);
    for (int t = 0; t <= G; t++) {
        const int g = start + t;
        for (int p = 0; p < P; p++) {
            gf_pop_gen &r = T.params[(size_t)t * P + p];
            r.N = demography(p, g);
            r.F = FI(p, g);
            r.h = dominance(p, g);
            r.mu = mu_rate(p, g);
            r.sel_slope = 0.0f;
            r.sel_intercept = 0.0f;
            lower_selection(sel_coeff, p, g, r.sel_slope, r.sel_intercept);
            for (int k = 0; k < P; k++) T.mig[(size_t)t * P * P + (size_t)k * P + p] = mig_prop(k, p, g);
        }
        T.pres[t] = preserve_mutations(g) ? 1 : 0;
        T.samp[t] = take_sample(g) ? 1 : 0;
    }
}

} /* namespace go_fish_details */

namespace Sim_Model {
template <typename Functor_sel, typename Functor_sel_pop>
inline void selection_population_specific<Functor_sel, Functor_sel_pop>::linear_form(int population, int generation, float &sl, float &ic) const {
    if (pop == population) go_fish_details::lower_selection(s_pop, population, generation - generation_shift, sl, ic);
    else go_fish_details::lower_selection(s, population, generation - generation_shift, sl, ic);
}
template <typename Functor_sel1, typename Functor_sel2>
inline void selection_piecewise<Functor_sel1, Functor_sel2>::linear_form(int population, int generation, float &sl, float &ic) const {
    if (generation >= inflection_point + generation_shift) go_fish_details::lower_selection(s2, population, generation - generation_shift, sl, ic);
    else go_fish_details::lower_selection(s1, population, generation - generation_shift, sl, ic);
}
} /* namespace Sim_Model */

/* ================= run_sim ================= */
namespace GO_Fish {

static_assert(sizeof(mutID) == sizeof(gf_mutid), "mutID and gf_mutid must share one layout");

template <typename Functor_mutation, typename Functor_demography, typename Functor_migration, typename Functor_selection,
          typename Functor_inbreeding, typename Functor_dominance, typename Functor_preserve, typename Functor_timesample>
inline void run_sim(allele_trajectories &all_results, const Functor_mutation mu_rate, const Functor_demography demography,
                    const Functor_migration mig_prop, const Functor_selection sel_coeff, const Functor_inbreeding FI,
                    const Functor_dominance dominance, const Functor_preserve preserve_mutations,
                    const Functor_timesample take_sample, const allele_trajectories &prev_sim) {
    allele_trajectories::sim_constants &in = all_results.sim_input_constants;
    const int P = in.num_populations;
    int start = 0;
    int mode = GF_INIT_MSE;
    const allele_trajectories::time_sample *src = NULL;
    if (!in.init_mse) {
        const int si = in.prev_sim_sample;
        if (si >= 0 && si < prev_sim.num_samples && prev_sim.time_samples) {
            const float pn = prev_sim.sim_run_constants.num_sites;
            const int pp = prev_sim.sim_run_constants.num_populations;
            if (in.num_sites == pn && P == pp) {
                mode = GF_INIT_PREV;
                src = prev_sim.time_samples[si];
                if (src->num_mutations > 0) start = src->sampled_generation;
            } else {
                std::fprintf(stderr, "run_sim error: prev_sim parameters do not match current simulation parameters: prev_sim num_sites %f\tcurrent_sim num_sites %f,\tprev_sim num_populations %d\tcurrent_sim num_populations %d\n",
                             pn, in.num_sites, pp, P);
                std::exit(1);
            }
        } else if (si < 0) {
            mode = GF_INIT_BLANK;
        } else {
            if (!prev_sim.time_samples) detail::fail("run_sim error: requested time sample from empty prev_sim\n");
            std::fprintf(stderr, "run_sim error: requested sample index out of bounds for prev_sim: sample %d\t[0\t %d)\n", si, prev_sim.num_samples);
            std::exit(1);
        }
    }
    const int G = in.num_generations;
    if (P < 1 || P > GF_MAX_POPULATIONS) {
        std::fprintf(stderr, "run_sim error: num_populations %d outside [1, %d]\n", P, GF_MAX_POPULATIONS);
        std::exit(1);
    }
    go_fish_details::sim_tables T;
    go_fish_details::build_tables(T, P, G, start, mu_rate, demography, mig_prop, sel_coeff, FI, dominance,
                                  preserve_mutations, take_sample);
    std::vector<uint8_t> ext(P, 0);
    gf_sim_desc d;
    std::memset(&d, 0, sizeof d);
    d.num_populations = P;
    d.num_sites = in.num_sites;
    d.seed1 = in.seed1;
    d.seed2 = in.seed2;
    d.num_generations = G;
    d.compact_interval = in.compact_interval;
    d.init_mode = mode;
    d.start_generation = start;
    if (mode == GF_INIT_PREV && src->num_mutations > 0) {
        d.init_num_mutations = src->num_mutations;
        d.init_freq = src->mutations_freq;
        d.init_ids = (const gf_mutid *)prev_sim.mutations_ID;
        for (int p = 0; p < P; p++) ext[p] = src->extinct[p] ? 1 : 0;
        d.init_extinct = ext.data();
    }
    d.params = T.params.data();
    d.migration = T.mig.data();
    d.preserve = T.pres.data();
    d.sample = T.samp.data();
    /* the reference's per-generation parameter checks, for the whole run,
     * before any device is touched */
    if (gf_check_sim(&d) != GF_OK) {
        std::fprintf(stderr, "%s\n", gf_last_error());
        std::exit(1);
    }
    /* device selection (reference set_cuda_device, shared.cuh:27-37) */
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { std::fprintf(stderr, "run_sim error: no HIP device\n"); std::exit(1); }
    if (in.device >= 0 && in.device < ndev) (void)hipSetDevice(in.device);
    int dev = 0;
    (void)hipGetDevice(&dev);
    in.device = dev;
    d.device = dev;
    gf_result *res = NULL;
    int32_t rc;
    if constexpr (go_fish_details::has_linear_form<Functor_selection>::value) {
        rc = gf_run(&d, &res);
    } else {
        gf_selection_hooks hooks;
        hooks.user = const_cast<void *>(static_cast<const void *>(&sel_coeff));
        hooks.mse_table = &go_fish_details::selection_hooks<Functor_selection>::mse_table;
        hooks.advance = &go_fish_details::selection_hooks<Functor_selection>::advance;
        rc = gf_run_with_selection(&d, &hooks, &res);
    }
    if (rc != GF_OK) {
        std::fprintf(stderr, "%s\n", gf_last_error());
        std::exit(1);
    }
    const int ns = gf_result_num_samples(res);
    all_results.initialize_sim_result_vector(ns);
    for (int s = 0; s < ns; s++) {
        allele_trajectories::time_sample *t = all_results.time_samples[s];
        int M = 0, gen = 0;
        t->Nchrom_e = new int[P];
        t->extinct = new bool[P];
        std::vector<uint8_t> e(P);
        gf_result_sample_info(res, s, &M, &gen, t->Nchrom_e, e.data());
        for (int p = 0; p < P; p++) t->extinct[p] = e[p] != 0;
        t->num_mutations = M;
        t->sampled_generation = gen;
        t->mutations_freq = new float[(size_t)P * M];
        gf_result_sample_freq(res, s, t->mutations_freq);
    }
    const int nid = gf_result_num_ids(res);
    all_results.all_mutations = nid;
    if (nid > 0) {
        all_results.mutations_ID = (mutID *)std::malloc(sizeof(mutID) * nid);
        gf_result_ids(res, (gf_mutid *)all_results.mutations_ID);
    }
    gf_result_free(res);
}

/** run_sim without a previous simulation (mutation-selection equilibrium or blank start) */
template <typename Functor_mutation, typename Functor_demography, typename Functor_migration, typename Functor_selection,
          typename Functor_inbreeding, typename Functor_dominance, typename Functor_preserve, typename Functor_timesample>
inline void run_sim(allele_trajectories &all_results, const Functor_mutation mu_rate, const Functor_demography demography,
                    const Functor_migration mig_prop, const Functor_selection sel_coeff, const Functor_inbreeding FI,
                    const Functor_dominance dominance, const Functor_preserve preserve_mutations,
                    const Functor_timesample take_sample) {
    run_sim(all_results, mu_rate, demography, mig_prop, sel_coeff, FI, dominance, preserve_mutations, take_sample,
            allele_trajectories());
}

} /* namespace GO_Fish */

#endif /* GO_FISH_API_H_ */
