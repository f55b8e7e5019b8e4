"""The C++ API layer on the host: Sim_Model functors lowered to model tables
by run_sim's own table builder (include/go_fish.cuh), the functors'
formulas against the reference's (template_inline_simulation_functors.cuh),
and the API's error exits (stderr message + exit status 1), through the
test program tests/cpp/sim_model.hip.

The three-population scenario built from functors must give exactly the
tables of parallelpopgen_amd/scenarios.three_population, which the GPU
tests run against the oracle; so the scenario's C++ lowering is what those
runs check."""
import math
import os
import subprocess

import numpy as np
import pytest

from parallelpopgen_amd import scenarios as S
from parallelpopgen_amd.capi import POP_GEN_DTYPE

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "tests", "cpp", "_build", "sim_model")


def _exe():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/_build/sim_model is not built (make -C tests/cpp)")
    q = subprocess.run(["make", "-q", "-C", os.path.join(REPO, "tests", "cpp"), EXE], capture_output=True)
    if q.returncode != 0:
        pytest.fail("tests/cpp/_build/sim_model is older than its sources (make -C tests/cpp)")
    return EXE


def test_sim_model_program_builds():
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "tests", "cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.access(EXE, os.X_OK)


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("sim_model"))
    r = subprocess.run([_exe(), "tables", d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return d


def test_three_population_functors_give_the_scenario_tables(tables):
    G, P = 2000, 3
    params = np.fromfile(os.path.join(tables, "three_pop_params.bin"), POP_GEN_DTYPE).reshape(G + 1, P)
    mig = np.fromfile(os.path.join(tables, "three_pop_mig.bin"), np.float32).reshape(G + 1, P, P)
    pres = np.fromfile(os.path.join(tables, "three_pop_pres.bin"), np.uint8)
    samp = np.fromfile(os.path.join(tables, "three_pop_samp.bin"), np.uint8)
    d = S.three_population()
    want = d.params.reshape(G + 1, P)
    for f in ("N", "F", "h", "mu", "sel_slope", "sel_intercept"):
        a, b = params[f], want[f]
        if a.dtype.kind == "f":
            a, b = a.view(np.uint32), b.view(np.uint32)
        bad = np.argwhere(a != b)
        assert len(bad) == 0, (f, bad[:5], params[f][tuple(bad[0])], want[f][tuple(bad[0])])
    assert np.array_equal(mig.view(np.uint32), d.migration.reshape(G + 1, P, P).view(np.uint32))
    assert np.array_equal(pres, d.preserve) and np.array_equal(samp, d.sample)
    # the scenario as described: growth to 2e5 at generation 500, the
    # bottleneck in [500, 600), F = 0.1 only in the third population
    assert params["N"][0, 0] == 20000 and params["N"][500, 0] == 200000 and params["N"][499, 0] < 200000
    assert set(params["N"][500:600, 1]) == {2000} and params["N"][600, 1] == 20000
    assert np.all(params["F"][:, 2] == np.float32(0.1)) and np.all(params["F"][:, :2] == 0)


def _functors(tables):
    out = {}
    with open(os.path.join(tables, "functors.txt")) as f:
        for line in f:
            t = line.split()
            out.setdefault(t[0], []).append([float(v) for v in t[1:]])
    return {k: np.array(v) for k, v in out.items()}


def _arg(pi, g, shift, rho):
    """pi * (generation - shift) + rho evaluated in float, as the functors do"""
    f32 = np.float32
    return ((f32(pi) * (g - shift).astype(f32)).astype(f32) + f32(rho)).astype(f32).astype(np.float64)


def test_functor_formulas(tables):
    F = _functors(tables)
    f32 = np.float32
    g = F["logistic"][:, 0].astype(np.int64)
    # demography_logistic_growth (reference :305-310): term = exp(rate (g - shift)) in float,
    # N = round(K N0 term / (K + N0 (term - 1)))
    rate, N0, K, shift = 0.01, 1000, 100000, 7
    term = np.exp(f32(rate) * (g - shift).astype(np.float64))
    exact = K * N0 * term / (K + N0 * (term - 1))
    got = F["logistic"][:, 1]
    assert np.all(np.abs(got - exact) <= 0.5 + 1e-6 * exact)
    assert got[0] < 1000 and got[-1] > 99000          # shifted start below N0, saturates at K
    # demography_sine_wave (reference :291): (int)A * sin(pi (g - shift) + rho) + D -- the
    # cast binds to A alone: A = 2.9 acts as 2
    A, pi, D, rho, shift = 2.9, 0.1, 1000, 0.25, 3
    v = int(A) * np.sin(_arg(pi, g, shift, rho)) + D
    got = F["demography_sine"][:, 1]
    clear = np.abs(v - np.round(v)) > 1e-4
    assert np.array_equal(got[clear], np.trunc(v[clear]))
    assert got.min() == 998 and got.max() <= 1002     # 1000 + 2.9 sin(...) would reach 997
    # F_mu_h_sine_wave (reference :190): A sin(pi (g - shift) + rho) + D
    A, pi, D, rho, shift = 0.05, 2 * math.pi / 50, 0.1, 0.3, 5
    want = f32(A) * np.sin(_arg(pi, g, shift, rho)) + f32(D)
    assert np.allclose(F["F_sine"][:, 1], want, rtol=0, atol=1e-7)
    # F_mu_h_piecewise (reference :209-222): p2 (with the generation shifted) from the
    # inflection point + shift on, p1 before
    ip, sh = 100, 3
    want_pw = np.where(g >= ip + sh, f32(A) * np.sin(_arg(pi, g - sh, shift, rho)) + f32(D), f32(0.1))
    assert np.allclose(F["F_piecewise"][:, 1], want_pw, rtol=0, atol=1e-7)
    # bool_piecewise / bool_pulse: the reference's documented example samples generation 0
    # and [100, 110] (template_inline_simulation_functors.cuh:620-628)
    on = F["bool_piecewise"][:, 1].astype(bool)
    assert set(np.flatnonzero(on)) == {0} | set(range(100, 111))
    # selection_sine_wave's linear form: slope 0, intercept A sin(pi (g - shift) + rho) + D
    A, pi, D, rho, shift = 0.002, 0.05, -0.001, 0.7, 11
    assert np.all(F["selection_sine"][:, 1] == 0)
    want = f32(A) * np.sin(_arg(pi, g, shift, rho)) + f32(D)
    assert np.allclose(F["selection_sine"][:, 2], want, rtol=0, atol=1e-9)


# error exits that need no device: case -> the reference's message (its
# file:line) that stderr must contain
HOST_DEATHS = {
    "mu_negative": "check_sim_parameters: mutation error: mu_rate < 0",                      # go_fish_impl.cuh:436
    "inbreeding_above_1": "check_sim_parameters: inbreeding error: inbreeding coefficient > 1",  # :441
    "inbreeding_below_0": "check_sim_parameters: inbreeding error: inbreeding coefficient < 0",  # :440
    "migration_negative": "check_sim_parameters: migration error: migration rate < 0",        # :444
    "migration_sum": "check_sim_parameters: migration error: migration rate does not sum to 1",  # :450
    "migration_from_unfounded": "migration error, migration from non-existant population",    # :445
    "revived_extinct": "check_sim_parameters: demography error: extinct population with population size > 0",  # :438
    "prev_sim_empty": "run_sim error: requested time sample from empty prev_sim",             # :617
    "num_populations": "run_sim error: num_populations 9 outside",                            # (no counterpart: 8 at most)
    "num_mutations_empty": "num_mutations error: empty allele_trajectories",                  # inline_go_fish_data_struct.hpp:94
    "sampled_generation_empty": "sampled_generation error: empty allele_trajectories",        # :102
    "extinct_empty": "extinct error: empty allele_trajectories",                              # :108
    "Nchrom_empty": "effective_number_of_chromosomes error: empty allele_trajectories",       # :115
    "frequency_empty": "frequency error: empty allele_trajectories",                          # :132
    "mutation_ID_empty": "mutation_ID error: empty allele_trajectories",                      # :142
    "delete_time_sample_empty": "delete_time_sample error: empty allele_trajectories",        # :168
    "sfs_empty": "error transferring allele_trajectories to spectrum: empty allele_trajectories",  # spectrum.cu:59
    "histogram_empty": "error transferring allele_trajectories to spectrum: empty allele_trajectories",
}


@pytest.mark.parametrize("case", sorted(HOST_DEATHS))
def test_error_exit(case):
    r = subprocess.run([_exe(), "die", case], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert HOST_DEATHS[case] in r.stderr, r.stderr
    assert "no error exit" not in r.stdout


def test_parameter_checks_name_generation_and_population(tables):
    """the messages carry the generation and population as the reference's
    do (the extinct population revives at generation 15, population 1)"""
    r = subprocess.run([_exe(), "die", "revived_extinct"], capture_output=True, text=True, timeout=60)
    assert "generation 15\t population 1" in r.stderr, r.stderr


# error exits after a run on the GPU
GPU_DEATHS = {
    "sfs_sample_size": "site_frequency_spectrum error: requested sample_size out of range [1,population_size)",  # spectrum.cu:229
    "sfs_indices": "site_frequency_spectrum error: requested indices out of bounds",         # spectrum.cu:224
    "frequency_bounds": "frequency error: index out of bounds",                               # inline_go_fish_data_struct.hpp:134
    "mutation_ID_bounds": "mutation_ID error: requested mutation index out of bounds",        # :141
    "extinct_bounds": "extinct error: index out of bounds",                                   # :110
    "prev_sim_mismatch": "run_sim error: prev_sim parameters do not match current simulation parameters",  # go_fish_impl.cuh:609
    "prev_sim_index": "run_sim error: requested sample index out of bounds for prev_sim",     # :618
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(GPU_DEATHS))
def test_error_exit_after_a_run(case):
    print("sim_model run_then_die", case, flush=True)
    r = subprocess.run([_exe(), "run_then_die", case], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert GPU_DEATHS[case] in r.stderr, r.stderr
