"""GPU run_sim path vs the CPU oracle on the same tables: every time sample's
frequencies, the final mutation IDs and the per-generation live counts must
be identical bit for bit."""
import numpy as np
import pytest

from parallelpopgen_amd import scenarios as S
from parallelpopgen_amd.capi import GF_INIT_BLANK, GF_INIT_PREV, POP_GEN_DTYPE, SimDesc

pytestmark = pytest.mark.gpu


def assert_same(a, b, what=""):
    assert len(a["samples"]) == len(b["samples"]), what
    for i, (x, y) in enumerate(zip(a["samples"], b["samples"])):
        assert x["generation"] == y["generation"], (what, i)
        assert x["num_mutations"] == y["num_mutations"], (what, i, x["num_mutations"], y["num_mutations"])
        assert x["Nchrom_e"] == y["Nchrom_e"] and x["extinct"] == y["extinct"], (what, i)
        fx, fy = x["freq"].view(np.uint32), y["freq"].view(np.uint32)
        if not np.array_equal(fx, fy):
            bad = np.argwhere(fx != fy)
            raise AssertionError("%s sample %d: %d mismatches, first %s gpu=%r oracle=%r" % (
                what, i, len(bad), bad[0], x["freq"][tuple(bad[0])], y["freq"][tuple(bad[0])]))
    assert np.array_equal(a["ids"], b["ids"]), what
    assert np.array_equal(a["live"], b["live"]), (what, np.argwhere(a["live"] != b["live"])[:5])


def test_plain_small_full(gpu, oracle):
    d = S.plain_small()
    g = gpu.run(d)
    o = oracle.run(d)
    assert_same(g, o, "plain_small")
    assert g["samples"][-1]["num_mutations"] > 7000


@pytest.mark.parametrize("kpl", [1, 5, 16, 64])
def test_generations_per_launch_invariance(gpu, oracle, kpl):
    d = S.plain_small(G=150, compact_interval=20)
    ref = oracle.run(d)
    gpu.set_generations_per_launch(kpl)
    try:
        g = gpu.run(d)
    finally:
        gpu.set_generations_per_launch(0)
    assert_same(g, ref, "kpl=%d" % kpl)


@pytest.mark.parametrize("interval", [0, 1, 7, 35])
def test_compaction_interval(gpu, oracle, interval):
    d = S.plain_small(G=120, compact_interval=interval)
    assert_same(gpu.run(d), oracle.run(d), "interval=%d" % interval)


def test_compaction_does_not_change_single_population_results(gpu):
    a = gpu.run(S.plain_small(G=200, compact_interval=10))
    b = gpu.run(S.plain_small(G=200, compact_interval=45))
    assert np.array_equal(a["samples"][-1]["freq"], b["samples"][-1]["freq"])
    assert np.array_equal(a["ids"], b["ids"])


def test_plain_large_selected_scaled(gpu, oracle):
    # gamma = -10, h = 0.5 at N = 1e5 with L scaled down so the oracle is quick
    d = S.plain_large(L=2e6, G=300)
    assert_same(gpu.run(d), oracle.run(d), "plain_large L=2e6")


def test_three_population_scaled(gpu, oracle):
    d = S.three_population(scale=0.1, L=2e8, G=700, preserve_every=100, sample_every=200)
    g, o = gpu.run(d), oracle.run(d)
    assert_same(g, o, "three_population")
    assert len(g["samples"]) == 4


def test_blank_start_small_N_binomial(gpu, oracle):
    # N < 100 chromosomes exercises the binomial CDF inversion
    d = S.constant(N=30, L=1e9, G=400, mu=1e-8, F=0.0, h=0.2, s=-0.02, init_mode=GF_INIT_BLANK,
                   compact_interval=13)
    assert_same(gpu.run(d), oracle.run(d), "blank N=30")


def test_positive_selection_dominance_inbreeding(gpu, oracle):
    d = S.constant(N=5000, L=5e7, G=300, F=0.3, h=1.0, s=0.004, compact_interval=9)
    assert_same(gpu.run(d), oracle.run(d), "s>0")
    d = S.constant(N=5000, L=5e7, G=300, F=1.0, h=0.0, s=-0.01, compact_interval=9)
    assert_same(gpu.run(d), oracle.run(d), "haploid")


def split_scenario(G=400, L=5e7, split=100, interval=25):
    """population 1 splits off population 0 at generation `split`, grows,
    then both exchange migrants; population 0 goes extinct at generation
    300.  A split generation that is not a multiple of 4 founds the
    population inside a block of four generations that share one Philox
    output."""
    P = 2
    params = np.zeros((G + 1, P), POP_GEN_DTYPE)
    g = np.arange(G + 1)
    params["N"][:, 0] = np.where(g < 300, 8000, 0)
    params["N"][:, 1] = np.where(g < split, 0, 2000 + 10 * (g - split))
    params["F"] = 0.0
    params["h"] = 0.5
    params["mu"] = 2e-9
    params["sel_slope"][:, 1] = -0.002
    params["sel_intercept"][:, 1] = 0.001
    mig = np.zeros((G + 1, P, P), np.float32)
    mig[:, 0, 0] = 1
    mig[:, 1, 1] = 1
    mig[split, 0, 1] = 1.0   # at the split pop 1 is made of pop 0 chromosomes
    mig[split, 1, 1] = 0.0
    m = np.float32(0.01)
    for t in range(split + 1, 300):
        mig[t] = [[1 - m, m], [m, 1 - m]]
    sample = np.zeros(G + 1, np.uint8)
    sample[[150, 299]] = 1
    return SimDesc(P, L, 0x1234, 0x5678, G, params, mig, None, sample, interval)


@pytest.mark.parametrize("split", [100, 101, 103])
def test_split_migration_extinction(gpu, oracle, split):
    d = split_scenario(split=split)
    g, o = gpu.run(d), oracle.run(d)
    assert_same(g, o, "split at %d" % split)
    assert g["samples"][-1]["extinct"] == [True, False]


_split_ref = {}


@pytest.mark.parametrize("interval", [0, 25, 100])
@pytest.mark.parametrize("kpl", [0, 1, 20, 64])
def test_several_populations_launch_length(gpu, oracle, interval, kpl):
    """two populations (split at 103: inside a block of four generations
    sharing one Philox output) with launches of 1, 20 and 64 generations and
    the library's own choice (0), compaction every 25 or 100 generations or
    never: every launch length gives the oracle's run bit for bit, so
    per-generation live counters in lanes up to 63, a population founded
    inside a long launch and tile tickets over long launches are covered"""
    d = split_scenario(split=103, interval=interval)
    if interval not in _split_ref:
        _split_ref[interval] = oracle.run(d)
    gpu.set_generations_per_launch(kpl)
    try:
        g = gpu.run(d)
    finally:
        gpu.set_generations_per_launch(0)
    assert_same(g, _split_ref[interval], "interval=%d kpl=%d" % (interval, kpl))


def test_prev_sim_restart(gpu, oracle):
    a = gpu.run(S.plain_small(G=100, compact_interval=10))
    last = a["samples"][-1]
    G = 80
    params = np.zeros((G + 1, 1), POP_GEN_DTYPE)
    params["N"] = 10000
    params["h"] = 0.5
    params["mu"] = 1e-9
    params["sel_intercept"] = -1e-4
    mig = np.ones((G + 1, 1, 1), np.float32)
    sample = np.zeros(G + 1, np.uint8)
    sample[0] = 1
    d = SimDesc(1, 2e7, 42, 7, G, params, mig, None, sample, 15, GF_INIT_PREV, start_generation=100,
                init_freq=last["freq"], init_ids=a["ids"], init_extinct=np.zeros(1, np.uint8))
    g, o = gpu.run(d), oracle.run(d)
    assert_same(g, o, "prev")
    assert g["samples"][0]["generation"] == 100 and g["samples"][-1]["generation"] == 180
    assert np.all(g["ids"]["origin_threadID"][: last["num_mutations"]] < 0)  # sampled start -> preserved


def test_spectrum_matches_oracle(gpu, oracle):
    d = S.plain_small(G=50)
    r = gpu.run(d)
    row = r["samples"][-1]["freq"][0]
    Nc = r["samples"][-1]["Nchrom_e"][0]
    h_g = gpu.population_frequency_histogram(row, Nc, d.num_sites)
    h_o = oracle.population_frequency_histogram(row, Nc, d.num_sites)
    assert np.array_equal(h_g, h_o)
    for n in [2, 50, 100, 1000, 5001]:
        s_g = gpu.site_frequency_spectrum(row, n, d.num_sites)
        s_o = oracle.site_frequency_spectrum(row, n, d.num_sites)
        assert np.array_equal(s_g.view(np.uint32), s_o.view(np.uint32)), n


def test_sfs_all_frequencies_matches_oracle(gpu, oracle):
    """sample SFS of rows covering the whole frequency range (uniform, near
    0 and 1, on the 1/N grid, at level boundaries (n+1)p integral, and 0 / 1
    themselves): each lane visits only the levels around its mode, and the
    result must equal the oracle's loop over every level bit for bit"""
    rng = np.random.default_rng(11)
    N = 20000
    rows = [rng.uniform(0, 1, 10000),
            np.exp(rng.uniform(np.log(1e-7), 0, 10000)),
            1 - np.exp(rng.uniform(np.log(1e-7), 0, 10000)),
            rng.integers(0, N + 1, 10000) / N]
    for n in [2, 7, 100, 1000, 5001]:
        grid = (np.arange(1, n + 1) / (n + 1)).astype(np.float32)
        row = np.concatenate(rows + [grid, np.array([0.0, 1.0, 0.5])]).astype(np.float32)
        s_g = gpu.site_frequency_spectrum(row, n, 1e9)
        s_o = oracle.site_frequency_spectrum(row, n, 1e9)
        assert np.array_equal(s_g.view(np.uint32), s_o.view(np.uint32)), n


def test_final_sfs_on_device_matches_host_path(gpu, oracle):
    """gf_result_site_frequency_spectrum bins the device-resident final
    state: identical to the spectrum of the copied-out sample (and to the
    oracle's), for every population of a multi-population run."""
    d = S.three_population(scale=0.1, L=2e8, G=300, preserve_every=0, sample_every=0)
    pairs = [(0, 50), (1, 20), (2, 100)]
    r = gpu.run(d, final_sfs=pairs)
    last = r["samples"][-1]
    for pop, n in pairs:
        host = gpu.site_frequency_spectrum(last["freq"][pop], n, d.num_sites)
        dev = r["final_sfs"][(pop, n)]
        assert np.array_equal(dev.view(np.uint32), host.view(np.uint32)), (pop, n)
        ref = oracle.site_frequency_spectrum(last["freq"][pop], n, d.num_sites)
        assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32)), (pop, n)
    with pytest.raises(RuntimeError):
        gpu.run(d, final_sfs=[(0, 10 ** 7)])


def test_joint_sfs_matches_oracle(gpu, oracle):
    """joint two-population SFS: GPU = oracle bit for bit, from host rows and
    from the device-resident final state; ragged counts and frequencies of 0
    and 1 included"""
    rng = np.random.default_rng(31)
    for M, na, nb in [(0, 5, 5), (1, 3, 9), (4097, 20, 30), (30000, 100, 60)]:
        fa = rng.random(M).astype(np.float32)
        fb = rng.random(M).astype(np.float32)
        if M > 10:
            fa[:5] = 0.0
            fb[3:9] = 1.0
        g = gpu.joint_site_frequency_spectrum(fa, fb, na, nb, 1e6)
        o = oracle.joint_site_frequency_spectrum(fa, fb, na, nb, 1e6)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), (M, na, nb)
    d = split_scenario(G=299)   # both populations alive at the end
    r = gpu.run(d, final_joint_sfs=[(0, 1, 40, 30)])
    last = r["samples"][-1]
    j = r["final_sfs"][(0, 1, 40, 30)]
    o = oracle.joint_site_frequency_spectrum(last["freq"][0], last["freq"][1], 40, 30, d.num_sites)
    assert np.array_equal(j.view(np.uint32), o.view(np.uint32))


def test_large_launch_chunked_tile_scan(gpu, oracle):
    """launches of more than two chunks of 16,384 tiles (about 1e7 slots:
    1.6e10 sites at N = 1e4 hold 6.7e6 live mutations and gain 3.2e5 a
    generation) take the chunk-sum + parallel tile scan; the run must equal
    the oracle's bit for bit, every generation's live count included"""
    d = S.constant(N=10000, L=1.6e10, G=12, s=S.selection_from_gamma(-2.0, 10000, 0.0), sample_every=6)
    g = gpu.run(d)
    oracle.set_threads(16)
    try:
        o = oracle.run(d)
    finally:
        oracle.set_threads(1)
    assert g["samples"][0]["num_mutations"] > 6_000_000
    assert_same(g, o, "chunked scan")


@pytest.mark.parametrize("Nc", [10000, 16384, 20000, 40961, 200000, 400000])
def test_population_histogram_bin_ranges(gpu, oracle, Nc):
    """population histogram in LDS passes over bin ranges: one range
    (Nc = 1e4), the old single-pass limit (16,384), a range boundary
    (40,961 = one full range + one bin) and the three-population scenario's
    sizes up to 4e5 bins; frequencies on the 1/Nc grid, off it, 0, 1 and
    the half-way points that round away from zero; GPU = oracle"""
    rng = np.random.default_rng(Nc)
    M = 3_000_000
    grid = rng.integers(0, Nc + 1, M // 2) / Nc
    neutral = np.exp(rng.uniform(np.log(1.0 / Nc), 0.0, M - M // 2 - 4))
    half = (np.array([0.5, 1.5, Nc - 0.5, Nc / 2 + 0.5]) / Nc)
    row = np.concatenate([grid, neutral, half]).astype(np.float32)
    row[:7] = [0.0, 1.0, 0.0, 1.0, 1.0 / Nc, (Nc - 1.0) / Nc, 0.5]
    h_g = gpu.population_frequency_histogram(row, Nc, 5e8)
    h_o = oracle.population_frequency_histogram(row, Nc, 5e8)
    assert h_g.shape == (Nc,)
    assert np.array_equal(h_g, h_o)
