"""Batched simulator: every simulation of a batch is bit-identical to the
plain gf_run of the same parameters, and the batch SFS matches the oracle."""
import numpy as np
import pytest

from parallelpopgen_amd import scenarios as S
from parallelpopgen_amd.capi import BATCH_SIM_DTYPE

pytestmark = pytest.mark.gpu


def make_batch(n, rng, G_scale=0.04, L_scale=1.0):
    sims = np.zeros(n, BATCH_SIM_DTYPE)
    N = np.exp(rng.uniform(np.log(500), np.log(5000), n)).astype(np.int32)
    gamma = rng.uniform(-50, 10, n)
    F = rng.uniform(0, 0.2, n).astype(np.float32)
    sims["N"] = N
    sims["F"] = F
    sims["h"] = rng.choice([0.0, 0.5, 1.0], n)
    sims["mu"] = 1e-9
    sims["s"] = [S.selection_from_gamma(g, Ni, Fi) for g, Ni, Fi in zip(gamma, N, F)]
    sims["s"][::5] = 0.0
    sims["num_sites"] = (np.exp(rng.uniform(np.log(2e6), np.log(5e7), n)) * L_scale).astype(np.float32)
    sims["seed1"] = 1000 + np.arange(n)
    sims["seed2"] = 77
    sims["num_generations"] = np.maximum(1, (4 * N / 10 * G_scale)).astype(np.int32)
    sims["num_generations"][3] = 0
    return sims


def plain_desc(sim, compact_interval):
    return S.constant(N=int(sim["N"]), L=float(sim["num_sites"]), G=int(sim["num_generations"]),
                      mu=float(sim["mu"]), F=float(sim["F"]), h=float(sim["h"]), s=np.float32(sim["s"]),
                      seed1=int(sim["seed1"]), seed2=int(sim["seed2"]), compact_interval=compact_interval)


@pytest.mark.parametrize("interval", [35, 8])
def test_batch_bit_identical_to_plain(gpu, interval):
    rng = np.random.default_rng(21)
    sims = make_batch(24, rng, G_scale=0.5)
    b = gpu.batch_run(sims, compact_interval=interval, sfs_n=50, keep_states=True)
    for k, sim in enumerate(sims):
        d = plain_desc(sim, 35)
        r = gpu.run(d)
        last = r["samples"][-1]
        f, ids = b["states"][k]
        assert b["num_mutations"][k] == last["num_mutations"], k
        assert np.array_equal(f.view(np.uint32), last["freq"][0].view(np.uint32)), k
        assert np.array_equal(ids, r["ids"]), k
        sfs = gpu.site_frequency_spectrum(last["freq"][0], 50, d.num_sites)
        assert np.array_equal(b["sfs"][k].view(np.uint32), sfs.view(np.uint32)), k


def test_batch_sfs_matches_oracle(gpu, oracle):
    rng = np.random.default_rng(22)
    sims = make_batch(6, rng, G_scale=0.2)
    b = gpu.batch_run(sims, compact_interval=35, sfs_n=50, keep_states=True)
    for k, sim in enumerate(sims):
        d = plain_desc(sim, 35)
        o = oracle.run(d)
        last = o["samples"][-1]
        assert np.array_equal(b["states"][k][0].view(np.uint32), last["freq"][0].view(np.uint32)), k
        s_o = oracle.site_frequency_spectrum(last["freq"][0], 50, d.num_sites)
        assert np.array_equal(b["sfs"][k].view(np.uint32), s_o.view(np.uint32)), k


def test_batch_live_count_matches_plain(gpu):
    rng = np.random.default_rng(23)
    sims = make_batch(10, rng, G_scale=0.3)
    b = gpu.batch_run(sims, compact_interval=35, sfs_n=20)
    tot = 0
    for sim in sims:
        r = gpu.run(plain_desc(sim, 35))
        tot += int(r["live"][1:].sum())
    assert b["live_mutation_generations"] == tot


def test_batch_size_change_matches_plain(gpu):
    """N1 != N: the batch draws the starting state at N and simulates every
    generation from 1 on at N1 (an instantaneous size change), exactly as
    gf_run does with the size table N(0) = N, N(g >= 1) = N1."""
    rng = np.random.default_rng(24)
    sims = make_batch(12, rng, G_scale=0.5)
    sims["N1"] = np.maximum(1, (sims["N"] * np.exp(rng.uniform(-1.5, 2.5, len(sims))))).astype(np.int32)
    sims["N1"][0] = sims["N"][0]
    b = gpu.batch_run(sims, compact_interval=8, sfs_n=40, keep_states=True)
    for k, sim in enumerate(sims):
        d = plain_desc(sim, 35)
        d.params["N"][1:] = sim["N1"]
        r = gpu.run(d)
        last = r["samples"][-1]
        f, ids = b["states"][k]
        assert b["num_mutations"][k] == last["num_mutations"], k
        assert np.array_equal(f.view(np.uint32), last["freq"][0].view(np.uint32)), k
        assert np.array_equal(ids, r["ids"]), k


def test_batch_requires_compaction(gpu):
    """every block of a batch compacts: a compaction interval of 0 (never
    compact) is refused rather than silently ignored"""
    sims = make_batch(4, np.random.default_rng(25))
    with pytest.raises(gpu.GoFishError):
        gpu.batch_run(sims, compact_interval=0, sfs_n=20)


@pytest.mark.parametrize("K", [1, 7, 32])
def test_batch_block_length_invariance(gpu, K):
    """the block length (graph replays of K generations, K capped by
    gf_set_generations_per_launch; otherwise the autotuner's choice) never
    changes a simulation: every K gives gf_run's states bit for bit"""
    rng = np.random.default_rng(26)
    sims = make_batch(16, rng, G_scale=0.5)
    gpu.set_generations_per_launch(K)
    try:
        b = gpu.batch_run(sims, compact_interval=35, sfs_n=30, keep_states=True)
    finally:
        gpu.set_generations_per_launch(0)
    assert b["tuning"]["generations_per_block"] == K
    for k, sim in enumerate(sims):
        r = gpu.run(plain_desc(sim, 35))
        last = r["samples"][-1]
        f, ids = b["states"][k]
        assert b["num_mutations"][k] == last["num_mutations"], (K, k)
        assert np.array_equal(f.view(np.uint32), last["freq"][0].view(np.uint32)), (K, k)
        assert np.array_equal(ids, r["ids"]), (K, k)


@pytest.mark.parametrize("scale", [1.0, 0.02])
def test_batch_growth_relayouts(gpu, scale):
    """populations that grow 20-80 fold after generation 0 keep more
    mutations than their equilibrium start: with regions cut to 2 % of the
    library's sizing they outgrow their tiles, stall, and are laid out again
    (several times) while the others go on; every result still equals
    gf_run's, with the default sizing too"""
    rng = np.random.default_rng(27)
    sims = make_batch(20, rng, G_scale=0.6)
    sims["N1"] = (sims["N"] * rng.uniform(20, 80, len(sims))).astype(np.int32)
    gpu.set_batch_region_scale(scale)
    try:
        b = gpu.batch_run(sims, compact_interval=16, sfs_n=40, keep_states=True)
    finally:
        gpu.set_batch_region_scale(1.0)
    if scale < 1:
        assert b["tuning"]["relayouts"] >= 2
    for k, sim in enumerate(sims):
        d = plain_desc(sim, 35)
        d.params["N"][1:] = sim["N1"]
        r = gpu.run(d)
        last = r["samples"][-1]
        f, ids = b["states"][k]
        assert b["num_mutations"][k] == last["num_mutations"], k
        assert np.array_equal(f.view(np.uint32), last["freq"][0].view(np.uint32)), k
        assert np.array_equal(ids, r["ids"]), k


@pytest.mark.parametrize("interval", [8, 35])
def test_random_batch_bit_identical_to_plain(gpu, interval):
    """A seeded random batch of 96 simulations over a wider range than the
    sweep: sizes 40 to 5,000 (below 100 the drift is the exact binomial
    inversion), instantaneous size changes, F from 0 to 1 (haploid F = 1
    included), h from 0 to 1, gamma from -20 to 10, 0 to 300 generations;
    every simulation's final state and IDs equal gf_run's bit for bit."""
    rng = np.random.default_rng(4242 + interval)
    n = 96
    sims = np.zeros(n, BATCH_SIM_DTYPE)
    N = np.exp(rng.uniform(np.log(40), np.log(5000), n)).astype(np.int32)
    F = rng.choice([0.0, 1.0, 0.5], n).astype(np.float32)
    F[::3] = rng.uniform(0, 1, len(F[::3])).astype(np.float32)
    sims["N"] = N
    sims["N1"] = np.where(rng.uniform(size=n) < 0.3, np.maximum(40, N * np.exp(rng.uniform(-1, 2, n))), N).astype(np.int32)
    sims["F"] = F
    sims["h"] = rng.choice([0.0, 0.5, 1.0, 0.25], n)
    sims["mu"] = np.float32(1e-8)
    gamma = rng.uniform(-20, 10, n)
    sims["s"] = [S.selection_from_gamma(g, Ni, Fi) for g, Ni, Fi in zip(gamma, N, F)]
    sims["s"][::4] = 0.0
    sims["num_sites"] = (rng.uniform(20, 200, n) / (2e-8 * np.maximum(N, sims["N1"]))).astype(np.float32)
    sims["seed1"] = rng.integers(1, 2 ** 31 - 1, n)
    sims["seed2"] = rng.integers(1, 2 ** 31 - 1, n)
    sims["num_generations"] = rng.integers(0, 300, n)
    b = gpu.batch_run(sims, compact_interval=interval, sfs_n=30, keep_states=True)
    for k, sim in enumerate(sims):
        r = gpu.run(S.sweep_plain_desc(sim, interval))
        last = r["samples"][-1]
        f, ids = b["states"][k]
        assert b["num_mutations"][k] == last["num_mutations"], k
        assert np.array_equal(f.view(np.uint32), last["freq"][0].view(np.uint32)), k
        assert np.array_equal(ids, r["ids"]), k
