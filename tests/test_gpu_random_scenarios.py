"""Seeded random small scenarios, GPU run_sim path against the CPU oracle
bit for bit (every time sample, the final IDs, the per-generation live
counts).  Each case draws: 1 to 4 populations; sizes from 40 to 5,000
(below 100 the drift is the exact binomial inversion) that stay constant,
ramp or dip; inbreeding F in [0, 1] (including the haploid F = 1);
dominance h in [0, 1]; linear frequency-dependent selection with gamma =
2Ns from -20 to 10 at the starting size, or beyond the clamp s >= -1 for
part of the frequency range (then from a blank start); equal migration; a
mutation-selection-equilibrium (or blank) start; compaction every 0, 1, 3, 17, 35 or
64 generations; preserve and sample pulses; and a launch length of the
library's choice or 1, 7 or 64 generations.  The same seed gives the same
cases on every run."""
import numpy as np
import pytest

from parallelpopgen_amd import scenarios as S
from parallelpopgen_amd.capi import GF_INIT_BLANK, GF_INIT_MSE, POP_GEN_DTYPE, SimDesc
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

CASES = 64


def random_scenario(i):
    rng = np.random.default_rng(7000 + i)
    P = int(rng.choice([1, 1, 2, 3, 4]))
    clamp = False
    G = int(rng.integers(40, 250))
    g = np.arange(G + 1)
    params = np.zeros((G + 1, P), POP_GEN_DTYPE)
    for p in range(P):
        n0 = int(rng.choice([int(rng.integers(40, 100)), int(rng.integers(100, 5000))]))
        shape = int(rng.integers(0, 3))
        if shape == 0:
            N = np.full(G + 1, n0)
        elif shape == 1:                      # linear ramp to another size
            n1 = int(rng.integers(40, 5000))
            N = np.rint(n0 + (n1 - n0) * g / G).astype(np.int64)
        else:                                 # a dip for a window of generations
            a = int(rng.integers(1, G))
            b = min(G + 1, a + int(rng.integers(5, 60)))
            N = np.full(G + 1, n0)
            N[a:b] = max(40, n0 // int(rng.integers(2, 10)))
        params["N"][:, p] = N
        params["F"][:, p] = float(rng.choice([0.0, 1.0, np.float32(rng.uniform(0, 1))]))
        params["h"][:, p] = float(rng.choice([0.0, 0.5, 1.0, np.float32(rng.uniform(0, 1))]))
        params["mu"][:, p] = np.float32(10 ** rng.uniform(-9, -7.5))
        kind = int(rng.integers(0, 4))
        unit = 1.0 / (2.0 * n0)                 # s for gamma = 1 at the starting size
        if kind == 1:
            params["sel_intercept"][:, p] = np.float32(rng.uniform(-20, 10) * unit)
        elif kind == 2:
            params["sel_slope"][:, p] = np.float32(rng.uniform(-20, 20) * unit)
            params["sel_intercept"][:, p] = np.float32(rng.uniform(-10, 5) * unit)
        elif kind == 3:                         # clamped at s = -1 for part of the range
            params["sel_slope"][:, p] = np.float32(-2.0)
            params["sel_intercept"][:, p] = np.float32(0.5)
            clamp = True
    m = float(rng.choice([0.0, np.float32(rng.uniform(0, 0.02))])) if P > 1 else 0.0
    mig = np.broadcast_to(S.migration_constant_equal(m, P), (G + 1, P, P)).copy()
    nmax = float(params["N"].max())
    mu = float(params["mu"].max())
    L = float(np.float32(rng.uniform(20, 200) / (2 * nmax * mu)))
    interval = int(rng.choice([0, 1, 3, 17, 35, 64]))
    # the equilibrium start integrates selection over frequency: strong
    # selection (the clamp case) starts blank
    init = GF_INIT_BLANK if (clamp or rng.uniform() < 0.2) else GF_INIT_MSE
    preserve = np.zeros(G + 1, np.uint8)
    pe = int(rng.choice([0, 0, 13, 50]))
    if pe:
        preserve[(g % pe == 0) & (g > 0)] = 1
    sample = np.zeros(G + 1, np.uint8)
    se = int(rng.choice([0, 0, 7, 40]))
    if se:
        sample[(g % se == 0) & (g > 0)] = 1
    seed1, seed2 = (int(v) for v in rng.integers(1, 2 ** 31 - 1, 2))
    kpl = int(rng.choice([0, 0, 1, 7, 64]))
    d = SimDesc(P, L, seed1, seed2, G, params, mig, preserve, sample, interval, init)
    return d, kpl


@pytest.mark.parametrize("i", range(CASES))
def test_random_scenario_matches_oracle(gpu, oracle, i):
    d, kpl = random_scenario(i)
    want = oracle.run(d)
    gpu.set_generations_per_launch(kpl)
    try:
        got = gpu.run(d)
    finally:
        gpu.set_generations_per_launch(0)
    assert_same(got, want, "case %d (P=%d, G=%d, interval=%d, kpl=%d)" % (i, d.P, d.G, d.compact_interval, kpl))
