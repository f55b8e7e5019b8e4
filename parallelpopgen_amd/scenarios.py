"""Model tables for the benchmark scenarios, built the way the C++ API layer
(include/go_fish.cuh) builds them from Sim_Model functors: one gf_pop_gen row
per (generation, population), a P x P migration matrix per generation, and
the preserve / sample flags.

    plain_small()      1 population, N=10,000, L=2e7, neutral, 1,000 generations
    plain_large()      1 population, N=100,000, L=5e8, gamma=-10, h=0.5, 5,000 generations
    three_population() 3 populations, growth, bottleneck, sine-wave selection,
                       migration 1e-4, F=0.1 in the third, 2,000 generations
    constant(...)      any single- or multi-population constant scenario
"""
import math

import numpy as np

from .capi import POP_GEN_DTYPE, SimDesc, GF_INIT_MSE

SEED2_DEFAULT = 0xdecafbad  # reference sim_constants default (inline_go_fish_data_struct.hpp:213)


def _f32(x):
    return np.float32(x)


_LIBM = None


def _libm():
    """the C library's float functions, which the C++ API's functors call
    on the host (expf in demography_exponential_growth, sinf in the sine-wave
    models): numpy's float32 exp and sin may differ from them in the last
    place"""
    global _LIBM
    if _LIBM is None:
        import ctypes
        import ctypes.util
        lib = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
        for name in ("expf", "sinf"):
            getattr(lib, name).restype = ctypes.c_float
            getattr(lib, name).argtypes = [ctypes.c_float]
        _LIBM = lib
    return _LIBM


def expf(x):
    return np.float32(_libm().expf(float(np.float32(x))))


def sinf(x):
    return np.float32(_libm().sinf(float(np.float32(x))))


def selection_from_gamma(gamma, N0, F0):
    """s = gamma / (2*N/(1+F)) in float, as Sim_Model::selection_constant's
    effective-selection constructor (template_inline_simulation_functors.cuh:52)."""
    return _f32(_f32(gamma) / _f32(_f32(2 * int(N0)) / _f32(1.0 + _f32(F0))))


def migration_constant_equal(m, P):
    """Sim_Model::migration_constant_equal (functors :432-435)."""
    M = np.full((P, P), _f32(m) if P > 1 else 0, np.float32)
    np.fill_diagonal(M, _f32(1.0) - _f32((P - 1) * _f32(m)))
    return M


def constant(N, L, G, mu=1e-9, F=0.0, h=0.5, s=0.0, P=1, m=0.0, seed1=0xbeeff00d, seed2=SEED2_DEFAULT,
             compact_interval=35, init_mode=GF_INIT_MSE, preserve_every=0, sample_every=0, device=-1, **kw):
    params = np.zeros((G + 1, P), POP_GEN_DTYPE)
    params["N"] = N
    params["F"] = F
    params["h"] = h
    params["mu"] = mu
    params["sel_slope"] = 0.0
    params["sel_intercept"] = s
    mig = np.broadcast_to(migration_constant_equal(m, P), (G + 1, P, P)).copy()
    preserve = np.zeros(G + 1, np.uint8)
    sample = np.zeros(G + 1, np.uint8)
    if preserve_every:
        preserve[np.arange(G + 1) % preserve_every == 0] = 1
    if sample_every:
        sample[np.arange(G + 1) % sample_every == 0] = 1
    return SimDesc(P, L, seed1, seed2, G, params, mig, preserve, sample, compact_interval, init_mode,
                   device=device, **kw)


def plain_small(seed1=42, seed2=SEED2_DEFAULT, G=1000, L=2e7, N=10000, compact_interval=35):
    """Plain method, small: N=10,000 diploids, F=0, mu=1e-9, L=2e7, neutral
    (h=0.5), neutral-equilibrium start, 1,000 generations, seed1=42."""
    return constant(N=N, L=L, G=G, mu=1e-9, F=0.0, h=0.5, s=0.0, seed1=seed1, seed2=seed2,
                    compact_interval=compact_interval)


def plain_large(seed1=0xbeeff00d, seed2=SEED2_DEFAULT, G=5000, L=5e8, N=100000, gamma=-10.0, compact_interval=35):
    """Plain method, large: N=100,000, F=0, mu=1e-9, L=5e8, gamma=2Ns=-10,
    h=0.5, 5,000 generations."""
    s = selection_from_gamma(gamma, N, 0.0)
    return constant(N=N, L=L, G=G, mu=1e-9, F=0.0, h=0.5, s=s, seed1=seed1, seed2=seed2,
                    compact_interval=compact_interval)


def three_population(seed1=0xbeeff00d, seed2=SEED2_DEFAULT, G=2000, L=2e8, N0=20000, scale=1.0,
                     compact_interval=35, preserve_every=100, sample_every=0):
    """Three populations (0-based indices 0, 1, 2 for the scenario's
    populations 1, 2, 3):
      index 0: exponential growth 2e4 -> 2e5 over generations 0..500
               (rate ln(10)/500, Sim_Model::demography_exponential_growth),
               constant 2e5 afterwards;
      index 1: 2e4, bottleneck to 2e3 for generations [500, 600), then 2e4;
      index 2: constant 2e4 with F = 0.1;
      migration_constant_equal(1e-4, 3); mu = 1e-9; h = 0.3 everywhere;
      selection_sine_wave with gamma amplitude 5, period 200 generations,
      offset 0 (A = 5 / (2*N0/(1+F0)) from population 0 at generation 0).
    `scale` multiplies N (and divides generations-independent nothing):
    used to make CPU-sized parity cases."""
    P = 3
    N0 = int(N0 * scale)
    params = np.zeros((G + 1, P), POP_GEN_DTYPE)
    rate = _f32(math.log(10.0) / 500.0)
    g = np.arange(G + 1)
    # (int)roundf(N0 * expf(rate * t)), half away from zero
    grow = np.array([int(math.floor(float(_f32(_f32(N0) * expf(_f32(rate * _f32(t))))) + 0.5)) for t in g], np.int64)
    grow[g >= 500] = N0 * 10
    params["N"][:, 0] = grow
    params["N"][:, 1] = np.where((g >= 500) & (g < 600), N0 // 10, N0)
    params["N"][:, 2] = N0
    params["F"][:, 2] = 0.1
    params["h"] = 0.3
    params["mu"] = 1e-9
    A = _f32(_f32(5.0) / _f32(_f32(2 * N0) / _f32(1.0)))
    pi = _f32(2 * math.pi / 200.0)
    s_t = np.array([_f32(_f32(A * sinf(_f32(pi * _f32(t)))) + _f32(0.0)) for t in g], np.float32)
    params["sel_slope"] = 0.0
    params["sel_intercept"] = s_t[:, None]
    mig = np.broadcast_to(migration_constant_equal(1e-4, P), (G + 1, P, P)).copy()
    preserve = np.zeros(G + 1, np.uint8)
    if preserve_every:
        preserve[(g % preserve_every == 0) & (g > 0)] = 1
    sample = np.zeros(G + 1, np.uint8)
    if sample_every:
        sample[(g % sample_every == 0) & (g > 0)] = 1
    return SimDesc(P, L, seed1, seed2, G, params, mig, preserve, sample, compact_interval, GF_INIT_MSE)


def expected_neutral_mse(N, L, mu=1e-9, F=0.0):
    """E[number of segregating mutations] at neutral equilibrium:
    2 mu L * Nc * H(Nc-1) (sum over classes of 2 mu L / (i/Nc))."""
    Nc = int(2 * N / (1 + F))
    return 2 * mu * L * Nc * sum(1.0 / i for i in range(1, Nc))


def harmonic(n):
    """H(n) = sum_{i=1..n} 1/i for an integer array n >= 1 (exact sum up to
    100, asymptotic expansion beyond: error < 1e-12)."""
    n = np.asarray(n, np.float64)
    small = np.cumsum(1.0 / np.arange(1, 101))
    big = np.log(np.maximum(n, 1)) + 0.5772156649015329 + 1 / (2 * n) - 1 / (12 * n ** 2) + 1 / (120 * n ** 4)
    return np.where(n <= 100, small[np.clip(n.astype(np.int64), 1, 100) - 1], big)


def parameter_sweep(n=4096, seed=2024, seed_base=1000, mu=1e-9, sfs_n=50):
    """Batched parameter sweep of single-population simulations (one row of
    gf_batch_sim per simulation):
      N ~ log-uniform[1e3, 1e5] diploids, gamma = 2Ns ~ U[-50, 10]
      (s = gamma / Nc as Sim_Model::selection_constant), h in {0, 0.5, 1},
      F ~ U[0, 0.2], mu = 1e-9, 4N/10 generations from mutation-selection
      equilibrium, seeds (seed_base + index, 0x5eed).
    The number of sites L is set per simulation so that the expected
    number of segregating sites at neutral equilibrium,
    2 mu L Nc H(Nc - 1), is log-uniform in [1e3, 1e6]."""
    from .capi import BATCH_SIM_DTYPE
    rng = np.random.default_rng(seed)
    N = np.exp(rng.uniform(math.log(1e3), math.log(1e5), n)).astype(np.int32)
    gamma = rng.uniform(-50.0, 10.0, n)
    h = rng.choice([0.0, 0.5, 1.0], n)
    F = rng.uniform(0.0, 0.2, n).astype(np.float32)
    target = np.exp(rng.uniform(math.log(1e3), math.log(1e6), n))
    Nc = np.floor((2.0 * N) / (1.0 + F.astype(np.float64))).astype(np.int64)
    L = target / (2.0 * mu * Nc * harmonic(Nc - 1))
    sims = np.zeros(n, BATCH_SIM_DTYPE)
    sims["N"] = N
    sims["F"] = F
    sims["h"] = h
    sims["mu"] = mu
    sims["s"] = [selection_from_gamma(g, Ni, Fi) for g, Ni, Fi in zip(gamma, N, F)]
    sims["num_sites"] = L.astype(np.float32)
    sims["seed1"] = seed_base + np.arange(n)
    sims["seed2"] = 0x5eed
    sims["num_generations"] = (4 * N.astype(np.int64) // 10).astype(np.int32)
    return sims


def sweep_plain_desc(sim, compact_interval=35):
    """the gf_run description of one row of a batched sweep"""
    d = constant(N=int(sim["N"]), L=float(sim["num_sites"]), G=int(sim["num_generations"]), mu=float(sim["mu"]),
                 F=float(sim["F"]), h=float(sim["h"]), s=np.float32(sim["s"]), seed1=int(sim["seed1"]),
                 seed2=int(sim["seed2"]), compact_interval=compact_interval)
    if sim["N1"] > 0:
        d.params["N"][1:] = sim["N1"]
    return d
