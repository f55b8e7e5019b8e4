"""Pins the oracle's model functions independently of the GPU library.

Every check below compares the oracle (oracle/gofish_oracle.cpp) with
something it was not derived from: the reference's own formula written out
in float64 in the reference's algebraic form, closed forms of special cases,
sampling moments, or a brute-force Python loop.  Each check is a function of
an oracle library handle; test_planted_mistakes_are_caught builds the oracle
with one plausible slip at a time (ORC_PLANT) and shows that the check named
for it fails.

Reference formulas (paths relative to the reference tree):
  Eq. 13 with F and h, migration   3P/_internal/go_fish_impl.cuh:93-106
  mse(), integrand, trapezoid,     3P/_internal/go_fish_impl.cuh:20-71, :201-223
  class lambda and its quirks
  approximate binomial drift       3P/_internal/shared.cu:10-18
  keep rule, preserve negation     3P/_internal/go_fish_impl.cuh:137-176,
                                   3P/_internal/go_fish_impl.cu:147-189
  population histogram            3P/_internal/spectrum.cu:74-83
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from model_reference import mse_lambda_ref
from oracle import oracle as O
from parallelpopgen_amd.capi import MUTID_DTYPE, POP_GEN_DTYPE

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
U = 2.0 ** -24   # float32 unit roundoff


# ---------------------------------------------------------------- helpers

def eq13_ref(x, s, h, F):
    """Eq. 13 exactly as the reference writes it (go_fish_impl.cuh:105), in
    float64, with s = max(s, -1) (:104)."""
    x, s, h, F = (np.asarray(v, np.float64) for v in (x, s, h, F))
    s = np.maximum(s, -1.0)
    num = s * x * x + x + (F + h - h * F) * s * x * (1 - x)
    den = x * x * s + (F + 2 * h - 2 * h * F) * s * x * (1 - x) + 1
    return num / den


def eq13_tolerance(x, s, h, F):
    """Absolute error bound of a float32 evaluation of Eq. 13 with a handful
    of roundings in numerator and denominator: 8 u times the sum of the
    magnitudes of the terms, over |den| (the conditioning of the two
    sums), times the result's scale."""
    x, s, h, F = (np.asarray(v, np.float64) for v in (x, s, h, F))
    s = np.maximum(s, -1.0)
    c1 = F + h - h * F
    c2 = F + 2 * h - 2 * h * F
    mag_n = x + np.abs(s) * x * (x + np.abs(c1) * (1 - x))
    mag_d = 1 + np.abs(s) * x * (x + np.abs(c2) * (1 - x))
    num = x + s * x * (x + c1 * (1 - x))
    den = 1 + s * x * (x + c2 * (1 - x))
    return 8 * U * (mag_n + np.abs(num / den) * mag_d) / np.abs(den) + 2.0 ** -149


def f32(v):
    return np.asarray(v, np.float32)


# ---------------------------------------------------------------- checks

def check_eq13(lib):
    """Eq. 13 against the reference formula in float64 over random (x, s,
    h, F); mean = x' N and var = (1 - x') mean of the draw; closed forms."""
    rng = np.random.default_rng(13)
    n = 24
    # a grid of (s, h, F), each with many x: |s| up to 0.9, and s < -1
    S = f32(np.concatenate([rng.uniform(-0.9, 1.0, n), [-0.5, -1e-4, 1e-4, 0.05, -3.0, -1.0]]))
    H = f32(np.concatenate([rng.uniform(0, 1, n), [0.0, 0.5, 1.0, 0.3, 0.5, 0.0]]))
    Fv = f32(np.concatenate([rng.uniform(0, 1, n), [0.0, 0.0, 0.1, 1.0, 0.2, 0.0]]))
    Nc = 200000
    for s, h, F in zip(S, H, Fv):
        x = f32(np.concatenate([rng.uniform(0, 1, 2000), np.arange(1, 200) / np.float32(Nc), [0.5, 1.0, 0.0]]))
        xs, mean, var = O.pre_draw(x.reshape(-1, 1), [Nc], [F], [h], [0.0], [s], [[1.0]], lib=lib)
        xs, mean, var = xs[:, 0], mean[:, 0], var[:, 0]
        with np.errstate(invalid="ignore"):
            ref = eq13_ref(x, s, h, F)
            tol = eq13_tolerance(x, s, h, F)
        # fixed and lethal (x = 1, s <= -1) is 0/0 in the reference too
        nan = ~np.isfinite(ref)
        assert np.all(np.isnan(xs[nan])) and np.all((x[nan] == 1) & (s <= -1)), ("Eq. 13 at 0/0", float(s))
        x, xs, mean, var, ref, tol = (a[~nan] for a in (x, xs, mean, var, ref, tol))
        bad = np.abs(xs - ref) > tol
        assert not bad.any(), ("Eq. 13", float(s), float(h), float(F), x[bad][:3], xs[bad][:3], ref[bad][:3])
        # the draw's moments, from the float64 x': mean N x', var N x'(1 - x')
        m_ref = ref * Nc
        v_ref = (1 - ref) * m_ref
        assert np.all(np.abs(mean - m_ref) <= tol * Nc + 2 * U * m_ref), ("mean", float(s), float(h), float(F))
        assert np.all(np.abs(var - v_ref) <= 2 * tol * Nc + 4 * U * m_ref + 1e-30), ("var", float(s), float(h), float(F))
    # closed forms (float64, different algebraic forms of the same model)
    x = f32(rng.uniform(1e-6, 1, 5000))
    for s in f32([-0.7, -0.05, 0.01, 0.4, 1.0]):
        sd = np.float64(s)
        cases = {
            "F=1 (haploid)": (0.37, 1.0, x * (1 + sd) / (1 + sd * x)),
            "h=0": (0.0, 0.0, (sd * x * x + x) / (sd * x * x + 1)),
            "h=1": (1.0, 0.0, x * (1 + sd) / (1 + 2 * sd * x - sd * x * x)),
            "h=1/2": (0.5, 0.0, x * (1 + sd / 2 + sd * x / 2) / (1 + sd * x)),
        }
        for name, (h, F, closed) in cases.items():
            xs = O.pre_draw(x.reshape(-1, 1), [Nc], [F], [h], [0.0], [s], [[1.0]], lib=lib)[0][:, 0]
            tol = eq13_tolerance(x, s, h, F)
            assert np.all(np.abs(xs - closed) <= tol), (name, float(s), np.max(np.abs(xs - closed) / tol))
    # s = 0: x' = x exactly
    xs = O.pre_draw(x.reshape(-1, 1), [Nc], [0.3], [0.2], [0.0], [0.0], [[1.0]], lib=lib)[0][:, 0]
    assert np.array_equal(xs, x), "s = 0"
    # s below -1 acts as s = -1; with F = 1 that is x' = 0 for x < 1
    a = O.pre_draw(x.reshape(-1, 1), [Nc], [1.0], [0.5], [0.0], [-2.5], [[1.0]], lib=lib)[0][:, 0]
    b = O.pre_draw(x.reshape(-1, 1), [Nc], [1.0], [0.5], [0.0], [-1.0], [[1.0]], lib=lib)[0][:, 0]
    assert np.array_equal(a, b), "s clamped at -1"
    assert np.all(a[x < 1] == 0.0), "s = -1, F = 1"


def check_migration(lib):
    """Frequency after migration: the weighted sum over source populations
    (m[k][j] = fraction of population j's parents from k); linear selection
    s = slope x + intercept evaluated at that frequency."""
    rng = np.random.default_rng(7)
    P = 3
    m = rng.uniform(0, 1, (P, P)) ** 3
    m = f32(m / m.sum(0, keepdims=True))             # columns sum to 1
    x = f32(rng.uniform(0, 1, (4000, P)))
    Nc = [40000, 30000, 20000]
    xs, _, _ = O.pre_draw(x, Nc, [0.0] * P, [0.5] * P, [0.0] * P, [0.0] * P, m, lib=lib)
    ref = x.astype(np.float64) @ m.astype(np.float64)
    assert np.all(np.abs(xs - ref) <= 2 * P * U * (np.abs(x) @ np.abs(m))), "migration sum"
    sa, sb, h, F = f32([0.3, -0.2, 0.1]), f32([-0.1, 0.05, 0.0]), f32([0.5, 0.0, 1.0]), f32([0.0, 0.1, 0.2])
    xs, _, _ = O.pre_draw(x, Nc, F, h, sa, sb, m, lib=lib)
    xm = f32(x.astype(np.float64) @ m.astype(np.float64))
    for p in range(P):
        s = np.float64(sa[p]) * xm[:, p] + np.float64(sb[p])
        ref = eq13_ref(xm[:, p], s, h[p], F[p])
        tol = eq13_tolerance(xm[:, p], s, h[p], F[p]) + 4 * P * U
        assert np.all(np.abs(xs[:, p] - ref) <= tol), ("migration + linear selection", p)


def _moments(j):
    j = j.astype(np.float64)
    mu = j.mean()
    d = j - mu
    var = (d * d).mean()
    m4 = (d ** 4).mean()
    return mu, var, np.sqrt(var / len(j)), np.sqrt(max(m4 - var * var, 1e-12) / len(j))


def check_drift_moments(lib):
    """Drawn count j over 1e5 mutation IDs: mean within 5 sigma of N x',
    variance within 5 sigma of N x'(1 - x') (+1/12 of the rounding in the
    normal regime), in the normal regime, both Poisson tails, and the
    binomial tail of N < 100; 0 and 1 absorbing."""
    n = 100000
    ids = np.zeros(n, MUTID_DTYPE)
    ids["origin_generation"] = 5
    ids["origin_threadID"] = np.arange(1, n + 1)
    cases = [  # (Nc, x, s, h, F, extra variance)
        (20000, 0.3, -0.01, 0.5, 0.0, 1.0 / 12),        # normal approximation
        (20000, 0.2, 0.02, 0.0, 0.1, 1.0 / 12),
        (20000, 3.0 / 20000, 0.0, 0.5, 0.0, 0.0),       # Poisson, lower tail
        (20000, 1.0 - 2.0 / 20000, 0.0, 0.5, 0.0, 0.0),  # Poisson, upper tail
        (50, 4.0 / 50, 0.0, 0.5, 0.0, 0.0),              # binomial (N < 100)
    ]
    for Nc, x0, s, h, F, extra in cases:
        x = np.full((n, 1), x0, np.float32)
        xs = eq13_ref(np.float32(x0), np.float32(s), np.float32(h), np.float32(F))
        out = O.drift_step(x, ids, 17, 0x1234567, 0x7654321, [Nc], [F], [h], [0.0], [s], [[1.0]], lib=lib)
        j = np.rint(out[:, 0].astype(np.float64) * Nc)
        mu, var, se_mu, se_var = _moments(j)
        e_mu, e_var = Nc * xs, Nc * xs * (1 - xs) + extra
        assert abs(mu - e_mu) < 5 * se_mu, ("drift mean", Nc, x0, s, mu, e_mu)
        assert abs(var - e_var) < 5 * se_var, ("drift variance", Nc, x0, s, var, e_var)
    for x0 in (0.0, 1.0):
        for s, h, F in ((0.0, 0.5, 0.0), (-1.0, 0.0, 0.0), (-0.99, 0.0, 0.0), (0.5, 1.0, 0.3), (-0.2, 0.5, 1.0)):
            if x0 == 1.0 and s == -1.0:
                continue
            x = np.full((5000, 1), x0, np.float32)
            out = O.drift_step(x, ids[:5000], 9, 11, 12, [20000], [F], [h], [0.0], [s], [[1.0]], lib=lib)
            assert np.all(out == x0), ("absorbing", x0, s, h, F)
    # fixed and lethal (s = -1): Eq. 13 is 0/0 there; the reference's NaN
    # mean draws NaN, which its float -> int conversion turns into 0
    x = np.full((100, 1), 1.0, np.float32)
    out = O.drift_step(x, ids[:100], 9, 11, 12, [20000], [0.0], [0.5], [0.0], [-1.0], [[1.0]], lib=lib)
    assert np.all(out == 0.0), "fixed and lethal"


def check_mse_lambda(lib):
    """Class lambda of the equilibrium start against the float64
    restatement over h in {0, 1/2, 1}, F in {0, 0.1, 1}, gamma in {-10, 4}
    (relative 2e-4: float32 exponent, trapezoids and a chunked float scan
    over 2e4 classes); a selection that falls below -1; the neutral total
    2 mu L Nc H(Nc - 1)."""
    N, mu, L = 10000, 1e-9, 2e8
    for h in (0.0, 0.5, 1.0):
        for F in (0.0, 0.1, 1.0):
            for gamma in (-10.0, 4.0):
                Nc = int(np.float32(2 * N) / np.float32(1 + F))
                s = np.float32(gamma / Nc)
                rec = np.zeros(1, POP_GEN_DTYPE)
                rec[0] = (N, F, h, mu, 0.0, s)
                got = O.mse_lambda(rec[0], L, lib=lib).astype(np.float64)
                ref = mse_lambda_ref(N, F, h, mu, L, lambda y: np.full_like(y, np.float64(s)))
                assert len(got) == len(ref) == Nc - 1
                rel = np.abs(got / ref - 1)
                assert rel.max() < 2e-4, ("mse lambda", h, F, gamma, int(np.argmax(rel)), rel.max())
    # linear selection reaching below -1 (tiny N keeps the exponent finite)
    N = 20
    rec = np.zeros(1, POP_GEN_DTYPE)
    rec[0] = (N, 0.0, 0.5, mu, -3.0, 0.5)
    got = O.mse_lambda(rec[0], L, lib=lib).astype(np.float64)
    ref = mse_lambda_ref(N, 0.0, 0.5, mu, L, lambda y: -3.0 * np.float64(np.float32(y)) + 0.5)
    # (classes near x = 1 have lambda below float's normal range, 1e-38)
    assert np.all(np.abs(got - ref) <= 2e-5 * ref + 1e-36), ("mse lambda, s < -1", np.abs(got / ref - 1).max())
    # neutral: 2 mu L / x per class, summing to 2 mu L Nc H(Nc - 1)
    N = 10000
    rec[0] = (N, 0.0, 0.5, mu, 0.0, 0.0)
    got = O.mse_lambda(rec[0], L, lib=lib).astype(np.float64)
    Nc = 2 * N
    total = 2 * mu * L * Nc * np.sum(1.0 / np.arange(1, Nc))
    assert abs(got.sum() / total - 1) < 1e-5, ("neutral total", got.sum(), total)


def check_mse_counts(lib):
    """Class counts are Poisson draws of lambda: the total within 5 sigma
    of the sum of lambda, every count >= 0, the low classes' mean counts
    within 5 sigma."""
    N, mu, L = 10000, 1e-9, 2e8
    for s in (0.0, np.float32(-10.0 / 20000)):
        rec = np.zeros(1, POP_GEN_DTYPE)
        rec[0] = (N, 0.0, 0.5, mu, 0.0, s)
        lam = O.mse_lambda(rec[0], L, lib=lib).astype(np.float64)
        c = O.mse_counts(rec[0], 0, L, 42, 4242, lib=lib).astype(np.float64)
        assert len(c) == len(lam) and c.min() >= 0
        assert abs(c.sum() - lam.sum()) < 5 * np.sqrt(lam.sum()), ("mse counts", s, c.sum(), lam.sum())
        hi = lam > 20   # normal-approximation classes: mean and spread
        z = (c[hi] - lam[hi]) / np.sqrt(lam[hi])
        assert abs(z.mean()) < 5 / np.sqrt(hi.sum()) and 0.8 < z.var() < 1.2, ("mse counts z", z.mean(), z.var())


def check_new_count(lib):
    """New-mutation counts over 4,000 generations: Poisson(mu Nc L) mean and
    variance within 5 sigma, in the Poisson (lambda 3) and normal (lambda
    200) regimes."""
    for Nc, mu, L in ((20000, 1e-9, 1.5e8), (20000, 1e-9, 1e10)):
        lam = np.float64(np.float32(np.float32(mu) * np.float32(Nc)) * np.float32(L))
        c = np.array([O.new_count(Nc, mu, L, 99, 100, g, 0, lib=lib) for g in range(1, 4001)], np.float64)
        mu_, var, se_mu, se_var = _moments(c)
        extra = 1.0 / 12 if lam > 6 else 0.0
        assert abs(mu_ - lam) < 5 * se_mu and abs(var - lam - extra) < 5 * se_var, ("new count", lam, mu_, var)


def check_pop_hist(lib):
    """Population histogram against numpy: bin = Nc x (float32 product)
    rounded half away from zero, bin Nc folded into 0, and
    (int)num_sites - M added to bin 0."""
    rng = np.random.default_rng(5)
    Nc = 1024
    k = np.arange(Nc)
    freq = np.concatenate([(k + np.float32(0.5)) / np.float32(Nc),   # exact ties
                           rng.uniform(0, 1, 3000), [0.0, 1.0, 1.0, 1 - 1e-4]]).astype(np.float32)
    num_sites = 5000.75
    got = O.population_frequency_histogram(freq, Nc, num_sites, lib=lib)
    v = (np.float32(Nc) * freq).astype(np.float64)
    idx = (np.sign(v) * np.floor(np.abs(v) + 0.5)).astype(np.int64)
    idx[idx == Nc] = 0
    ref = np.bincount(idx, minlength=Nc).astype(np.float64)
    ref[0] += int(num_sites) - len(freq)
    assert np.array_equal(got.astype(np.float64), ref), ("pop hist", np.nonzero(got != ref)[0][:5])


def check_keep_rule(lib):
    """One compaction against a brute-force loop: keep a slot iff it is not
    lost in every present population, not fixed in every present
    population, or preserved (threadID < 0); stable order; with preserve,
    every kept threadID becomes -|threadID|."""
    rng = np.random.default_rng(3)
    vals = np.float32([0.0, 1.0, 0.25, 0.5, 1e-5, 1.0 - 2 ** -20])
    for trial in range(120):
        P = int(rng.integers(1, 4))
        M = int(rng.integers(0, 40))
        x = vals[rng.integers(0, len(vals), (M, P))]
        x[rng.random((M, P)) < 0.4] = 0.0
        x[rng.random((M, P)) < 0.2] = 1.0
        ids = np.zeros(M, MUTID_DTYPE)
        ids["origin_generation"] = rng.integers(0, 100, M)
        ids["origin_population"] = rng.integers(0, P, M)
        ids["origin_threadID"] = rng.integers(1, 1000, M) * rng.choice([-1, 1, 1], M)
        present = rng.integers(0, 2, P)
        preserve = bool(rng.integers(0, 2))
        got_x, got_ids = O.compact_step(x, ids, present, preserve, lib=lib)
        keep = []
        for i in range(M):
            vp = [x[i, p] for p in range(P) if present[p]]
            zero = all(v <= 0 for v in vp)
            one = all(v >= 1 for v in vp)
            keep.append((not (zero or one)) or ids["origin_threadID"][i] < 0)
        keep = np.array(keep, bool)
        ref_ids = ids[keep].copy()
        if preserve:
            ref_ids["origin_threadID"] = -np.abs(ref_ids["origin_threadID"])
        assert np.array_equal(got_x, x[keep]) and np.array_equal(got_ids, ref_ids), ("keep rule", trial, P, present)


CHECKS = {
    "eq13": check_eq13,
    "migration": check_migration,
    "drift_moments": check_drift_moments,
    "mse_lambda": check_mse_lambda,
    "mse_counts": check_mse_counts,
    "new_count": check_new_count,
    "pop_hist": check_pop_hist,
    "keep_rule": check_keep_rule,
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_oracle_pin(name, oracle):
    CHECKS[name](oracle.load())


# ORC_PLANT value -> (the slip, the check that must catch it)
PLANTS = {
    1: ("Eq. 13 constants c1 and c2 swapped", "eq13"),
    2: ("F-term sign of c1: (F+h)+hF", "eq13"),
    3: ("MSE j = 1 - x instead of ((Nc - id) + 1)/Nc", "mse_lambda"),
    4: ("MSE upper integral from (id+1)/Nc", "mse_lambda"),
    5: ("MSE integrand exponent with +s", "mse_lambda"),
    6: ("histogram bins rounded half to even", "pop_hist"),
    7: ("histogram bin Nc not folded into 0", "pop_hist"),
    8: ("keep rule ignores the present mask", "keep_rule"),
    9: ("preserve does not negate threadIDs", "keep_rule"),
    10: ("draw variance x' mean instead of (1 - x') mean", "eq13"),
    11: ("selection not clamped at -1", "eq13"),
    12: ("mse() F-term (1 + F) instead of (1 - F)", "mse_lambda"),
    13: ("num_sites rounded instead of truncated", "pop_hist"),
    15: ("normal draw scaled by the variance, not its root", "drift_moments"),
}


def _build_plant(k, outdir):
    out = os.path.join(outdir, "liboracle_plant_%d.so" % k)
    cmd = ["g++", "-O2", "-ffp-contract=off", "-fno-fast-math", "-std=c++17", "-fPIC", "-shared", "-pthread",
           "-DORC_PLANT=%d" % k, "-o", out, os.path.join(REPO, "oracle", "gofish_oracle.cpp")]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("plants"))
    with ThreadPoolExecutor(max_workers=6) as ex:
        paths = dict(zip(PLANTS, ex.map(lambda k: _build_plant(k, d), PLANTS)))
    return {k: O.bind(p) for k, p in paths.items()}


@pytest.mark.parametrize("k", sorted(PLANTS))
def test_planted_mistakes_are_caught(k, planted):
    what, check = PLANTS[k]
    with pytest.raises(AssertionError):
        CHECKS[check](planted[k])
