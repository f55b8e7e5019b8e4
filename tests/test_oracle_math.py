"""Pins the oracle's random-number and elementary functions.

Philox: against the reference's own Random123 (tests/golden/philox_kat.txt,
written by tests/golden/make_philox_kat.py from the vendored philox.h).
Everything else: against closed forms, long-double / scipy references and
brute-force CDF inversion, with the error bounds stated in each test.
"""
import os

import numpy as np
import pytest
import scipy.special as sp

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load_kat():
    rows = []
    with open(os.path.join(GOLDEN, "philox_kat.txt")) as f:
        for line in f:
            if line.startswith("#"):
                continue
            t = line.split()
            rows.append([int(v) for v in t[:6]] + [int(v, 16) for v in t[6:]])
    return np.array(rows, dtype=np.uint64)


def test_philox_matches_reference_random123(oracle):
    kat = load_kat()
    assert len(kat) > 250
    for r in kat:
        out = oracle.philox(int(r[0]), int(r[1]), [int(v) for v in r[2:6]])
        assert list(out) == [int(v) for v in r[6:10]], r


def test_philox_published_kat(oracle):
    # Random123 kat_vectors: philox4x32 R=10, zero key/counter
    assert list(oracle.philox(0, 0, [0, 0, 0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]


def test_u01_closed_form(oracle):
    rng = np.random.default_rng(1)
    u = np.concatenate([rng.integers(0, 2**32, 200000, dtype=np.uint64),
                        [0, 1, 2**31, 2**32 - 1, 2**32 - 128, 2**32 - 129, 2**24, 2**24 + 1]]).astype(np.uint32)
    f = oracle.u01(u).astype(np.float64)
    exact = (u.astype(np.float64) + 0.5) / 2**32
    # float(u) rounds u to 24 bits, then exact scaling and one rounded add:
    # within 2^-24 relative of (u + 0.5) / 2^32
    assert np.all(np.abs(f - exact) <= exact * 2.0**-23)
    assert np.all((f > 0) & (f <= 1))
    assert oracle.u01(np.array([0], np.uint32))[0] == np.float32(2.0**-33)
    assert oracle.u01(np.array([2**32 - 1], np.uint32))[0] == np.float32(1.0)
    assert np.all(np.diff(oracle.u01(np.sort(u))) >= 0)


def ulp_err(got, x, ref):
    got = got.astype(np.longdouble)
    spacing = np.spacing(np.abs(ref.astype(np.float32))).astype(np.longdouble)
    spacing = np.maximum(spacing, np.longdouble(2.0**-149))
    return np.abs(got - ref) / spacing


def test_exp_error_bound(oracle):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-103.9, 88.7, 400000), rng.uniform(-1, 1, 100000),
                        np.linspace(-30, 5, 20001)]).astype(np.float32)
    got = oracle.exp_f(x)
    ref = np.exp(x.astype(np.longdouble))
    err = ulp_err(got, x, ref)
    assert err.max() <= 2.0, err.max()
    assert oracle.exp_f(np.array([0.0], np.float32))[0] == 1.0
    assert oracle.exp_f(np.array([-200.0], np.float32))[0] == 0.0
    assert np.isinf(oracle.exp_f(np.array([89.0], np.float32))[0])


def test_log_error_bound(oracle):
    rng = np.random.default_rng(3)
    x = np.concatenate([np.exp(rng.uniform(-100, 88, 400000)), rng.uniform(0.5, 2, 100000),
                        [1e-40, 1e-45, 3e38, 1.0]]).astype(np.float32)
    got = oracle.log_f(x)
    ref = np.log(x.astype(np.longdouble))
    err = ulp_err(got, x, ref)
    assert err.max() <= 2.0, err.max()
    assert oracle.log_f(np.array([1.0], np.float32))[0] == 0.0
    assert np.isneginf(oracle.log_f(np.array([0.0], np.float32))[0])


def _u01_values(bits):
    b = np.asarray(bits, np.uint64).astype(np.float32)
    return (b * np.float32(2.3283064365386963e-10) + np.float32(1.1641532182693481e-10)).astype(np.float32)


def test_expneg_table_error_bound(oracle):
    """the oracle's exp(-m) table (shared with the library) against long
    double exp: within 1.5 ulp on random m in [0, 6], at the segment knots
    and just below them, and at tiny m; exp(-0) = 1 exactly"""
    rng = np.random.default_rng(12)
    knots = np.arange(97) / 16.0
    m = np.concatenate([rng.uniform(0, 6, 300000), knots, np.nextafter(knots[1:], 0),
                        2.0 ** -np.arange(1, 60)]).astype(np.float32)
    got = oracle.expneg(m).astype(np.longdouble)
    want = np.exp(-m.astype(np.longdouble))
    ulp = np.spacing(want.astype(np.float32)).astype(np.longdouble)
    err = np.abs(got - want) / ulp
    assert err.max() <= 1.5, (float(err.max()), float(m[np.argmax(err)]))
    assert oracle.expneg(np.zeros(1, np.float32))[0] == 1.0


def test_normcdfinv_error_bound(oracle):
    # Domain: every value u01 can return, [2^-33, 1].  The segment table is
    # fitted to float accuracy (1.5 ulp where |z| >= 0.1, 4.8e-7 absolute
    # everywhere); we allow 1e-6 relative + 2e-6 absolute overall and 2 ulp
    # where |z| >= 0.1.
    rng = np.random.default_rng(4)
    bits = np.concatenate([rng.integers(0, 2 ** 32, 400000, dtype=np.uint64),
                           np.arange(0, 100000, dtype=np.uint64),                 # smallest u (deepest tail)
                           2 ** 32 - 1 - np.arange(0, 100000, dtype=np.uint64),   # largest u
                           np.exp(rng.uniform(0, np.log(2.0 ** 32), 100000)).astype(np.uint64)])
    p = _u01_values(bits)
    got = oracle.normcdfinv(p).astype(np.float64)
    ref = sp.ndtri(np.minimum(p.astype(np.float64), 1 - 2.0 ** -24))
    assert np.all(np.abs(got - ref) <= 1e-6 * np.abs(ref) + 2e-6)
    big = np.abs(ref) >= 0.1
    ulp = np.spacing(np.abs(ref[big]).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(got[big] - ref[big]) <= 2 * ulp)
    # every binade of s = min(p, 1 - p) is sampled, the deepest tail included
    s = np.minimum(p, 1 - p)
    assert np.sum(s < 2.0 ** -30) > 1000 and np.sum(s > 0.25) > 1000
    # symmetry where 1-p is exact (p = 1 is clamped, below)
    q = p[(p >= 0.5) & (p < 1.0)]
    assert np.array_equal(oracle.normcdfinv(q), -oracle.normcdfinv((1 - q).astype(np.float32)))
    # u == 1 (u01 of u >= 2^32-128) is clamped to 1 - 2^-24: finite
    z1 = oracle.normcdfinv(np.array([1.0], np.float32))[0]
    assert np.isfinite(z1) and abs(z1 - sp.ndtri(1 - 2.0 ** -24)) < 1e-5
    assert oracle.normcdfinv(np.array([0.5], np.float32))[0] == 0.0


def test_normcdfinv_outside_its_domain(oracle):
    # inputs below u01's smallest value 2^-33 (the table starts there), 0,
    # negatives and NaN are taken as 2^-33; above 1 - 2^-24 as 1 - 2^-24
    lo = oracle.normcdfinv(np.array([2.0 ** -33], np.float32))[0]
    x = np.array([1.03e-10, 1e-20, 1e-45, 0.0, -0.0, -1.0, -np.inf, np.nan], np.float32)
    z = oracle.normcdfinv(x)
    assert np.all(z == lo) and abs(lo - sp.ndtri(2.0 ** -33)) < 1e-5
    hi = oracle.normcdfinv(np.array([1 - 2.0 ** -24], np.float32))[0]
    assert np.all(oracle.normcdfinv(np.array([1.0, 1.5, np.inf], np.float32)) == hi)


def test_normcdfinv_monotone(oracle):
    # non-decreasing to within one ulp of the result over consecutive u01
    # values (consecutive inputs move the result by a fraction of an ulp in
    # the tails, so a segment knot may step back by one ulp): at the deep
    # tail, across segment knots around s ~ 1e-3 and 1 - s, and in the
    # centre
    starts = [0, 4190000, 2 ** 31 - 500000, 2 ** 32 - 4200000 - 1000000, 2 ** 32 - 1000000]
    for b0 in starts:
        z = oracle.normcdfinv(_u01_values(np.arange(b0, b0 + 1000000, dtype=np.uint64)))
        d = np.diff(z.astype(np.float64))
        ulp = np.spacing(np.abs(z[:-1])).astype(np.float64)
        assert np.all(d >= -ulp), b0
        assert np.sum(d < 0) <= 8, b0


def test_round_away(oracle):
    x = np.array([0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 0.49999997, 2.0**23 + 1, -3.7, 3.2, 0.0, -0.0], np.float32)
    got = oracle.round_away(x)
    ref = (np.sign(x.astype(np.float64)) * np.floor(np.abs(x.astype(np.float64)) + 0.5)).astype(np.float32)
    assert np.array_equal(got, ref)


EXITS = {11: 1, 14: 2, 17: 3, 20: 4, 22: 5, 24: 6, 26: 7, 28: 8, 30: 9, 32: 10, 34: 11, 36: 12, 38: 13, 40: 14,
         41: 15, 43: 16, 45: 17, 46: 18, 48: 19, 50: 20, 51: 21, 53: 22, 55: 23, 56: 24, 58: 25, 59: 26, 61: 27,
         62: 28, 64: 29, 65: 30, 67: 31, 68: 32}


def brute_inversion(r, mean, pmf):
    """smallest k with CDF(k) >= r, truncated by the early-exit table and 70;
    also returns the distance of r to the nearest CDF boundary."""
    cdf = np.cumsum(pmf, axis=1)
    out = np.empty(len(r), np.int64)
    dist = np.empty(len(r))
    for i, (ri, mi) in enumerate(zip(r, mean)):
        k = int(np.searchsorted(cdf[i], ri))
        for j in range(1, 70):
            if j in EXITS and mi <= EXITS[j] and j < k:
                k = j
                break
        out[i] = min(k, 70)
        dist[i] = np.min(np.abs(cdf[i][:71] - ri))
    return out, dist


def test_poiscdfinv_brute_force(oracle):
    rng = np.random.default_rng(5)
    n = 20000
    mean = np.concatenate([rng.uniform(0, 6, n // 2), rng.uniform(0, 40, n // 2)]).astype(np.float32)
    r = rng.uniform(0, 1, n).astype(np.float32)
    got = oracle.poiscdfinv(r, mean)
    k = np.arange(80)
    m = mean.astype(np.longdouble)[:, None]
    logpmf = -m + k * np.log(np.maximum(m, 1e-300)) - np.array([np.sum(np.log(np.arange(1, j + 1, dtype=np.longdouble))) for j in k])
    pmf = np.exp(logpmf)
    pmf[mean == 0] = 0
    pmf[mean == 0, 0] = 1
    ref, dist = brute_inversion(r.astype(np.longdouble), mean, pmf)
    bad = (got != ref) & (dist > 1e-5)
    assert not bad.any(), list(zip(mean[bad][:5], r[bad][:5], got[bad][:5], ref[bad][:5]))
    assert (got == ref).mean() > 0.999


def test_binomcdfinv_brute_force(oracle):
    rng = np.random.default_rng(6)
    n = 20000
    N = rng.integers(2, 100, n).astype(np.float32)
    x = rng.uniform(0, 0.3, n).astype(np.float32)
    mean = (x * N).astype(np.float32)
    r = rng.uniform(0, 1, n).astype(np.float32)
    got = oracle.binomcdfinv(r, mean, x, N)
    k = np.arange(80)
    pmf = np.zeros((n, 80), np.longdouble)
    for i in range(n):
        Ni, xi = int(N[i]), np.longdouble(x[i])
        kk = k[k <= Ni]
        pmf[i, kk] = sp.comb(Ni, kk, exact=False) * xi ** kk * (1 - xi) ** (Ni - kk)
    ref, dist = brute_inversion(r.astype(np.longdouble), mean, pmf)
    bad = (got != ref) & (dist > 2e-5)
    assert not bad.any(), list(zip(N[bad][:5], x[bad][:5], r[bad][:5], got[bad][:5], ref[bad][:5]))


@pytest.mark.parametrize("N,p", [(20000, 0.3), (20000, 1e-4), (20000, 0.9999), (50, 0.05), (50, 0.97), (2e5, 0.5)])
def test_binom_draw_moments(oracle, N, p):
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2**32, 400000, dtype=np.uint64).astype(np.uint32)
    mean = np.float32(p * N)
    var = np.float32((1 - np.float32(p)) * mean)
    j = oracle.binom_draw(bits, mean, var, np.float32(N)).astype(np.float64)
    se = np.sqrt(N * p * (1 - p) / len(j))
    assert abs(j.mean() - N * p) < 5 * se + 1e-3
    assert abs(j.var() / (N * p * (1 - p)) - 1) < 0.03
    assert j.min() >= 0 and j.max() <= N


def test_pois_draw_moments(oracle):
    rng = np.random.default_rng(8)
    for lam in [0.5, 3.0, 6.0, 50.0, 1e5]:
        bits = rng.integers(0, 2**32, 300000, dtype=np.uint64).astype(np.uint32)
        k = oracle.pois_draw(bits, np.float32(lam), np.float32(1e12)).astype(np.float64)
        assert abs(k.mean() - lam) < 5 * np.sqrt(lam / len(k)) + 1e-3
        assert abs(k.var() / lam - 1) < 0.03


def test_f2i_saturates_and_truncates(oracle):
    x = np.array([np.nan, np.inf, -np.inf, 2.0 ** 31, -2.0 ** 31, 3e9, -3e9, 2.5, -2.5, 0.99, -0.99, 123456.7],
                 np.float32)
    want = np.array([0, 2 ** 31 - 1, -2 ** 31, 2 ** 31 - 1, -2 ** 31, 2 ** 31 - 1, -2 ** 31, 2, -2, 0, 0, 123456],
                    np.int64)
    assert np.array_equal(oracle.f2i(x).astype(np.int64), want)
    rng = np.random.default_rng(3)
    y = rng.uniform(-2e9, 2e9, 100000).astype(np.float32)
    assert np.array_equal(oracle.f2i(y).astype(np.int64), np.trunc(y.astype(np.float64)).astype(np.int64))


def test_joint_sfs_oracle_against_binomial_sums(oracle):
    """joint SFS = sum of products of binomial sample probabilities (float64
    brute force); monomorphic sites at [0][0]; summing over one population
    gives that population's unfolded sample spectrum."""
    from scipy.stats import binom
    rng = np.random.default_rng(7)
    M, na, nb, L = 400, 12, 7, 5000.0
    fa = rng.random(M).astype(np.float32)
    fb = rng.random(M).astype(np.float32)
    fa[:20] = 0.0
    fb[20:30] = 1.0
    fa[30:35] = 1.0
    j = oracle.joint_site_frequency_spectrum(fa, fb, na, nb, L)
    pa = binom.pmf(np.arange(na + 1)[None, :], na, fa.astype(np.float64)[:, None])
    pb = binom.pmf(np.arange(nb + 1)[None, :], nb, fb.astype(np.float64)[:, None])
    ref = pa.T @ pb
    ref[0, 0] += L - M
    assert np.allclose(j, ref, rtol=2e-5, atol=2e-3)
    assert abs(j.sum() - L) < 1e-2
    # marginal over population b = population a's unfolded spectrum, which
    # the marginal SFS folds (k = n_a and p in {0,1} into bin 0)
    marg = j.sum(1).astype(np.float64)
    m = oracle.site_frequency_spectrum(fa, na, L).astype(np.float64)
    assert np.allclose(marg[1:na], m[1:na], rtol=1e-5, atol=1e-3)


def test_oracle_builds_agree(oracle):
    # liboracle_fma.so (-mfma, used where the CPU has FMA) and the portable
    # liboracle.so evaluate the same explicit operations: same bits
    import ctypes as C
    if not oracle._cpu_has_fma():
        pytest.skip("CPU without FMA: only the portable build is used")
    libs = [C.CDLL(oracle.LIB_PATH), C.CDLL(oracle.LIB_FMA_PATH)]
    p = _u01_values(np.random.default_rng(9).integers(0, 2 ** 32, 200000, dtype=np.uint64))
    outs = []
    for lib in libs:
        z = np.empty_like(p)
        lib.orc_normcdfinv_array(p.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p), C.c_int64(p.size))
        outs.append(z)
    assert np.array_equal(outs[0], outs[1])
