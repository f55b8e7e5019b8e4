"""The surrogate's CPU oracle (oracle/surrogate_oracle.py) against what the
mathematics fixes: with the bfloat16 rounding points switched off its
gradients equal central finite differences of its own loss, its loss is the
softmax cross-entropy written out term by term, and one Adam step from zero
moments moves every weight by lr * g / (|g| + eps)."""
import numpy as np
import pytest

from oracle import surrogate_oracle as SO


@pytest.fixture
def exact(monkeypatch):
    monkeypatch.setattr(SO, "bf16", lambda a: np.asarray(a, np.float64))
    return SO


def small_problem(seed=0, B=5):
    rng = np.random.default_rng(seed)
    offs, total = SO.layout()
    p = np.zeros(total)
    for l, (w, b) in enumerate(offs):
        fan_in = 6 if l == 0 else SO.INS[l]
        W = rng.uniform(-1, 1, (SO.OUTS[l], SO.INS[l])) * np.sqrt(6.0 / fan_in)
        if l == 0:
            W[:, 6:] = 0
        if l == SO.NHID:
            W[SO.OUT:, :] = 0
        p[w:w + W.size] = W.ravel()
        p[b:b + SO.OUTS[l]] = rng.normal(0, 0.1, SO.OUTS[l]) * (np.arange(SO.OUTS[l]) < (SO.OUT if l == SO.NHID else SO.OUTS[l]))
    x = rng.uniform(-1, 1, (B, 6))
    y = rng.dirichlet(np.ones(SO.NBINS), B)
    t = rng.normal(0, 1, B)
    return p, x, y, t


def test_gradient_matches_finite_differences(exact):
    p, x, y, t = small_problem()
    loss, g = exact.loss_and_grads(p, x, y, t, 0.3)
    offs, total = exact.layout()
    rng = np.random.default_rng(1)
    picks = []
    for l, (w, b) in enumerate(offs):
        nin = 6 if l == 0 else exact.INS[l]
        nout = exact.OUT if l == exact.NHID else exact.OUTS[l]
        for _ in range(6):
            picks.append(w + rng.integers(nout) * exact.INS[l] + rng.integers(nin))
        picks.append(b + rng.integers(nout))
    picks.append(offs[-1][0] + 99 * exact.INS[-1] + 3)   # the log-total head
    h = 1e-6
    for i in picks:
        q = p.copy()
        q[i] += h
        lp, _ = exact.loss_and_grads(q, x, y, t, 0.3)
        q[i] -= 2 * h
        lm, _ = exact.loss_and_grads(q, x, y, t, 0.3)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_loss_is_cross_entropy_plus_total_error(exact):
    p, x, y, t = small_problem(2, B=3)
    _, _, L = exact.forward(p, x)
    ce = []
    for b in range(3):
        z = L[b, :99]
        logp = z - np.log(np.sum(np.exp(z)))
        ce.append(-np.dot(y[b], logp) + 0.5 * (L[b, 99] - t[b]) ** 2)
    loss, _ = exact.loss_and_grads(p, x, y, t, 0.5)
    assert abs(loss - np.mean(ce)) < 1e-12


def test_padding_gets_no_gradient(exact):
    p, x, y, t = small_problem(3)
    _, g = exact.loss_and_grads(p, x, y, t, 0.1)
    offs, _ = exact.layout()
    W0 = g[offs[0][0]:offs[0][0] + 64 * 512].reshape(512, 64)
    assert np.all(W0[:, 6:] == 0)
    WL = g[offs[-1][0]:offs[-1][0] + 128 * 512].reshape(128, 512)
    assert np.all(WL[100:] == 0)


def test_first_adam_step():
    rng = np.random.default_rng(4)
    w = rng.normal(size=1000)
    g = rng.normal(size=1000) * 10.0 ** rng.uniform(-6, 1, 1000)
    w1, m, v = SO.adam(w, g, np.zeros(1000), np.zeros(1000), 1, 1e-3)
    assert np.allclose(w1 - w, -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-9, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_bf16_rounding():
    # bfloat16 keeps 7 fraction bits: the spacing just above 1 is 2^-7
    v = np.array([1.0, 1 + 2 ** -7, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -12, -3.0e-3, 65504.0],
                 np.float32)
    r = SO.bf16(v)
    assert r[0] == 1.0 and r[1] == 1 + 2 ** -7
    assert r[2] == 1.0                     # tie, to the even neighbour
    assert r[3] == 1 + 2 ** -6             # tie, to the even neighbour (upwards)
    assert r[4] == 1 + 2 ** -7             # above the tie
    assert abs(r[5] - (-3.0e-3)) <= 3.0e-3 * 2 ** -8
    assert r[6] == 65536.0
