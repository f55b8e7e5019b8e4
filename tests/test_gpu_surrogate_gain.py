"""What the trained surrogate gains over simulation, and its fallback.

One ensemble is trained on 65,536 batched simulations by the library's own
kernels (parallelpopgen_amd/surrogate.py fit), then scored on 2,048 fresh
draws, each simulated twice with independent seeds (replicates A and B):
  * accuracy: the mean per-bin Poisson deviance of A against the surrogate
    (with its fallback) is at most 1.5 times that of A against B, the
    simulator's own replicate noise;
  * speed: ensemble predictions per second are at least 100 times the
    batched simulator's SFS per second, both measured here on the same GPU
    (the simulator on 16,384 draws in its 8,192-draw calls, the network on
    65,536 inputs, both with inputs and outputs in host memory);
  * fallback: inputs outside the training box, and inputs whose ensemble
    disagreement is above the threshold, are answered by simulation --
    exactly those rows, with exactly the simulator's numbers -- and every
    other row by the network.
The nearest reference analogue is the comparison of simulated spectra with
an independent method (examples/example_dadi/run.cu:104-113 against
dadi_example.out)."""
import time

import numpy as np
import pytest

from parallelpopgen_amd import surrogate as SG

pytestmark = pytest.mark.gpu

TRAIN = 65536
HELD_OUT = 2048


@pytest.fixture(scope="module")
def trained():
    print("surrogate: simulating %d training draws" % TRAIN, flush=True)
    data = SG.make_dataset(TRAIN + 4096, seed=31)
    x, y, t = data["x"][:TRAIN], data["y"][:TRAIN], data["t"][:TRAIN]
    print("surrogate: training", flush=True)
    net, losses = SG.fit(x, y, t)
    thr = SG.threshold_from(net, data["x"][TRAIN:])
    return net, thr, losses


def test_training_converges(trained):
    _, _, losses = trained
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.9 * losses[0]


def test_deviance_within_replicate_noise(trained):
    net, thr, _ = trained
    ev = SG.evaluate(SG.Surrogate(net, thr), HELD_OUT, seed=32)
    ratio = ev["deviance_surrogate"] / ev["deviance_replicates"]
    print("held-out deviance: replicates %.4f, surrogate %.4f (ratio %.3f), network alone %.4f, fallback %.4f"
          % (ev["deviance_replicates"], ev["deviance_surrogate"], ratio, ev["deviance_network_only"],
             ev["fallback_fraction"]), flush=True)
    assert ratio <= 1.5
    # the threshold is the 99th percentile of held-out disagreement: about
    # 1 % of in-distribution inputs fall back
    assert 0.0 < ev["fallback_fraction"] < 0.05


def test_speedup_over_batched_simulator(trained):
    net, _, _ = trained
    rng = np.random.default_rng(33)
    n_sim = 16384
    p = SG.draw(n_sim, rng)
    seeds = rng.integers(1, 2 ** 31 - 1, n_sim)
    SG.simulate({k: v[:256] for k, v in p.items()}, seeds[:256])   # warm-up
    t0 = time.perf_counter()
    SG.simulate(p, seeds)
    sim_rate = n_sim / (time.perf_counter() - t0)
    pred_rate = SG.prediction_rate(net, n=65536)
    print("simulator %.3g SFS/s, surrogate %.3g SFS/s (%.0fx)" % (sim_rate, pred_rate, pred_rate / sim_rate),
          flush=True)
    assert pred_rate >= 100 * sim_rate


def test_fallback_rows_are_simulated(trained):
    net, thr, _ = trained
    rng = np.random.default_rng(34)
    n = 512
    p = SG.draw(n, rng)
    # out of the training box: selection and growth beyond the drawn ranges
    p["gamma"][:16] = -200.0
    p["growth"][16:32] = 300.0
    seeds = rng.integers(1, 2 ** 31 - 1, n)
    x = SG.features(p)
    _, _, dis = net.predict(x)
    # a threshold at the median disagreement sends about half of the
    # in-box rows to simulation as well
    sur = SG.Surrogate(net, float(np.median(dis[32:])))
    counts, fb = sur.predict(p, seeds)
    want = (~SG.in_hull(x)) | (dis > sur.threshold)
    assert np.array_equal(fb, want)
    assert np.all(fb[:32]) and 0.3 < fb[32:].mean() < 0.7
    idx = np.nonzero(fb)[0]
    sim = SG.simulate({k: v[idx] for k, v in p.items()}, seeds[idx])
    assert np.array_equal(counts[idx], sim.astype(counts.dtype))
    sfs, tot, _ = net.predict(x)
    keep = np.nonzero(~fb)[0]
    assert np.array_equal(counts[keep], SG.expected_counts(sfs, tot)[keep])
