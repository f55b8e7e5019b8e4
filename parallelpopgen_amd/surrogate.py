"""Learned surrogate of the batched simulator's sample SFS.

Parameter space (per draw): selection gamma ~ U(-50, 10) (2*N0*s/(1+F) at
the ancestral size), dominance h in {0, 0.5, 1}, inbreeding F ~ U(0, 0.2),
one instantaneous size change by a factor G ~ logU(1, 100) that lasted T ~
logU(10, 2000) generations before sampling.  Output: the expected SFS of an
n = 100 sample, as normalised bins 1..99 plus log(total / theta).

Simulation scale.  Draws are simulated at an ancestral size of
SIM_N0 = 1,000 diploids standing for REF_N0 = 10,000: generations are
divided and the per-generation mutation rate multiplied by
REF_N0 / SIM_N0 (the reference's population-scaling rule,
examples/usage_and_examples.dox:86-102, under which the normalised SFS is
invariant); gamma is kept.  theta = 4 * SIM_N0 * mu * L is fixed at
THETA per simulation.

Network: include/gofish_capi.h gf_sg_* (ensemble of MLPs trained on the
GPU by the library's own kernels).  Inference falls back to simulation for
inputs outside the training box or whose ensemble disagreement exceeds a
threshold fixed on held-out data.
"""
import ctypes as C
import math
import time

import numpy as np

from . import capi

REF_N0 = 10000
SIM_N0 = 1000
SCALE = REF_N0 // SIM_N0
MU = 1e-9 * SCALE
THETA = 200.0
N_SAMPLE = 100
NBINS = 99
GAMMA_RANGE = (-50.0, 10.0)
H_VALUES = (0.0, 0.5, 1.0)
F_RANGE = (0.0, 0.2)
GROWTH_RANGE = (1.0, 100.0)
T_RANGE = (10.0, 2000.0)
GAMMA_FLOOR = 1e-2

def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Network:
    """Ensemble of MLPs living on the GPU (C ABI gf_sg_*)."""

    def __init__(self, ensemble=4, seed=1, device=-1):
        self.lib = capi.load()
        self.E = ensemble
        h = C.c_void_p()
        capi.check(self.lib.gf_sg_create(ensemble, seed, device, C.byref(h)))
        self.h = h
        self.n_params = self.lib.gf_sg_num_params(h)

    def close(self):
        if self.h:
            self.lib.gf_sg_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.empty((self.E, self.n_params), np.float32)
        capi.check(self.lib.gf_sg_get_params(self.h, _p(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32).reshape(self.E, self.n_params)
        capi.check(self.lib.gf_sg_set_params(self.h, _p(p)))

    def gradients(self, x, y, t, w_tot=0.1):
        x, y, t = (np.ascontiguousarray(a, np.float32) for a in (x, y, t))
        g = np.empty((self.E, self.n_params), np.float32)
        loss = np.empty(self.E, np.float32)
        capi.check(self.lib.gf_sg_gradients(self.h, _p(x), _p(y), _p(t), len(x), w_tot, _p(g), _p(loss)))
        return loss, g

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty((self.E, len(x), 128), np.float32)
        capi.check(self.lib.gf_sg_forward(self.h, _p(x), len(x), _p(out)))
        return out

    def train(self, x, y, t, epochs=10, batch=512, lr=1e-3, w_tot=0.1, seed=0, shuffle=True):
        """`epochs` passes of Adam over mini-batches of `batch` rows (a
        multiple of 64); rows are visited in a fresh random order each
        epoch, or in storage order when shuffle is False.  Returns the mean
        training loss of every epoch."""
        x, y, t = (np.ascontiguousarray(a, np.float32) for a in (x, y, t))
        order = None
        if shuffle:
            rng = np.random.default_rng(seed)
            order = np.ascontiguousarray(np.stack([rng.permutation(len(x)) for _ in range(epochs)]), np.int32)
        losses = np.empty(epochs, np.float32)
        capi.check(self.lib.gf_sg_train(self.h, _p(x), _p(y), _p(t), len(x), epochs, batch, lr, w_tot,
                                        _p(order), _p(losses)))
        return losses

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        n = len(x)
        sfs = np.empty((n, NBINS), np.float32)
        tot = np.empty(n, np.float32)
        dis = np.empty(n, np.float32)
        capi.check(self.lib.gf_sg_predict(self.h, _p(x), n, _p(sfs), _p(tot), _p(dis)))
        return sfs, tot, dis

    def predict_device(self, x, sfs, tot=None, dis=None):
        """predict on device memory: x [n][6] and the outputs sfs [n][99],
        tot [n], dis [n] are float32 capi.DeviceArray (tot and dis
        optional)"""
        n = x.shape[0]
        for a, shape in ((x, (n, 6)), (sfs, (n, NBINS)), (tot, (n,)), (dis, (n,))):
            if a is not None and (a.shape != shape or a.dtype != np.float32):
                raise ValueError("predict_device: float32 device arrays of shapes [n][6], [n][99], [n]")
        ptr = lambda a: C.c_void_p(a.ptr) if a is not None else None
        capi.check(self.lib.gf_sg_predict_device(self.h, ptr(x), n, ptr(sfs), ptr(tot), ptr(dis)))


# ---------------- parameter draws, features, simulation ----------------

def draw(n, rng):
    return dict(
        gamma=rng.uniform(*GAMMA_RANGE, n),
        h=rng.choice(H_VALUES, n),
        F=rng.uniform(*F_RANGE, n),
        growth=np.exp(rng.uniform(math.log(GROWTH_RANGE[0]), math.log(GROWTH_RANGE[1]), n)),
        T=np.exp(rng.uniform(math.log(T_RANGE[0]), math.log(T_RANGE[1]), n)),
    )


def features(p):
    g = np.maximum(np.abs(p["gamma"]), GAMMA_FLOOR)
    return np.stack([np.log(p["growth"]) / math.log(GROWTH_RANGE[1]),
                     np.log(p["T"] / T_RANGE[0]) / math.log(T_RANGE[1] / T_RANGE[0]),
                     np.log10(g) / 2.0,
                     np.sign(p["gamma"]),
                     p["h"],
                     p["F"] / F_RANGE[1]], axis=1).astype(np.float32)


def in_hull(x):
    """training box of the features (with the exact bounds of draw())"""
    lo = np.array([0, 0, math.log10(GAMMA_FLOOR) / 2, -1, 0, 0], np.float32) - 1e-6
    hi = np.array([1, 1, math.log10(max(abs(GAMMA_RANGE[0]), GAMMA_RANGE[1])) / 2, 1, 1, 1], np.float32) + 1e-6
    return np.all((x >= lo) & (x <= hi), axis=1)


def batch_sims(p, seeds):
    n = len(p["gamma"])
    sims = np.zeros(n, capi.BATCH_SIM_DTYPE)
    Nc0 = (2.0 * SIM_N0) / (1.0 + p["F"])   # Sim_Model effective-selection conversion
    sims["N"] = SIM_N0
    sims["N1"] = np.maximum(1, np.round(SIM_N0 * p["growth"])).astype(np.int32)
    sims["F"] = p["F"]
    sims["h"] = p["h"]
    sims["mu"] = MU
    sims["s"] = (p["gamma"] / Nc0).astype(np.float32)
    sims["num_sites"] = THETA / (4.0 * SIM_N0 * MU)
    sims["seed1"] = seeds.astype(np.int64) & 0x7FFFFFFF
    sims["seed2"] = 0x5eed
    sims["num_generations"] = np.maximum(1, np.round(p["T"] / SCALE)).astype(np.int32)
    return sims


def simulate(p, seeds, chunk=8192, compact_interval=32):
    """expected n=100 sample SFS (bins 1..99) of every draw, by the batched simulator"""
    n = len(seeds)
    out = np.empty((n, NBINS), np.float64)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        sub = {k: v[a:b] for k, v in p.items()}
        r = capi.batch_run(batch_sims(sub, seeds[a:b]), compact_interval=compact_interval, sfs_n=N_SAMPLE)
        out[a:b] = r["sfs"][:, 1:]
    return out


def targets(sfs):
    tot = np.maximum(sfs.sum(1), 1e-6)
    y = (sfs / tot[:, None]).astype(np.float32)
    t = np.log(tot / THETA).astype(np.float32)
    return y, t


def expected_counts(sfs_norm, log_tot):
    return sfs_norm * (np.exp(log_tot) * THETA)[:, None]


def poisson_deviance(obs, mu):
    """mean per-bin Poisson deviance of observed (expected-SFS) counts
    against means mu, per draw"""
    obs = np.maximum(obs, 0.0)
    mu = np.maximum(mu, 1e-12)
    with np.errstate(divide="ignore", invalid="ignore"):
        term = np.where(obs > 0, obs * np.log(obs / mu), 0.0)
    return (2.0 * (term - (obs - mu))).mean(1)


class Surrogate:
    """Network + fallback to simulation (out-of-hull inputs or ensemble
    disagreement above `threshold`)."""

    def __init__(self, net, threshold):
        self.net = net
        self.threshold = float(threshold)

    def predict(self, p, seeds):
        x = features(p)
        sfs, tot, dis = self.net.predict(x)
        counts = expected_counts(sfs, tot)
        fallback = (~in_hull(x)) | (dis > self.threshold)
        if fallback.any():
            idx = np.nonzero(fallback)[0]
            sub = {k: v[idx] for k, v in p.items()}
            counts[idx] = simulate(sub, seeds[idx])
        return counts, fallback


# ---------------- data, training, evaluation ----------------

def make_dataset(n, seed, chunk=8192):
    """n parameter draws and their simulated SFS (one batched simulation
    each); returns features, targets and the simulation wall time"""
    rng = np.random.default_rng(seed)
    p = draw(n, rng)
    seeds = rng.integers(1, 2 ** 31 - 1, n)
    t0 = time.perf_counter()
    sfs = simulate(p, seeds, chunk=chunk)
    dt = time.perf_counter() - t0
    y, t = targets(sfs)
    return dict(params=p, seeds=seeds, x=features(p), y=y, t=t, sfs=sfs, sim_seconds=dt)


DEFAULT_SCHEDULE = ((1e-3, 12), (3e-4, 8), (1e-4, 5))


def fit(x, y, t, ensemble=4, seed=1, schedule=DEFAULT_SCHEDULE, batch=512, w_tot=1.0):
    """trains a fresh ensemble with Adam through a piecewise-constant
    learning-rate schedule of (lr, epochs) pairs; returns (net, losses)"""
    net = Network(ensemble=ensemble, seed=seed)
    n = len(x) // batch * batch
    losses = []
    for i, (lr, ep) in enumerate(schedule):
        losses += list(net.train(x[:n], y[:n], t[:n], epochs=ep, batch=batch, lr=lr, w_tot=w_tot, seed=seed + i))
    return net, np.array(losses)


def threshold_from(net, x_val, quantile=0.99):
    """disagreement threshold: the given quantile of the ensemble
    disagreement on held-out inputs"""
    _, _, dis = net.predict(x_val)
    return float(np.quantile(dis, quantile))


def evaluate(sur, n, seed):
    """fresh draws, each simulated twice with independent seeds (replicates
    A and B).  Returns the mean per-bin Poisson deviance of A against B (the
    replicate noise floor), of A against the surrogate (with its fallback),
    the fallback fraction, and the simulator and surrogate rates in SFS/s."""
    rng = np.random.default_rng(seed)
    p = draw(n, rng)
    sa = rng.integers(1, 2 ** 31 - 1, n)
    sb = rng.integers(1, 2 ** 31 - 1, n)
    sc = rng.integers(1, 2 ** 31 - 1, n)   # fallback simulations: a third, independent replicate
    t0 = time.perf_counter()
    A = simulate(p, sa)
    t_sim = time.perf_counter() - t0
    B = simulate(p, sb)
    t0 = time.perf_counter()
    pred, fb = sur.predict(p, sc)
    t_sur = time.perf_counter() - t0
    x = features(p)
    sfs, tot, _ = sur.net.predict(x)
    net_only = expected_counts(sfs, tot)
    return dict(
        deviance_replicates=float(poisson_deviance(A, B).mean()),
        deviance_surrogate=float(poisson_deviance(A, pred).mean()),
        deviance_network_only=float(poisson_deviance(A, net_only).mean()),
        fallback_fraction=float(fb.mean()),
        simulator_sfs_per_s=n / t_sim,
        surrogate_with_fallback_sfs_per_s=n / t_sur,
    )


def prediction_rate_device(net, n=1 << 20, repeats=3, seed=0):
    """ensemble predictions per second for n inputs already in device
    memory, outputs written to device memory"""
    rng = np.random.default_rng(seed)
    x = capi.DeviceArray.from_host(features(draw(n, rng)))
    sfs, tot, dis = capi.DeviceArray((n, NBINS)), capi.DeviceArray(n), capi.DeviceArray(n)
    net.predict_device(x, sfs, tot, dis)   # the batch buffers are allocated here
    best = None
    for _ in range(repeats):
        t0 = time.perf_counter()
        net.predict_device(x, sfs, tot, dis)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    for a in (x, sfs, tot, dis):
        a.free()
    return n / best


def prediction_rate(net, n=1 << 20, repeats=3, seed=0):
    """ensemble predictions per second for n inputs already on the host
    (includes the host<->device copies of inputs and outputs)"""
    rng = np.random.default_rng(seed)
    x = features(draw(n, rng))
    net.predict(x[:8192])
    best = None
    for _ in range(repeats):
        t0 = time.perf_counter()
        net.predict(x)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return n / best
