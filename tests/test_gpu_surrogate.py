"""Surrogate network on the GPU (C ABI gf_sg_*) against the float64 CPU
oracle with the same bfloat16 rounding points (oracle/surrogate_oracle.py).

Tolerances: the GPU accumulates in FP32 in MFMA order and evaluates tanh /
exp in FP32, so a bfloat16 rounding of an activation can land on the other
side of a tie; logits agree to 1e-2 of their RMS, per-layer gradients to
3e-2 in relative L2 norm, losses to 1e-3.  A wrong accumulator layout or a
transposed operand gives errors of order 1."""
import numpy as np
import pytest

from oracle import surrogate_oracle as SO
from parallelpopgen_amd import surrogate as SG

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    y = rng.dirichlet(np.ones(SO.NBINS) * 0.5, n).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    return x, y, t


def test_forward_matches_oracle():
    net = SG.Network(ensemble=2, seed=11)
    x, _, _ = inputs(128, 1)
    logits = net.forward(x)
    params = net.get_params()
    for e in range(2):
        _, _, L = SO.forward(params[e].astype(np.float64), x.astype(np.float64))
        g = logits[e]
        assert np.all(g[:, SO.OUT:] == 0.0)
        err = np.abs(g[:, :SO.OUT] - L[:, :SO.OUT]).max()
        rms = np.sqrt(np.mean(L[:, :SO.OUT] ** 2))
        assert err <= 1e-2 * rms, (e, err, rms)
    net.close()


def test_ensemble_members_differ_and_seed_reproduces():
    a = SG.Network(ensemble=3, seed=5).get_params()
    b = SG.Network(ensemble=3, seed=5).get_params()
    c = SG.Network(ensemble=3, seed=6).get_params()
    assert np.array_equal(a, b)
    assert not np.array_equal(a[0], a[1]) and not np.array_equal(a, c)


@pytest.mark.parametrize("B", [64, 192, 256])
def test_gradients_match_oracle(B):
    net = SG.Network(ensemble=2, seed=12)
    x, y, t = inputs(B, 2)
    loss, g = net.gradients(x, y, t, w_tot=0.2)
    params = net.get_params()
    offs, _ = SO.layout()
    for e in range(2):
        lo, go = SO.loss_and_grads(params[e].astype(np.float64), x.astype(np.float64), y.astype(np.float64),
                                   t.astype(np.float64), 0.2)
        assert abs(loss[e] - lo) <= 1e-3 * abs(lo), (e, loss[e], lo)
        for l, (w, b) in enumerate(offs):
            nw = SO.INS[l] * SO.OUTS[l]
            rw = rel(g[e, w:w + nw], go[w:w + nw])
            rb = rel(g[e, b:b + SO.OUTS[l]], go[b:b + SO.OUTS[l]])
            assert rw < 3e-2 and rb < 3e-2, (e, l, rw, rb)
    net.close()


def test_adam_step_matches_oracle():
    net = SG.Network(ensemble=2, seed=13)
    x, y, t = inputs(64, 3)
    w0 = net.get_params().astype(np.float64)
    _, g = net.gradients(x, y, t, w_tot=0.1)
    lr = 1e-3
    net.train(x, y, t, epochs=1, batch=64, lr=lr, w_tot=0.1, shuffle=False)
    w1 = net.get_params().astype(np.float64)
    for e in range(2):
        exp, _, _ = SO.adam(w0[e], g[e].astype(np.float64), 0.0, 0.0, 1, lr)
        assert np.abs((w1[e] - w0[e]) - (exp - w0[e])).max() <= 1e-3 * lr + 1e-7
    # zero padding stays zero
    offs, _ = SO.layout()
    W0 = w1[:, offs[0][0]:offs[0][0] + 64 * 512].reshape(2, 512, 64)
    assert np.all(W0[:, :, 6:] == 0)
    net.close()


def test_training_reduces_loss_on_a_smooth_map():
    rng = np.random.default_rng(4)
    n = 4096
    x = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    A = rng.normal(0, 1.0, (6, SO.NBINS))
    z = x @ A
    y = np.exp(z - z.max(1, keepdims=True))
    y = (y / y.sum(1, keepdims=True)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2]).astype(np.float32)
    ent = float(-(y * np.log(np.maximum(y, 1e-30))).sum(1).mean())
    net = SG.Network(ensemble=4, seed=14)
    losses = net.train(x, y, t, epochs=40, batch=256, lr=1e-3, w_tot=1.0)
    assert np.all(np.isfinite(losses))
    excess0, excess1 = losses[0] - ent, losses[-1] - ent
    assert excess1 < 0.1 * excess0, (losses[0], losses[-1], ent)
    net.close()


def test_predict_is_the_ensemble_mean_of_forward():
    net = SG.Network(ensemble=3, seed=15)
    x, _, _ = inputs(192, 5)
    L = net.forward(x).astype(np.float64)
    sfs, tot, dis = net.predict(x[:150])        # a ragged count
    Lb = L[:, :150, :SO.NBINS]
    p = np.exp(Lb - Lb.max(2, keepdims=True))
    p /= p.sum(2, keepdims=True)
    mean = p.mean(0)
    assert np.allclose(sfs, mean, rtol=1e-4, atol=1e-7)
    assert np.allclose(tot, L[:, :150, SO.NBINS].mean(0), rtol=1e-5, atol=1e-5)
    sd = np.sqrt(np.maximum((p * p).mean(0) - mean ** 2, 0))
    ratio = np.where(mean >= 1e-3, sd / np.maximum(mean, 1e-30), 0).max(1)
    assert np.allclose(dis, ratio, rtol=2e-3, atol=2e-4)
    net.close()


def test_large_batch_forward_kernel_gives_the_same_bits():
    """Batches of 4,096 rows and more run the forward products through the
    128 x 128-tile kernel (double-buffered LDS, XCD-aware tile order, no
    pre-activation store); each output accumulates its k-steps in the same
    order as the 64 x 64-tile kernel that smaller batches use, so the logits
    of a 4,096-row batch equal those of the same rows sent 64 at a time, bit
    for bit, for every ensemble member."""
    net = SG.Network(ensemble=2, seed=21)
    x, _, _ = inputs(4096, 9)
    big = net.forward(x)
    small = np.concatenate([net.forward(x[i:i + 64]) for i in range(0, 4096, 64)], axis=1)
    assert np.array_equal(big.view(np.uint32), small.view(np.uint32))
    net.close()


def test_predict_pipelined_batches_match_one_batch():
    """Prediction over 2^17 + 100 inputs runs three batches (65,536 rows at
    most) through two output buffers, the read-back of one batch beside the
    compute of the next: the rows of each batch equal, bit for bit, a
    prediction of that batch's inputs alone."""
    net = SG.Network(ensemble=4, seed=22)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, ((1 << 17) + 100, 6)).astype(np.float32)
    sfs, tot, dis = net.predict(x)
    for lo, hi in [(0, 65536), (65536, 131072), (131072, len(x))]:
        s1, t1, d1 = net.predict(x[lo:hi])
        assert np.array_equal(sfs[lo:hi], s1) and np.array_equal(tot[lo:hi], t1) and np.array_equal(dis[lo:hi], d1), lo
    net.close()


def test_predict_device_matches_host_predict():
    """gf_sg_predict_device (inputs and outputs in device memory, batches
    of 65,536 rows, the last ragged) gives the bits of gf_sg_predict; the
    optional outputs may be left out; host pointers are refused."""
    import ctypes as C
    from parallelpopgen_amd.capi import DeviceArray
    net = SG.Network(ensemble=4, seed=23)
    rng = np.random.default_rng(4)
    n = (1 << 17) + 100
    x = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    sfs, tot, dis = net.predict(x)
    xd = DeviceArray.from_host(x)
    sd = DeviceArray.from_host(np.full((n, SG.NBINS), -1.0, np.float32))
    td, dd = DeviceArray.from_host(np.full(n, -1.0, np.float32)), DeviceArray.from_host(np.full(n, -1.0, np.float32))
    net.predict_device(xd, sd, td, dd)
    assert np.array_equal(sd.to_host(), sfs) and np.array_equal(td.to_host(), tot)
    assert np.array_equal(dd.to_host(), dis)
    x1 = DeviceArray.from_host(x[:1000])
    s1 = DeviceArray((1000, SG.NBINS))
    net.predict_device(x1, s1)
    assert np.array_equal(s1.to_host(), sfs[:1000])
    host_out = np.empty((10, SG.NBINS), np.float32)
    rc = net.lib.gf_sg_predict_device(net.h, C.c_void_p(xd.ptr), 10, host_out.ctypes.data_as(C.c_void_p), None, None)
    assert rc != 0 and "device" in net.lib.gf_last_error().decode()
    for a in (xd, sd, td, dd, x1, s1):
        a.free()
    net.close()


@pytest.mark.parametrize("n", [1, 127, 129, 4095, 4097])
def test_predict_device_ragged_counts(n):
    """Counts below, at and just past the 128-row tile and the 4,096-row
    switch to the large-tile kernel: device and host prediction agree bit
    for bit and the rows past n are not written."""
    from parallelpopgen_amd.capi import DeviceArray
    net = SG.Network(ensemble=2, seed=24)
    x = np.random.default_rng(n).uniform(-1, 1, (n, 6)).astype(np.float32)
    sfs, tot, dis = net.predict(x)
    xd = DeviceArray.from_host(x)
    guard = np.full((n + 64, SG.NBINS), 7.0, np.float32)
    sd = DeviceArray.from_host(guard)
    td = DeviceArray.from_host(np.full(n, -1.0, np.float32))
    dd = DeviceArray.from_host(np.full(n, -1.0, np.float32))
    rc = net.lib.gf_sg_predict_device(net.h, SG.C.c_void_p(xd.ptr), n, SG.C.c_void_p(sd.ptr),
                                      SG.C.c_void_p(td.ptr), SG.C.c_void_p(dd.ptr))
    assert rc == 0
    out = sd.to_host()
    assert np.array_equal(out[:n], sfs) and np.all(out[n:] == 7.0)
    assert np.array_equal(td.to_host(), tot) and np.array_equal(dd.to_host(), dis)
    for a in (xd, sd, td, dd):
        a.free()
    net.close()


def test_invalid_arguments_are_rejected():
    net = SG.Network(ensemble=1, seed=1)
    x, y, t = inputs(100, 6)      # not a multiple of 64
    with pytest.raises(RuntimeError):
        net.gradients(x, y, t)
    net.close()
