"""CPU ORACLE of the surrogate network (test infrastructure only).

float64 numpy restatement of the library's surrogate forward / backward /
Adam (parallelpopgen_amd/csrc/surrogate.hip) with the same BF16 rounding
points: weights, layer inputs (activations) and back-propagated
pre-activation gradients are rounded to bfloat16 (round to nearest even),
everything else is accumulated in float64.  Differences from the GPU come
only from FP32 accumulation order and FP32 tanh/exp, so gradients agree to
about 1e-3 relative.

Parameter layout per ensemble member (floats): for l = 0..4, W_l [out][in]
then b_l [out]; in = (64, 512, 512, 512, 512), out = (512, 512, 512, 512, 128);
inputs are the 6 features padded with zeros to 64, outputs 0..98 are the
SFS-bin logits, 99 the log-total head, 100..127 padding.
"""
import numpy as np

IN_PAD, HID, NHID, OUT, OUT_PAD, NBINS = 64, 512, 4, 100, 128, 99
INS = [IN_PAD] + [HID] * NHID
OUTS = [HID] * NHID + [OUT_PAD]


def layout():
    offs, o = [], 0
    for l in range(NHID + 1):
        w = o
        o += INS[l] * OUTS[l]
        b = o
        o += OUTS[l]
        offs.append((w, b))
    return offs, o


def bf16(a):
    """round float32 -> bfloat16 (nearest even), returned as float64"""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def gelu(x):
    c = np.sqrt(2 / np.pi)
    return 0.5 * x * (1 + np.tanh(c * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    c = np.sqrt(2 / np.pi)
    u = c * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * c * (1 + 3 * 0.044715 * x * x)


def unpack(params):
    offs, total = layout()
    Ws, bs = [], []
    for l, (w, b) in enumerate(offs):
        Ws.append(params[w:w + INS[l] * OUTS[l]].reshape(OUTS[l], INS[l]).astype(np.float64))
        bs.append(params[b:b + OUTS[l]].astype(np.float64))
    return Ws, bs


def forward(params, x):
    Ws, bs = unpack(params)
    B = x.shape[0]
    X = np.zeros((B, IN_PAD))
    X[:, :x.shape[1]] = x
    X = bf16(X)
    acts, Zs = [X], []
    for l in range(NHID + 1):
        Z = acts[-1] @ bf16(Ws[l]).T + bs[l]
        if l < NHID:
            Zs.append(Z.astype(np.float32).astype(np.float64))
            acts.append(bf16(gelu(Z)))
        else:
            logits = Z
    return acts, Zs, logits


def loss_and_grads(params, x, y, t, w_tot):
    """returns (loss, flat gradient) for one ensemble member"""
    Ws, bs = unpack(params)
    B = x.shape[0]
    acts, Zs, L = forward(params, x)
    Lb = L[:, :NBINS]
    mx = Lb.max(1, keepdims=True)
    ls = np.log(np.exp(Lb - mx).sum(1, keepdims=True)) + mx
    p = np.exp(Lb - ls)
    ce = -(y * (Lb - ls)).sum(1)
    e = L[:, NBINS] - t
    loss = float(np.mean(ce + w_tot * e * e))
    D = np.zeros((B, OUT_PAD))
    D[:, :NBINS] = (p - y) / B
    D[:, NBINS] = 2 * w_tot * e / B
    D = bf16(D)
    offs, total = layout()
    g = np.zeros(total)
    for l in range(NHID, -1, -1):
        w0, b0 = offs[l]
        g[w0:w0 + INS[l] * OUTS[l]] = (D.T @ acts[l]).ravel()
        g[b0:b0 + OUTS[l]] = D.sum(0)
        if l == 0:
            break
        D = bf16((D @ bf16(Ws[l])) * gelu_grad(Zs[l - 1]))
    return loss, g


def adam(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    w = w - lr * (m / c1) / (np.sqrt(v / c2) + eps)
    return w, m, v
