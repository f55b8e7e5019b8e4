"""Fits the polynomial coefficients of normcdfinv_f (gf_math.hpp) and of its
restatement in oracle/gofish_oracle.cpp, and reports the error of the
float32 evaluation against scipy.special.ndtri.

    python parallelpopgen_amd/csrc/fit_normcdfinv.py

Method (the change of variable of Giles, "Approximating the erfinv
function", GPU Computing Gems Jade Edition, 2011): with s = min(p, 1 - p)
and w = -log2(4 s (1 - s)), the ratio R = ndtri(p) / (p - 1/2) is a smooth
function of w on [0, 8] and of sqrt(w) beyond, so one polynomial per region
(no division, no branch for most lanes) gives it to float accuracy.  The
coefficients here are fitted by weighted least squares on Chebyshev nodes
(relative error weights), then rounded to float32.

log2 of v = 4 s (1 - s) in (0, 1]: v = m 2^e with m in [sqrt(1/2), sqrt(2)),
log2(m) = f L(f), f = m - 1, L a fitted polynomial.
"""
import numpy as np
import scipy.special as sp

W_SPLIT = 8.0                 # central region: w < 8 (s > ~9.8e-4)
W_MID = 4.0                   # central polynomial in (w - 4)
Y_LO = float(np.sqrt(W_SPLIT))
Y_HI = float(np.sqrt(31.0)) + 0.01   # w <= 31 for s >= 2^-33 (u01's smallest value)
Y_MID = 0.5 * (Y_LO + Y_HI)
DEG_LOG, DEG_C, DEG_T = 9, 9, 8
F_LO, F_HI = float(np.sqrt(0.5)) - 1.0, float(np.sqrt(2.0)) - 1.0


def ratio_of_w(w2):
    """R(w) = ndtri(s) / (s - 1/2) for w = -log2(4 s (1 - s)), s <= 1/2"""
    w2 = np.asarray(w2, np.float64)
    e = -np.expm1(-w2 * np.log(2.0))          # 1 - v = (1 - 2 s)^2
    r = np.sqrt(e)                            # 1 - 2 s
    s = 0.5 * (1.0 - r)
    with np.errstate(invalid="ignore", divide="ignore"):
        out = 2.0 * np.abs(sp.ndtri(s)) / r
    return np.where(w2 < 1e-12, np.sqrt(2.0 * np.pi), out)


def _fit(fun, lo, hi, mid, deg, n=40000):
    t = np.cos(np.pi * (np.arange(n) + 0.5) / n)
    x = 0.5 * (lo + hi) + 0.5 * (hi - lo) * t
    y = fun(x)
    c = np.polynomial.polynomial.polyfit(x - mid, y, deg, w=1.0 / np.abs(y))
    return c.astype(np.float32)


def coefficients():
    log_c = _fit(lambda f: np.log2(1.0 + f) / np.where(f == 0, 1, f) + np.where(f == 0, 1 / np.log(2.0), 0),
                 F_LO, F_HI, 0.0, DEG_LOG)
    cen_c = _fit(ratio_of_w, 0.0, W_SPLIT, W_MID, DEG_C)
    tail_c = _fit(lambda y: ratio_of_w(y * y), Y_LO, Y_HI, Y_MID, DEG_T)
    return log_c, cen_c, tail_c


def _f32(x):
    return np.asarray(x, np.float64).astype(np.float32)


def _fma(a, b, c):
    # a*b is exact in float64 for float32 operands; one rounding of the sum
    # to float64 and one to float32 (double rounding is immaterial for an
    # error estimate)
    return _f32(np.float64(a) * np.float64(b) + np.float64(c))


def _horner(c, x):
    acc = np.full_like(x, c[-1], dtype=np.float32)
    for k in range(len(c) - 2, -1, -1):
        acc = _fma(acc, x, np.float32(c[k]))
    return acc


def evaluate(p, log_c, cen_c, tail_c):
    """float32 model of normcdfinv_f (same operations and order)"""
    p = np.minimum(_f32(p), np.float32(1.0 - 2.0 ** -24))
    q = _f32(p - np.float32(0.5))
    s = np.where(q < 0, p, _f32(np.float32(1.0) - p))
    v = _f32(_f32(np.float32(4.0) * s) * _f32(np.float32(1.0) - s))
    b = v.view(np.uint32)
    e = (b >> 23).astype(np.int32) - 127
    m = ((b & 0x7FFFFF) | 0x3F800000).view(np.float32)
    big = m > np.float32(1.41421356)
    m = np.where(big, _f32(np.float32(0.5) * m), m)
    e = e + big.astype(np.int32)
    f = _f32(m - np.float32(1.0))
    lg = _f32(f * _horner(log_c, f))
    w = _f32(_f32(-e.astype(np.float32)) - lg)
    cen = w < np.float32(W_SPLIT)
    rc = _horner(cen_c, _f32(w - np.float32(W_MID)))
    y = _f32(np.sqrt(np.maximum(w, 0).astype(np.float64)))
    rt = _horner(tail_c, _f32(y - np.float32(Y_MID)))
    return _f32(q * np.where(cen, rc, rt))


def report(log_c, cen_c, tail_c):
    rng = np.random.default_rng(7)
    bits = np.concatenate([rng.integers(0, 2 ** 32, 2000000, dtype=np.uint64),
                           np.arange(0, 200000, dtype=np.uint64),
                           2 ** 32 - 1 - np.arange(0, 200000, dtype=np.uint64)])
    u = _f32(_f32(bits.astype(np.float32) * np.float32(2.3283064365386963e-10)) + np.float32(1.1641532182693481e-10))
    z = evaluate(u, log_c, cen_c, tail_c).astype(np.float64)
    ref = sp.ndtri(np.minimum(u.astype(np.float64), 1.0 - 2.0 ** -24))
    err = np.abs(z - ref)
    return float(np.max(err / (np.abs(ref) + 1e-30) * (np.abs(ref) > 1e-3))), float(np.max(err))


def c_array(name, c):
    body = ", ".join(float(x).hex().replace("0x", "0x") + "f" for x in c)
    return "static constexpr float %s[%d] = {%s};" % (name, len(c), body)


if __name__ == "__main__":
    lc, cc, tc = coefficients()
    rel, ab = report(lc, cc, tc)
    print("max relative error %.3e (|z| > 1e-3), max absolute error %.3e" % (rel, ab))
    print(c_array("kNcdfLog2", lc))
    print(c_array("kNcdfCentral", cc))
    print(c_array("kNcdfTail", tc))
    print("Y_MID = %r (float32 %s)" % (Y_MID, float(np.float32(Y_MID)).hex()))
