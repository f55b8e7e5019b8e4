"""Float64 restatements of reference formulas shared by several tests
(paths relative to the reference tree)."""
import numpy as np


def mse_lambda_ref(N, F, h, mu, L, sel):
    """Float64 restatement of the reference's equilibrium class counts
    (go_fish_impl.cuh:20-71, :201-223): the upper integral of class id
    starts at id/Nc, j = ((Nc - id) + 1)/Nc, the integrand's exponent uses
    -max(s, -1); lambda = 2 mu L / x if s == 0."""
    Nc = int(np.float32(2 * N) / (np.float32(1) + np.float32(F)))
    F, h = np.float64(np.float32(F)), np.float64(np.float32(h))
    step = 1.0 / Nc

    def mse(x, s):
        return np.exp(2 * N * s * x * ((2 * h + (1 - 2 * h) * x) * (1 - F) + 2 * F) / (1 + F))

    ids = np.arange(Nc)
    a = 1.0 - ids * step
    f = lambda y: mse(y, -np.maximum(sel(y), -1.0))
    area = step * (f(a) + f(a - step)) / 2
    S = np.cumsum(area)
    I = S[Nc - 1 - ids]
    k = np.arange(Nc - 1)
    x = (k + 1) / Nc
    j = ((Nc - k) + 1) / Nc
    s = sel(x)
    twomuL = 2 * np.float64(np.float32(mu)) * L
    with np.errstate(over="ignore", invalid="ignore"):
        lam = np.where(s == 0, twomuL / x, twomuL * mse(x, s) * I[k] / (I[0] * x * j))
    return lam
