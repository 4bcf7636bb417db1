"""Fit of the polynomial Q(t) behind the FP32 erfc of the mixed pair kernel.

erfc(x) = t Q(t) exp(-x^2) with t = 1 / (1 + x / 2), so Q(t) = erfcx(x) / t,
a smooth function on t in [0, 1] (x from infinity down to 0; Q(0) =
1 / (2 sqrt(pi))).  The coefficients are a weighted least-squares fit for
the relative error on Chebyshev nodes, the weights raised where the error
is largest (iteratively reweighted, towards the minimax fit).

    python -m mdbas_amd.erfc_fit        # prints the coefficients, low order first

The kernel (csrc/mdg_pair.hip, mdg_erfc_q) holds the degree-9 result; the
CPU test tests/test_erfc_fit.py checks those constants against scipy's
erfc in an FP32 emulation of the kernel's evaluation order.
"""
from __future__ import annotations

import numpy as np

DEGREE = 9

# Coefficients of Q, low order first, as used by mdg_erfc_q.
COEFFS = (0.28209474704728354, 0.28210421927957413, 0.24651220482179143,
          0.18041489004864147, 0.05826160866113146, 0.07946913878476664,
          -0.1985488152101896, 0.030753194918781607, 0.06446153649429703,
          -0.02552288340259605)


def q_exact(t):
    from scipy.special import erfcx
    t = np.asarray(t, dtype=np.float64)
    safe = np.where(t > 0, t, 1.0)
    x = 2.0 * (1.0 - safe) / safe
    return np.where(t > 0, erfcx(x) / safe, 0.5 / np.sqrt(np.pi))


def fit(deg: int = DEGREE, nodes: int = 4000, sweeps: int = 30):
    t = 0.5 * (1.0 - np.cos(np.pi * (np.arange(nodes) + 0.5) / nodes))
    y = q_exact(t)
    w = 1.0 / y
    V = np.vander(t, deg + 1, increasing=True)
    for _ in range(sweeps):
        c, *_ = np.linalg.lstsq(V * w[:, None], y * w, rcond=None)
        err = (V @ c - y) / y
        w = w * (1.0 + 5.0 * np.abs(err) / np.abs(err).max())
    return c


def erfc_f32(x, coeffs=COEFFS):
    """The kernel's evaluation in FP32 (numpy float32 rounding per step;
    the device's exp and reciprocal differ from numpy's by an ulp or so)."""
    f = np.float32
    x = np.asarray(x, dtype=np.float32)
    t = (f(1.0) / (f(0.5) * x + f(1.0))).astype(np.float32)
    p = np.float32(coeffs[-1])
    for c in coeffs[-2::-1]:
        p = (t * p + f(c)).astype(np.float32)
    e = np.exp(-(x * x).astype(np.float32)).astype(np.float32)
    return (t * p * e).astype(np.float32)


if __name__ == "__main__":
    c = fit()
    tt = np.linspace(0.0, 1.0, 200001)
    rel = np.abs(np.polyval(c[::-1], tt) / q_exact(tt) - 1.0).max()
    print("max relative error of the fit:", rel)
    print(", ".join(repr(float(v)) for v in c))
