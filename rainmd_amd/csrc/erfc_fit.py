"""Fits the polynomials the kernels use for erfc and erf (device_common.h,
erfcWithGaussian and erfWithGaussian) and prints their coefficients.

erfc(x) = exp(-x^2) t P(t) with t = 1 / (1 + x/2) for x in [0, 4]: P, of
degree 8, fits erfcx(x) / t = exp(x^2) erfc(x) / t as a polynomial in t,
minimising the relative error (least squares on Chebyshev nodes in x,
reweighted towards the minimax: Lawson's iteration).

erf(x) = x Q(x^2) for x in [0, 0.8]: Q, of degree 5 in x^2, fits erf(x) / x
the same way.

The script also evaluates both in FP32 exactly as the kernels do (FMA Horner
steps, the hardware-rounded reciprocal emulated by a correctly rounded one)
against FP64, and checks the coefficients device_common.h holds against the
fit's. Run: python erfc_fit.py
"""
import math
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ERFC_DEGREE, ERFC_XMAX = 8, 4.0
ERF_DEGREE, ERF_XMAX = 5, 0.8


def lawson(u, y, degree, iterations=200, nodes_weight=None):
    """Polynomial in u of `degree` minimising max |p(u)/y - 1| (Lawson)."""
    v = np.vander(u, degree + 1, increasing=True)
    w = np.ones_like(u) if nodes_weight is None else nodes_weight
    for _ in range(iterations):
        s = np.sqrt(w) / np.abs(y)
        c, *_ = np.linalg.lstsq(v * s[:, None], y * s, rcond=None)
        err = np.abs(v @ c / y - 1.0)
        w = w * err
        w /= w.sum()
    return c


def chebyshev_nodes(lo, hi, n):
    k = np.arange(n)
    return lo + 0.5 * (hi - lo) * (1.0 - np.cos(math.pi * (k + 0.5) / n))


def erfcx(x):
    from scipy.special import erfcx as f

    return f(x)


def fit_erfc():
    x = chebyshev_nodes(0.0, ERFC_XMAX, 6000)
    t = 1.0 / (1.0 + 0.5 * x)
    return lawson(t, erfcx(x) / t, ERFC_DEGREE)


def fit_erf():
    x = chebyshev_nodes(0.0, ERF_XMAX, 4000)
    from scipy.special import erf

    return lawson(x * x, erf(x) / x, ERF_DEGREE)


def horner_f32(c, u):
    u = np.float32(u)
    p = np.float32(c[-1])
    for a in c[-2::-1]:
        p = np.float32(np.float64(p) * np.float64(u) + np.float64(np.float32(a)))  # one rounding: an FMA
    return p


def erfc_f32(c, x):
    x = np.float32(x)
    t = np.float32(1.0 / (np.float64(0.5) * np.float64(x) + 1.0))
    p = horner_f32(c, t)
    hi = np.float32(x * x)
    lo = np.float32(np.float64(x) * np.float64(x) - np.float64(hi))
    e = np.float32(np.float32(math.exp(-float(hi))) * np.float32(1.0 - lo))
    return float(np.float32(np.float32(e * t) * p))


def device_coefficients(name):
    """The coefficients of the Horner chain in device_common.h's `name`,
    highest degree first as they appear."""
    with open(os.path.join(HERE, "device_common.h")) as f:
        text = f.read()
    body = text[text.index(name):]
    body = body[:body.index("}")]
    vals = [float(v) for v in re.findall(r"([-+]?\d\.\d+e[-+]\d+)f", body)]
    return vals


def main():
    from scipy.special import erfc, erf

    c = fit_erfc()
    xs = np.linspace(0.0, ERFC_XMAX, 20001)
    ts = 1.0 / (1.0 + 0.5 * xs)
    fit_err = np.abs(np.exp(-xs * xs) * ts * np.polyval(c[::-1], ts) / erfc(xs) - 1.0).max()
    print("erfc P (t^8 .. t^0):", ", ".join("%.9e" % v for v in c[::-1]))
    print("  max relative fit error %.2e" % fit_err)
    f32 = np.array([erfc_f32(c, x) for x in xs[::10]])
    rel = f32 / erfc(xs[::10]) - 1.0
    print("  FP32 evaluation: max %.2e, mean %.2e" % (np.abs(rel).max(), rel.mean()))
    held = device_coefficients("erfcWithGaussian")[:ERFC_DEGREE + 1]
    print("  device_common.h holds:", ", ".join("%.9e" % v for v in held))
    print("  largest difference to the fit: %.2e" % np.abs(np.array(held) - np.float32(c[::-1])).max())
    q = fit_erf()
    x2 = np.linspace(1e-6, ERF_XMAX, 20001)
    err = np.abs(x2 * np.polyval(q[::-1], x2 * x2) / erf(x2) - 1.0).max()
    print("erf Q ((x^2)^5 .. 1):", ", ".join("%.9e" % v for v in q[::-1]))
    print("  max relative fit error %.2e" % err)
    held = device_coefficients("erfWithGaussian")[:ERF_DEGREE + 1]
    print("  device_common.h holds:", ", ".join("%.9e" % v for v in held))
    print("  largest difference to the fit: %.2e" % np.abs(np.array(held) - np.float32(q[::-1])).max())


if __name__ == "__main__":
    main()
