"""Fits the polynomials of the thermostat noise's sin(2 pi u) and cos(2 pi u)
(integrate.hip, sinCosTwoPi) and prints their coefficients.

The kernel reduces u in [0, 1] to r = u - q/4 with q = rint(4u), so
r in [-1/8, 1/8] (exact in FP32), and evaluates
    sin(2 pi r) = r S(r^2),    cos(2 pi r) = C(r^2)
with S of degree 4 and C of degree 4 in r^2, fitted for the smallest
maximum absolute error on [0, 1/64] (Lawson's iteration on Chebyshev nodes,
C's constant pinned to 1 so cos(0) is exact), then rotates by the quadrant q.
The script evaluates the kernel's FP32 arithmetic (FMA Horner steps) against
FP64 sin and cos over every u the counter hash can produce in a sample, and
checks the coefficients integrate.hip holds. Run: python sincos_fit.py
"""
import math
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEGREE = 4
R2MAX = 1.0 / 64.0


def lawson(u, y, degree, weight, fixed0=None, iterations=300):
    """Polynomial in u of `degree` minimising max |weight (p(u) - y)|."""
    first = 0 if fixed0 is None else 1
    v = np.vander(u, degree + 1, increasing=True)[:, first:]
    target = y - (0.0 if fixed0 is None else fixed0)
    w = np.ones_like(u)
    for _ in range(iterations):
        s = np.sqrt(w) * weight
        c, *_ = np.linalg.lstsq(v * s[:, None], target * s, rcond=None)
        err = np.abs(weight * (v @ c - target))
        w = w * err
        w /= w.sum()
    return c if fixed0 is None else np.concatenate([[fixed0], c])


def nodes(n):
    k = np.arange(n)
    return 0.5 * R2MAX * (1.0 - np.cos(math.pi * (k + 0.5) / n))


def fit():
    r2 = nodes(4000)
    r = np.sqrt(r2)
    # sin: the absolute error of r S(r^2) is r times that of S, so S is fitted
    # with the weight r
    s = lawson(r2, np.sin(2 * math.pi * r) / np.where(r > 0, r, 1.0), DEGREE, weight=np.maximum(r, 1e-30))
    c = lawson(r2, np.cos(2 * math.pi * r), DEGREE, weight=np.ones_like(r2), fixed0=1.0)
    return s, c


def horner_f32(coef, x):
    p = np.float32(coef[-1])
    for a in coef[-2::-1]:
        p = np.float32(np.float64(p) * np.float64(x) + np.float64(np.float32(a)))
    return p


def sincos_f32(s, c, u):
    """The kernel's arithmetic in FP32 for an array of FP32 u."""
    u = np.asarray(u, dtype=np.float32)
    q = np.rint(np.float32(4.0) * u).astype(np.float32)
    r = (np.float64(q) * -0.25 + np.float64(u)).astype(np.float32)  # exact
    r2 = (r * r).astype(np.float32)
    ps = np.full_like(r2, np.float32(s[-1]))
    for a in s[-2::-1]:
        ps = (ps.astype(np.float64) * r2 + np.float64(np.float32(a))).astype(np.float32)
    pc = np.full_like(r2, np.float32(c[-1]))
    for a in c[-2::-1]:
        pc = (pc.astype(np.float64) * r2 + np.float64(np.float32(a))).astype(np.float32)
    sr = (r * ps).astype(np.float32)
    k = q.astype(np.int64) & 3
    sin = np.select([k == 0, k == 1, k == 2], [sr, pc, -sr], -pc)
    cos = np.select([k == 0, k == 1, k == 2], [pc, -sr, -pc], sr)
    return sin.astype(np.float32), cos.astype(np.float32)


def device_coefficients():
    with open(os.path.join(HERE, "integrate.hip")) as f:
        text = f.read()
    body = text[text.index("void sinCosTwoPi("):]
    body = body[:body.index("\n}\n")]
    vals = [float(v) for v in re.findall(r"([-+]?\d\.\d+e[-+]\d+)f", body)]
    return vals[:DEGREE + 1], vals[DEGREE + 1:2 * (DEGREE + 1)]


def noise_uniforms(count, seed=1):
    """Uniforms of the counter hash's form ((h >> 9) + 0.5) 2^-23."""
    h = np.random.default_rng(seed).integers(0, 1 << 23, size=count)
    return ((h.astype(np.float64) + 0.5) * 2.0 ** -23).astype(np.float32)


def main():
    s, c = fit()
    print("sin S (r^2)^4 .. 1:", ", ".join("%.9e" % v for v in s[::-1]))
    print("cos C (r^2)^4 .. 1:", ", ".join("%.9e" % v for v in c[::-1]))
    u = noise_uniforms(2_000_000)
    fs, fc = sincos_f32(s, c, u)
    es = np.abs(fs - np.sin(2 * math.pi * u.astype(np.float64))).max()
    ec = np.abs(fc - np.cos(2 * math.pi * u.astype(np.float64))).max()
    print("FP32 evaluation, max absolute error: sin %.2e, cos %.2e" % (es, ec))
    try:
        hs, hc = device_coefficients()
        print("integrate.hip holds: sin", ", ".join("%.9e" % v for v in hs), "cos", ", ".join("%.9e" % v for v in hc))
    except ValueError:
        print("integrate.hip holds no sinCosTwoPi coefficients yet")


if __name__ == "__main__":
    main()
