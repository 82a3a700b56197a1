"""Fits the rational function the direct-space kernel uses for the screened
Coulomb force (nonbonded.hip, ewaldG) and prints its coefficients.

The force scalar of a pair, qq (erfc(z)/r + 2 alpha/sqrt(pi) exp(-z^2)) / r^2
with z = alpha r, equals qq (1/r^3 - alpha^3 G(w)) with w = z^2 and

    G(w) = (erf(z) - 2 z / sqrt(pi) exp(-w)) / z^3,

which is smooth in w (G(0) = 4 / (3 sqrt(pi))). G is fitted on [0, W_MAX]
by P(w) / Q(w), both of degree 5 with Q(0) = 1, minimising the relative
error (linearised least squares on Chebyshev nodes, reweighted towards the
minimax). Run: python ewald_force_fit.py
"""
import math

import numpy as np

W_MAX = 12.0  # (alpha rc)^2 up to 12: Ewald tolerances down to ~1e-6
DEGREE = 5


def g_exact(w):
    z = math.sqrt(w)
    if z < 1e-3:
        return 4.0 / (3.0 * math.sqrt(math.pi)) * (1.0 - 3.0 * w / 5.0 + 3.0 * w * w / 14.0)
    return (math.erf(z) - 2.0 * z / math.sqrt(math.pi) * math.exp(-w)) / z ** 3


def fit(n_p=DEGREE, n_q=DEGREE, w_max=W_MAX, iterations=60, nodes=4000):
    k = np.arange(nodes)
    x = 0.5 * w_max * (1.0 - np.cos(math.pi * (k + 0.5) / nodes))
    y = np.array([g_exact(v) for v in x])
    weight = 1.0 / np.abs(y)
    q_at = np.ones_like(x)
    for _ in range(iterations):
        a = np.hstack([np.vander(x, n_p + 1, increasing=True),
                       -(y[:, None] * np.vander(x, n_q + 1, increasing=True)[:, 1:])])
        s = weight / np.abs(q_at)
        sol, *_ = np.linalg.lstsq(a * s[:, None], y * s, rcond=None)
        p = sol[:n_p + 1]
        q = np.concatenate([[1.0], sol[n_p + 1:]])
        q_at = np.polyval(q[::-1], x)
        err = (np.polyval(p[::-1], x) / q_at - y) / y
        weight = weight * (np.abs(err) / np.abs(err).max()) ** 0.3 + 1e-300
    return p, q


def main():
    p, q = fit()
    xs = np.linspace(0.0, W_MAX, 200001)
    exact = np.array([g_exact(v) for v in xs])
    rel = (np.polyval(p[::-1], xs) / np.polyval(q[::-1], xs) - exact) / exact
    print("max relative error (FP64) %.3e; min Q %.3f" % (np.abs(rel).max(), np.polyval(q[::-1], xs).min()))
    print("P:", ", ".join("%.9e" % c for c in p))
    print("Q:", ", ".join("%.9e" % c for c in q))


if __name__ == "__main__":
    main()
