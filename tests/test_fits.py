"""The kernels' approximations come from the fit scripts committed beside
them: the erfc and erf polynomials (csrc/erfc_fit.py, device_common.h), the
rational screened-Coulomb force (csrc/ewald_force_fit.py, nonbonded.hip) and
the thermostat noise's sine and cosine (csrc/sincos_fit.py, integrate.hip)
are what the scripts produce, and the fits meet their stated accuracy."""
import math
import os
import re
import sys

import numpy as np
from scipy import special

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "rainmd_amd", "csrc")
sys.path.insert(0, CSRC)

import erfc_fit  # noqa: E402
import ewald_force_fit  # noqa: E402
import sincos_fit  # noqa: E402


def _held(path, name, count):
    with open(os.path.join(CSRC, path)) as f:
        text = f.read()
    body = text[text.index(name):]
    body = body[:body.index("}")]
    return [float(v) for v in re.findall(r"([-+]?\d\.\d+e[-+]\d+)f", body)][:count]


def test_erfc_and_erf_coefficients_are_the_fit_s():
    c = erfc_fit.fit_erfc()[::-1]
    assert np.allclose(_held("device_common.h", "erfcWithGaussian(float x", 9), np.float32(c), rtol=1e-8, atol=0)
    q = erfc_fit.fit_erf()[::-1]
    assert np.allclose(_held("device_common.h", "erfWithGaussian(float x", 6), np.float32(q), rtol=1e-8, atol=0)
    xs = np.linspace(0.0, 4.0, 4001)
    t = 1.0 / (1.0 + 0.5 * xs)
    assert np.abs(np.exp(-xs * xs) * t * np.polyval(c, t) / special.erfc(xs) - 1).max() < 2e-8
    x = np.linspace(1e-6, 0.8, 2001)
    assert np.abs(x * np.polyval(q, x * x) / special.erf(x) - 1).max() < 5e-9


def test_rational_force_coefficients_are_the_fit_s():
    p, q = ewald_force_fit.fit()
    held = _held("nonbonded.hip", "float ewaldG(float w)", 12)
    assert np.allclose(held[:6], np.float32(p[::-1]), rtol=2e-6, atol=1e-12)
    assert np.allclose(held[6:11], np.float32(q[::-1][:-1]), rtol=2e-6, atol=1e-12)
    ws = np.linspace(0.0, ewald_force_fit.W_MAX, 2001)
    exact = np.array([ewald_force_fit.g_exact(w) for w in ws])
    assert np.abs(np.polyval(p[::-1], ws) / np.polyval(q[::-1], ws) / exact - 1).max() < 1e-7


def test_noise_sine_and_cosine_are_the_fit_s():
    s, c = sincos_fit.fit()
    hs, hc = sincos_fit.device_coefficients()
    # the kernel's Horner chains run from the highest power down
    assert np.allclose(hs, s[::-1], rtol=1e-9, atol=0)
    assert np.allclose(hc, c[::-1], rtol=1e-9, atol=0)
    # the kernel's FP32 arithmetic over the counter hash's uniforms, and the
    # quadrant rotation at its edges (u = 1/8, 3/8, ... and near 0 and 1)
    u = np.concatenate([sincos_fit.noise_uniforms(500_000, seed=3),
                        ((np.arange(0, 1 << 23, 4099) + 0.5) * 2.0 ** -23).astype(np.float32),
                        np.array([0.5 * 2.0 ** -23, 1 - 0.5 * 2.0 ** -23, 0.125, 0.375, 0.625, 0.875],
                                 dtype=np.float32)])
    fs, fc = sincos_fit.sincos_f32(s, c, u)
    exact = 2 * np.pi * u.astype(np.float64)
    assert np.abs(fs - np.sin(exact)).max() < 1.5e-7
    assert np.abs(fc - np.cos(exact)).max() < 1.5e-7
