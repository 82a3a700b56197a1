"""The NPT statistics judge: what a barostat run's volume series says about
the ensemble it sampled, and whether two runs (the plain barostat and the
learned screen) sampled the same one.

The methodology is the reference's (README.md:46-51, docs/STATUS.md:57,
tools/oracle/check_npt_box.py): the density distribution, the isothermal
compressibility from volume fluctuations, kappa_T = <dV^2> / (k_B T <V>),
and a two-sample Kolmogorov-Smirnov test between runs; plus, for the
screen, the fraction of exact evaluations it skipped. Correlated samples
are handled with block averages and an integrated autocorrelation time, and
the KS test is run on samples thinned to one per correlation time.
"""
import math

import numpy as np

BOLTZMANN = 0.008314462618            # kJ/mol/K
BAR_TO_KJ_PER_MOL_NM3 = 0.06022140857
AVOGADRO = 6.02214076e23


def density(volumes_nm3, total_mass_amu):
    """Mass density in g/cm^3 for each volume."""
    v = np.asarray(volumes_nm3, dtype=np.float64)
    return total_mass_amu / AVOGADRO / (v * 1e-21)


def compressibility(volumes_nm3, temperature):
    """Isothermal compressibility in 1/bar from volume fluctuations."""
    v = np.asarray(volumes_nm3, dtype=np.float64)
    kappa = v.var() / (BOLTZMANN * temperature * v.mean())  # nm^3 / (kJ/mol)
    return kappa * BAR_TO_KJ_PER_MOL_NM3


def autocorrelation_time(x, max_lag=None):
    """Integrated autocorrelation time (in samples) with the self-consistent
    window tau_w = 5 * tau (Sokal); 1 for an uncorrelated series."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    if n < 4 or x.var() == 0:
        return 1.0
    y = x - x.mean()
    f = np.fft.rfft(y, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    acf /= acf[0]
    max_lag = max_lag or n // 2
    tau = 1.0
    for lag in range(1, max_lag):
        tau += 2.0 * acf[lag]
        if lag >= 5.0 * tau:
            break
    return max(tau, 1.0)


def block_average(x, blocks=10):
    """Mean and its standard error from `blocks` contiguous block means."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x) // blocks * blocks
    if n < blocks:
        return float(x.mean()), float("nan")
    means = x[:n].reshape(blocks, -1).mean(1)
    return float(x.mean()), float(means.std(ddof=1) / math.sqrt(blocks))


def kolmogorov_q(lam):
    """Q_KS(lambda) = 2 sum_{k>=1} (-1)^(k-1) exp(-2 k^2 lambda^2)."""
    if lam < 1e-3:
        return 1.0
    total = 0.0
    for k in range(1, 101):
        term = 2.0 * (-1) ** (k - 1) * math.exp(-2.0 * k * k * lam * lam)
        total += term
        if abs(term) < 1e-12:
            break
    return min(max(total, 0.0), 1.0)


def ks_two_sample(a, b):
    """Two-sample Kolmogorov-Smirnov statistic D and its asymptotic p-value."""
    a = np.sort(np.asarray(a, dtype=np.float64))
    b = np.sort(np.asarray(b, dtype=np.float64))
    both = np.concatenate([a, b])
    cdf_a = np.searchsorted(a, both, side="right") / len(a)
    cdf_b = np.searchsorted(b, both, side="right") / len(b)
    d = float(np.max(np.abs(cdf_a - cdf_b)))
    ne = len(a) * len(b) / (len(a) + len(b))
    lam = (math.sqrt(ne) + 0.12 + 0.11 / math.sqrt(ne)) * d
    return d, kolmogorov_q(lam)


def thin(x, tau):
    """One sample per ceil(tau) samples."""
    step = max(1, int(math.ceil(tau)))
    return np.asarray(x)[::step]


def summarise(volumes_nm3, total_mass_amu, temperature, blocks=10):
    v = np.asarray(volumes_nm3, dtype=np.float64)
    rho = density(v, total_mass_amu)
    mean, err = block_average(rho, blocks)
    tau = autocorrelation_time(v)
    return {"samples": int(len(v)), "density": mean, "density_error": err, "density_std": float(rho.std()),
            "compressibility": compressibility(v, temperature), "autocorrelation_time": tau,
            "effective_samples": float(len(v) / tau)}


def judge(plain_volumes, screened_volumes, total_mass_amu, temperature, exact_evaluations=None, moves=None):
    """Compare the plain barostat's volume series with the screened one's:
    density means (difference in units of the combined block error),
    compressibilities, and a KS test on samples thinned by the larger
    correlation time. With the screen's counts, the fraction of moves that
    needed no exact evaluation."""
    plain = summarise(plain_volumes, total_mass_amu, temperature)
    screened = summarise(screened_volumes, total_mass_amu, temperature)
    tau = max(plain["autocorrelation_time"], screened["autocorrelation_time"])
    d, p = ks_two_sample(thin(density(plain_volumes, total_mass_amu), tau),
                         thin(density(screened_volumes, total_mass_amu), tau))
    combined = math.hypot(plain["density_error"], screened["density_error"])
    out = {"plain": plain, "screened": screened, "ks_statistic": d, "ks_p": p,
           "density_difference_sigma": abs(plain["density"] - screened["density"]) / combined if combined > 0 else 0.0,
           "compressibility_ratio": screened["compressibility"] / plain["compressibility"]}
    if exact_evaluations is not None and moves:
        out["skipped_fraction"] = 1.0 - exact_evaluations / moves
    return out
