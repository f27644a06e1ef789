"""The restated spin-lifetime fit (denmat_dynm_development_amd/taufit.py)
against the reference's own fit outputs for its GaAs example run (sfit.out,
tau_t.dat; fixture written by tests/golden/make_gaas_tau.py), a closed form,
scipy's curve_fit, and finite differences of its implicit derivative."""
import os

import numpy as np
import scipy.optimize

from denmat_dynm_development_amd import taufit

HERE = os.path.dirname(os.path.abspath(__file__))


def _gaas():
    return np.load(os.path.join(HERE, "golden", "gaas_tau_fit.npz"))


def test_global_fit_matches_reference_sfit():
    d = _gaas()
    t, s, fit = taufit.fit_curve(d["t_au"], d["s"])
    assert np.array_equal(t, d["sfit_t"]) or np.abs(t - d["sfit_t"]).max() < 1e-12
    assert np.abs(s - d["sfit_s"]).max() == 0.0
    # MINPACK's stopping tolerance (1.49e-8 relative on beta) bounds the agreement
    assert np.abs(fit - d["sfit_fit"]).max() / np.abs(d["sfit_fit"]).max() < 1e-8
    beta_ref = -np.log(d["sfit_fit"][-1] / d["sfit_fit"][0]) / d["sfit_t"][-1]
    assert abs(taufit.fit_beta(d["t_au"], d["s"]) / beta_ref - 1) < 1e-8
    assert abs(taufit.fit_tau(d["t_au"], d["s"]) - 34.58) < 0.01  # the run's printed lifetime


def test_log_fit_matches_reference_tau_t():
    d = _gaas()
    t, tau = taufit.log_fit(d["t_au"], d["s"])
    assert np.abs(t - d["taut_t"]).max() < 1e-12
    assert np.abs(tau / d["taut"] - 1).max() < 1e-7  # tau_t.dat is printed to 8 digits


def test_exact_exponential_and_curve_fit():
    rng = np.random.default_rng(3)
    t_au = np.linspace(0, 4e6, 300)
    tp = t_au * taufit.AU_TO_PS
    s = -2.5e-11 * np.exp(-0.037 * tp)
    assert abs(taufit.fit_beta(t_au, s) / 0.037 - 1) < 1e-12
    noisy = s * (1 + 0.02 * rng.normal(size=s.shape)) + 1e-13 * np.sin(tp)
    noisy[0] = s[0]
    popt, _ = scipy.optimize.curve_fit(lambda x, b: noisy[0] * np.exp(-b * x), tp, noisy, p0=[0.01],
                                       xtol=1e-14, ftol=1e-14)
    assert abs(taufit.fit_beta(t_au, noisy) / popt[0] - 1) < 1e-9


def test_implicit_derivative_matches_finite_differences():
    d = _gaas()
    t, s = d["t_au"], d["s"]
    tau, g = taufit.dtau_ds(t, s)
    h = 1e-6 * np.abs(s).max()
    for j in [0, 5, 100, 500, len(s) - 1]:
        sp, sm = s.copy(), s.copy()
        sp[j] += h
        sm[j] -= h
        fd = (taufit.fit_tau(t, sp) - taufit.fit_tau(t, sm)) / (2 * h)
        assert abs(g[j] - fd) <= 1e-4 * np.abs(g).max(), (j, g[j], fd)
    # a planted sign error in the s_0 term would be caught: its magnitude dominates
    assert abs(g[0]) > 10 * np.abs(g[1:]).max()
