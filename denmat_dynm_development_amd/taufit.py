"""Spin-lifetime fit of an S(t) series: the reference's post-processing
(Examples/RealTime_Example_GaAs_20230105/.../fit.py), restated.

- Global fit (fit.py:13,24-43): s(t) = s_0 exp(-beta t) with the amplitude
  fixed to the first sample, t in ps from the first sample (a.u. x
  2.4188843265857e-5), beta by least squares; tau = 1/beta.  Solved here by
  a safeguarded Newton iteration on dR/dbeta = 0 instead of MINPACK.
- Time-resolved lifetimes (fit.py:54-64): a straight line through each pair
  of consecutive (t, ln|s|) points, tau_t = -1/slope.
- The implicit derivative d tau / d s_m of the global fit (for the adjoint of
  a loss on tau): at the optimum G(beta, s) = sum_m r_m s_0 t_m e^{-beta t_m}
  = 0 with r_m = s_m - s_0 e^{-beta t_m}, so dbeta/ds = -(dG/ds) / (dG/dbeta).
"""
import numpy as np

AU_TO_PS = 2.4188843265857e-5


def _prep(t_au, s, tstart=0.0, tend=np.inf):
    t = np.asarray(t_au, np.float64) * AU_TO_PS
    s = np.asarray(s, np.float64)
    mask = (t >= tstart) & (t <= tend)
    t, s = t[mask], s[mask]
    return t - t[0], s, mask


def _G(beta, t, s):
    e = np.exp(-beta * t)
    r = s - s[0] * e
    g = np.sum(r * s[0] * t * e)
    dg = np.sum((s[0] * t * e) ** 2) - np.sum(r * s[0] * t * t * e)
    return g, dg


def fit_beta(t_au, s, tstart=0.0, tend=np.inf, beta0=None, tol=1e-15, max_iter=200):
    """beta (1/ps) of the global fit; tau = 1/beta ps."""
    t, s, _ = _prep(t_au, s, tstart, tend)
    if beta0 is None:  # log-linear estimate as the start
        ratio = np.abs(s / s[0])
        ok = (ratio > 0) & (t > 0)
        beta0 = -np.polyfit(t[ok], np.log(ratio[ok]), 1)[0] if ok.sum() >= 2 else 0.01
        if not np.isfinite(beta0) or beta0 == 0:
            beta0 = 0.01
    beta = float(beta0)
    for _ in range(max_iter):
        g, dg = _G(beta, t, s)
        step = -g / dg if dg > 0 else -np.sign(g) * 0.5 * abs(beta)
        # keep the step inside a factor of 2 of the current rate
        lim = max(abs(beta), 1e-12)
        step = float(np.clip(step, -0.5 * lim, lim))
        beta += step
        if abs(step) <= tol * abs(beta):
            break
    return beta


def fit_tau(t_au, s, tstart=0.0, tend=np.inf):
    """tau (ps) of the global exponential fit (fit.py)."""
    return 1.0 / fit_beta(t_au, s, tstart, tend)


def fit_curve(t_au, s, tstart=0.0, tend=np.inf):
    """(t_ps, s, s_fit) as fit.py writes them to sfit.out."""
    beta = fit_beta(t_au, s, tstart, tend)
    t, sd, _ = _prep(t_au, s, tstart, tend)
    return t, sd, sd[0] * np.exp(-beta * t)


def log_fit(t_au, s, tstart=0.0, tend=np.inf):
    """(t_ps[:-1], tau_t) of the time-resolved log fit (fit.py, tau_t.dat)."""
    t, sd, _ = _prep(t_au, s, tstart, tend)
    ls = np.log(np.abs(sd))
    slope = (ls[1:] - ls[:-1]) / (t[1:] - t[:-1])
    return t[:-1], -1.0 / slope


def dtau_ds(t_au, s, tstart=0.0, tend=np.inf):
    """(tau, d tau / d s) for the samples of s (zeros outside [tstart, tend])."""
    t, sd, mask = _prep(t_au, s, tstart, tend)
    beta = fit_beta(t_au, s, tstart, tend)
    e = np.exp(-beta * t)
    r = sd - sd[0] * e
    _, dg_dbeta = _G(beta, t, sd)
    dg_ds = sd[0] * t * e                                  # through r_m, m >= 1 (t_0 = 0)
    dg_ds[0] = np.sum(t * e * (r - sd[0] * e))             # s_0 enters every residual
    dbeta = -dg_ds / dg_dbeta
    out = np.zeros(len(mask))
    out[mask] = -dbeta / beta ** 2
    return 1.0 / beta, out
