"""The one-step Lindblad spin relaxation against the golden-rule rate formula.

The initializer checks its P by computing T1 twice (lindbladInit_for-DMD-4.5.6/
ElectronPhonon.h:882-1036, T1_rate_dRho with obsdeg = drhosdeg = false and a
fixed smearing): once from the rate formula
    dS/dt = 2 pi / nk^2 sum_{k k' modes} -Re sum_b1b2 conj([S, G])_b1b2 [drho G (n + 1 - f') - (n + f) G drho']_b1b2
with G = g^-_kk' sqrt(delta(e_k - e_k' - w)), over every ordered pair, and once
from one Lindblad step on the first-order state drho = dRho_H1(S) (:864-878),
with dS = sum_k Tr(S_k drho_k) / nk (help_electron.cpp:13-23).  Its log prints
the same T1 both ways (Examples/.../lindbladInit.out: 217.073 ps).  Here the
identity is checked on random g, S and energies: the oracle's compute_P (P from
g^- and g^+_kk' = (g^-_k'k)^dagger) and evolve at t = 0 give the rate
formula's dS/dt to rounding, for the reference's perturbation and for any
Hermitian drho (the identity is linear in drho).  A planted mistake, g^+
without the dagger, breaks it."""
import numpy as np
import pytest


def _system(nb, seed, nk=5, nm=3):
    rng = np.random.default_rng(seed)
    T, sigma = 0.002, 0.003
    e = np.sort(rng.uniform(0, 0.01, (nk, nb)), axis=1)
    F = 1 / (np.exp((e - e.mean()) / T) + 1)

    def crand(*s):
        return rng.normal(size=s) + 1j * rng.normal(size=s)

    g = crand(nk, nk, nm, nb, nb) * 1e-3                      # g^-_kk' per ordered pair and mode
    w = rng.uniform(0.001, 0.004, (nk, nk, nm))
    w = 0.5 * (w + w.transpose(1, 0, 2))                       # w(q) = w(-q)
    S = crand(nk, nb, nb)
    S = 0.5 * (S + S.conj().transpose(0, 2, 1))
    drho = crand(nk, nb, nb) * 1e-3
    drho = 0.5 * (drho + drho.conj().transpose(0, 2, 1))
    return rng, T, sigma, e, F, g, w, S, drho


def _bose(x):
    return 1 / (np.exp(x) - 1)


def _drho_H1(e, F, T, H1):
    """dRho_H1 (ElectronPhonon.h:864-878): (f_i - f_j) / (e_i - e_j) H1_ij, the
    degenerate limit f(f - 1) / T H1_ij."""
    de = e[:, :, None] - e[:, None, :]
    df = F[:, :, None] - F[:, None, :]
    deg = np.abs(de) <= 1e-12
    fa = 0.5 * (F[:, :, None] + F[:, None, :])
    return np.where(deg, fa * (fa - 1) / T, df / np.where(deg, 1.0, de)) * H1


def _rate_Sdot(e, F, T, sigma, g, w, S, drho):
    """sum over ordered pairs and modes of T1_rate_dRho's contrib (ElectronPhonon.h:
    933-1001, no energy-window cut)."""
    nk = e.shape[0]
    pd = 1 / (sigma * np.sqrt(2 * np.pi))
    c = 0.0
    for k in range(nk):
        for kp in range(nk):
            dE = e[k][:, None] - e[kp][None, :]
            for lam in range(w.shape[2]):
                wq = w[k, kp, lam]
                nq = _bose(max(1e-3, wq / T))
                G = g[k, kp, lam] * np.sqrt(pd * np.exp(-(dE - wq) ** 2 / (2 * sigma ** 2)))
                SG = S[k] @ G - G @ S[kp]
                DG = drho[k] @ G @ np.diag(nq + 1 - F[kp]) - np.diag(nq + F[k]) @ G @ drho[kp]
                c -= np.sum((SG.conj() * DG).real)
    return c


def _lindblad_Sdot(oracle, e, F, T, sigma, g, w, S, drho, dagger=True):
    """sum_k Re Tr(S_k L(drho)_k) with L the linear part of evolve at f (prefac 1):
    (evolve(f + drho) - evolve(f - drho)) / 2, exact for the quadratic map."""
    nk, nb = e.shape
    k1, k2, P1, P2 = [], [], [], []
    for k in range(nk):
        for kp in range(k, nk):
            gp = g[kp, k].conj().transpose(0, 2, 1) if dagger else g[kp, k]
            a, b = oracle.compute_P(e[k], e[kp], w[k, kp], g[k, kp], gp, T, sigma)
            k1.append(k)
            k2.append(kp)
            P1.append(a)
            P2.append(b)
    k1, k2 = np.array(k1, np.uint64), np.array(k2, np.uint64)
    f = np.zeros((nk, nb, nb), complex)
    f[:, range(nb), range(nb)] = F
    up = oracle.evolve(k1, k2, np.array(P1), np.array(P2), e, (f + drho).reshape(nk, -1), t=0.0, prefac=1.0)
    dn = oracle.evolve(k1, k2, np.array(P1), np.array(P2), e, (f - drho).reshape(nk, -1), t=0.0, prefac=1.0)
    L = ((up - dn) / 2).reshape(nk, nb, nb)
    return sum(np.trace(S[k] @ L[k]).real for k in range(nk))


@pytest.mark.parametrize("nb", [1, 2, 3, 4])
def test_one_step_T1_equals_the_rate_formula(nb):
    from oracle import oracle
    rng, T, sigma, e, F, g, w, S, drho = _system(nb, 300 + nb)
    nk = e.shape[0]
    for d in (_drho_H1(e, F, T, S), drho):
        dS = sum(np.trace(S[k] @ d[k]).real for k in range(nk))
        T1_rate = -(dS / nk) / (2 * np.pi / nk ** 2 * _rate_Sdot(e, F, T, sigma, g, w, S, d))
        T1_lindblad = -dS / (2 * np.pi / nk * _lindblad_Sdot(oracle, e, F, T, sigma, g, w, S, d))
        assert T1_lindblad == pytest.approx(T1_rate, rel=1e-11), (T1_lindblad, T1_rate)


def test_planted_mistake_breaks_the_identity():
    from oracle import oracle
    rng, T, sigma, e, F, g, w, S, drho = _system(2, 310)
    d = _drho_H1(e, F, T, S)
    good = _rate_Sdot(e, F, T, sigma, g, w, S, d)
    bad = _lindblad_Sdot(oracle, e, F, T, sigma, g, w, S, d, dagger=False)  # g^+ without the dagger
    assert abs(bad / good - 1) > 1e-3
