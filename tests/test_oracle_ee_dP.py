"""The oracle's alg_linearize_dPee pieces, pinned without the reference:

- dP (calc_P with dP1, dP2, ElecElec_Model.h:185-262 with calc_A1rho_dP,
  :271-281).  P at a state is bilinear in (dm, dm1), so its derivative at
  (diag f_eq, diag(1 - f_eq)) along any pair of directions is exact from one
  central difference.  The reference's calc_A1rho_dP takes f_eq[i] from every
  element of row i (ddm[i][j] = dm[i][j] - f[i]) before a zhemm that reads the
  upper triangle: dP is that derivative along the Hermitian directions built
  from those row-subtracted upper triangles, and for nb = 1, where no row
  subtraction enters, the derivative along rho - f_eq.  For nb > 1 it is not
  the derivative along rho - f_eq, which the test records.
- The evolve-side term (evolve_eph.cpp:107-131,138,144-151) equals evolve of
  dP at diag(f_eq), diag(1 - f_eq) without phases (the pinned oracle_evolve)
  with cis((e_i - e_j) t) on element (i, j).
"""
import numpy as np
import pytest

from oracle import oracle
from test_oracle_ee_dynamic import ETHR, NKF, SIG, _debye, _herm_state, _one_minus, _system, _vq_pairs


def _row_subtracted(M, g):
    """Hermitian matrices from the upper triangles of M[i][j] - g[i] (real diagonal)."""
    nk, nb = g.shape
    X = M.reshape(nk, nb, nb) - g[:, :, None]
    up = np.triu(X, 1)
    D = up + up.conj().transpose(0, 2, 1)
    D[:, np.arange(nb), np.arange(nb)] = X[:, np.arange(nb), np.arange(nb)].real
    return D.reshape(nk, nb * nb)


def _diag(v):
    nk, nb = v.shape
    d = np.zeros((nk, nb * nb), complex)
    d[:, ::nb + 1] = v
    return d


@pytest.mark.parametrize("nb", [1, 2, 3])
@pytest.mark.parametrize("lindblad,exchange", [(True, False), (False, False), (True, True)])
def test_dP_is_the_derivative_along_the_row_subtracted_directions(nb, lindblad, exchange):
    rng = np.random.default_rng(40 + nb + 5 * lindblad + 11 * exchange)
    km, kv, e, f, U = _system(rng, nb=nb, nkeep=10)
    nk = e.shape[0]
    pr = [(a, b) for a in range(nk) for b in range(a, nk)]
    k1 = np.array([a for a, _ in pr], np.uint64)
    k2 = np.array([b for _, b in pr], np.uint64)
    m = _debye()
    vq = _vq_pairs(m, kv, k1, k2)
    am = m if exchange else None
    dm = _herm_state(rng, f, 0.05)
    dm1 = _one_minus(dm, nb) + _herm_state(rng, np.zeros_like(f), 0.02)  # the two sides independent
    P1, P2, D1, D2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, lindblad, antisym_model=am,
                                       dm=dm, dm1=dm1, with_dP=True)
    F1, F2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, lindblad, antisym_model=am)
    s = np.abs(F1).max()
    assert s > 0 and np.abs(P1 - F1).max() <= 1e-14 * s and np.abs(P2 - F2).max() <= 1e-14 * s  # P stays at f_eq
    dR, dL = _row_subtracted(dm, f), _row_subtracted(dm1, 1 - f)

    def P_at(eps):
        return oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, lindblad, antisym_model=am,
                                 dm=_diag(f) + eps * dR, dm1=_diag(1 - f) + eps * dL)

    Pp, Pm = P_at(1.0), P_at(-1.0)
    J1, J2 = (Pp[0] - Pm[0]) / 2, (Pp[1] - Pm[1]) / 2
    sd = np.abs(J1).max()
    assert sd > 0
    assert np.abs(D1 - J1).max() <= 1e-12 * sd and np.abs(D2 - J2).max() <= 1e-12 * sd
    # the derivative along rho - f_eq (dm1 = 1 - dm): dP for nb = 1 only
    dm1c = _one_minus(dm, nb)
    _, _, C1, C2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, lindblad, antisym_model=am,
                                     dm=dm, dm1=dm1c, with_dP=True)
    d = dm - _diag(f)
    Q = [oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, lindblad, antisym_model=am,
                           dm=_diag(f) + sg * d, dm1=_one_minus(_diag(f) + sg * d, nb)) for sg in (1.0, -1.0)]
    G1 = (Q[0][0] - Q[1][0]) / 2
    if nb == 1:
        assert np.abs(C1 - G1).max() <= 1e-12 * np.abs(G1).max()
    else:
        assert np.abs(C1 - G1).max() > 1e-3 * np.abs(G1).max()


@pytest.mark.parametrize("nb", [1, 2, 3, 4])
def test_dP_term_is_evolve_of_dP_at_f_eq_with_row_phases(nb):
    rng = np.random.default_rng(60 + nb)
    nk = 9
    pr = [(a, b) for a in range(nk) for b in range(a, nk) if a == b or rng.uniform() < 0.6]
    k1 = np.array([a for a, _ in pr], np.uint64)
    k2 = np.array([b for _, b in pr], np.uint64)
    n4 = nb ** 4
    dP1 = rng.normal(size=(len(pr), n4)) + 1j * rng.normal(size=(len(pr), n4))
    dP2 = rng.normal(size=(len(pr), n4)) + 1j * rng.normal(size=(len(pr), n4))
    e = np.sort(rng.uniform(0, 2e-3, size=(nk, nb)), axis=1)
    f = rng.uniform(0.05, 0.95, size=(nk, nb))
    prefac, t = 0.37, 4.1e3
    got = oracle.ee_dP_term(k1, k2, dP1, dP2, e, f, t=t, prefac=prefac)
    H = oracle.evolve(k1, k2, dP1, dP2, e, _diag(f), dm1=_diag(1 - f), t=0.0, prefac=prefac, expt=False)
    ph = np.exp(1j * (e[:, :, None] - e[:, None, :]) * t).reshape(nk, nb * nb)
    ref = H * ph
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    # the term at t = 0 is Hermitian per k, and the phase matters off the diagonal
    H0 = oracle.ee_dP_term(k1, k2, dP1, dP2, e, f, t=0.0, prefac=prefac).reshape(nk, nb, nb)
    assert np.abs(H0 - H0.conj().transpose(0, 2, 1)).max() <= 1e-14 * np.abs(H0).max()
    if nb > 1:
        assert np.abs(got - H0.reshape(nk, -1)).max() > 1e-3 * np.abs(H0).max()
