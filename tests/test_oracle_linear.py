"""The oracle's linearized operator (set_Lsc + evolve_linear, restated from
src_v4.5.7/ElectronPhonon_linear.cpp:3-207 and evolve_eph_linear.cpp:3-53),
pinned by what the mathematics fixes: the nonlinear evolve is quadratic in
rho, so its central difference at diag(f_eq) is its exact Jacobian, and
evolve_linear must equal that Jacobian when every k has its diagonal pair;
rows without one lose exactly their own-row (Lscii) term, as in the
reference."""
import numpy as np
import pytest

from helpers import FS, random_P, random_energies, random_pairs, random_rho, rel_err
from oracle import oracle


def _case(nk, nb, rng, all_diag=True):
    k1, k2 = random_pairs(nk, 4, rng)
    pairs = set(zip(k1.tolist(), k2.tolist()))
    if all_diag:
        pairs |= {(k, k) for k in range(nk)}
    pr = sorted(pairs)
    k1 = np.array([a for a, _ in pr], np.uint64)
    k2 = np.array([b for _, b in pr], np.uint64)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    f = rng.uniform(0.05, 0.95, (nk, nb))
    return k1, k2, e, P1, P2, f


def _diag(f):
    nk, nb = f.shape
    d = np.zeros((nk, nb, nb), np.complex128)
    d[:, np.arange(nb), np.arange(nb)] = f
    return d.reshape(nk, nb * nb)


@pytest.mark.parametrize("nb", [1, 2, 3, 4])
@pytest.mark.parametrize("t_fs", [0.0, 37.0])
def test_linear_operator_is_the_jacobian_of_evolve(nb, t_fs):
    rng = np.random.default_rng(40 + nb)
    nk = 11
    k1, k2, e, P1, P2, f = _case(nk, nb, rng)
    prefac, t, h = 2 * np.pi / nk, t_fs * FS, 1e-3
    d = random_rho(nk, nb, rng, scale=1.0, diag=0.0) + _diag(rng.normal(size=(nk, nb)))
    fp = oracle.evolve(k1, k2, P1, P2, e, _diag(f) + h * d, t=t, prefac=prefac)
    fm = oracle.evolve(k1, k2, P1, P2, e, _diag(f) - h * d, t=t, prefac=prefac)
    jac = (fp - fm) / (2 * h)
    Lij, Lji = oracle.set_Lsc(k1, k2, P1, P2, f)
    lin = oracle.evolve_linear(k1, k2, Lij, Lji, e, _diag(f) + d, f, t=t, prefac=prefac)
    assert rel_err(lin, jac) < 1e-11, rel_err(lin, jac)
    # linear in drho, Hermitian rows
    lin2 = oracle.evolve_linear(k1, k2, Lij, Lji, e, _diag(f) + 2 * d, f, t=t, prefac=prefac)
    assert rel_err(lin2, 2 * lin) < 1e-14
    m = lin.reshape(nk, nb, nb)
    assert np.abs(m - m.conj().transpose(0, 2, 1)).max() == 0.0
    # a planted mistake: the wrong f in the forward matrix misses the Jacobian
    Lbad, _ = oracle.set_Lsc(k1, k2, P1, P2, np.roll(f, 1, axis=0))
    bad = oracle.evolve_linear(k1, k2, Lbad, Lji, e, _diag(f) + d, f, t=t, prefac=prefac)
    assert rel_err(bad, jac) > 1e-3


def test_rows_without_their_diagonal_pair_drop_lscii():
    rng = np.random.default_rng(7)
    nk, nb = 9, 2
    k1, k2, e, P1, P2, f = _case(nk, nb, rng, all_diag=False)
    keep = ~((k1 == k2) & (k1 == 4))  # row 4 has no diagonal pair
    k1, k2, P1, P2 = k1[keep], k2[keep], P1[keep], P2[keep]
    assert not np.any((k1 == 4) & (k2 == 4))
    prefac, h = 2 * np.pi / nk, 1e-3
    d = np.zeros((nk, nb * nb), np.complex128)
    d[4] = random_rho(1, nb, rng, scale=1.0, diag=0.5)[0]
    jac = (oracle.evolve(k1, k2, P1, P2, e, _diag(f) + h * d, prefac=prefac)
           - oracle.evolve(k1, k2, P1, P2, e, _diag(f) - h * d, prefac=prefac)) / (2 * h)
    Lij, Lji = oracle.set_Lsc(k1, k2, P1, P2, f)
    lin = oracle.evolve_linear(k1, k2, Lij, Lji, e, _diag(f) + d, f, prefac=prefac)
    others = np.arange(nk) != 4
    assert rel_err(lin[others], jac[others]) < 1e-11
    assert np.abs(lin[4]).max() == 0.0 and np.abs(jac[4]).max() > 0
