"""The oracle's e-e P beyond the equilibrium form, pinned without the reference:

- P at a density matrix (eeMode = Pee_update, calc_A1rho): at dm = diag(f) it
  is the equilibrium P; for a general Hermitian state it equals a brute-force
  sum written from the reference's defining formula (ElecElec_Model.h:188-192,
  (I-rho)_{k3,b6,b5} A1_{..,(k3,b5),(k4,b7)} rho_{k4,b7,b8} conj(A2_{..,(k3,b6),(k4,b8)}))
  with explicit index loops; P2(k1, k2) = conj(P1(k2, k1)) holds at any state.
- The linearized e-e operator (set_Lsc with P_ee at f_eq, fac2 for exchange,
  P3..P6): with the P3/P4 terms added it is the exact Jacobian of the e-e
  right-hand side evolve(P_ee(rho), rho) at rho = diag(f_eq), for nb = 1..3,
  both scattering forms, with and without exchange; the reference subtracts
  P3/P4 (ElectronPhonon_linear.cpp:209-231), which the tests show is not the
  Jacobian unless exchange removes those terms.
- The static RPA v(q) per mesh cell agrees with the per-pair table.
"""
import numpy as np
import pytest

from oracle import oracle


def _system(rng, nb=2, km=(3, 3, 2), nkeep=12, nw=3):
    cells = np.array([(a, b, c) for a in range(km[0]) for b in range(km[1]) for c in range(km[2])])
    keep = np.sort(rng.choice(len(cells), nkeep, replace=False))
    kv = cells[keep] / np.array(km)
    e = np.sort(rng.uniform(0, 2e-3, size=(nkeep, nb)), axis=1)
    f = 1 / (np.exp((e - 1e-3) / 9.5e-4) + 1)
    U = rng.normal(size=(nkeep, nw, nb)) + 1j * rng.normal(size=(nkeep, nw, nb))
    return km, kv, e, f, U


def _debye():
    GGT = np.diag([1.3, 1.1, 0.9])
    return {"GGT": GGT, "screening": 1, "prefac_vq": 0.7, "prefac_vq_bare": 8.4, "qscr2_debye": 0.3,
            "fac0": 0.0, "fac2": 0.0, "fac4": 0.0}


def _vq_pairs(m, kv, k1, k2):
    return np.array([oracle.eimp_vq(m, oracle.eimp_q2(m["GGT"], kv[int(a)], kv[int(b)])) for a, b in zip(k1, k2)])


def _herm_state(rng, f, amp):
    nk, nb = f.shape
    A = rng.normal(size=(nk, nb, nb)) + 1j * rng.normal(size=(nk, nb, nb))
    r = amp * 0.5 * (A + A.conj().transpose(0, 2, 1))
    r[:, np.arange(nb), np.arange(nb)] += f
    return r.reshape(nk, nb * nb)


def _one_minus(dm, nb):
    d1 = -dm.copy()
    d1[:, ::nb + 1] += 1.0
    return d1


SIG, ETHR, NKF = 4e-4, 1.2e-3, 30.0


def test_P_at_equilibrium_state_equals_fixed_P():
    rng = np.random.default_rng(1)
    km, kv, e, f, U = _system(rng)
    nk, nb = e.shape
    k1, k2 = [x.ravel().astype(np.uint64) for x in np.meshgrid(np.arange(nk), np.arange(nk), indexing="ij")]
    m = _debye()
    vq = _vq_pairs(m, kv, k1, k2)
    dm = np.zeros((nk, nb * nb), complex)
    dm[:, ::nb + 1] = f
    for lb in (True, False):
        P1, P2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, lb)
        Q1, Q2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, lb, dm=dm, dm1=_one_minus(dm, nb))
        s = np.abs(P1).max()
        assert s > 0 and np.abs(Q1 - P1).max() <= 1e-14 * s and np.abs(Q2 - P2).max() <= 1e-14 * s


@pytest.mark.parametrize("nb", [1, 2, 3])
def test_P_at_a_state_against_brute_force(nb):
    """P1 and P2 at a Hermitian state with off-diagonal entries against the
    reference's defining sums written as explicit loops over (k3, b5..b8)."""
    rng = np.random.default_rng(10 + nb)
    km, kv, e, f, U = _system(rng, nb=nb, nkeep=9)
    nk = e.shape[0]
    k1, k2 = [x.ravel().astype(np.uint64) for x in np.meshgrid(np.arange(nk), np.arange(nk), indexing="ij")]
    m = _debye()
    vq = _vq_pairs(m, kv, k1, k2)
    dm = _herm_state(rng, f, 0.05)
    dm1 = _one_minus(dm, nb)
    P1, P2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, True, dm=dm, dm1=dm1)
    code = lambda x: tuple(np.rint(np.array(km) * (x - np.floor(x))).astype(int) % np.array(km))
    cell = {}
    for i in range(nk):
        cell.setdefault(code(kv[i]), i)
    pg = 1 / (SIG * np.sqrt(2 * np.pi))
    R = dm.reshape(nk, nb, nb)
    R1 = dm1.reshape(nk, nb, nb)
    for p in range(0, len(k1), 11):
        a, b = int(k1[p]), int(k2[p])
        W1 = np.zeros((nb,) * 4, complex)
        W2 = np.zeros((nb,) * 4, complex)
        if a != b:
            o12 = vq[p] * (U[a].conj().T @ U[b])
            for c3 in range(nk):
                c4 = cell.get(code(kv[a] + kv[c3] - kv[b]))
                if c4 is None:
                    continue
                o34 = U[c3].conj().T @ U[c4]
                # A[(p q),(x y)] = v <a p|b q> <c3 x|c4 y> sqrt-Gaussian(e_ap - e_bq + e_c3x - e_c4y)
                A = np.zeros((nb,) * 4, complex)
                for i in range(nb):
                    for j in range(nb):
                        for x in range(nb):
                            for y in range(nb):
                                de = e[a, i] - e[b, j] + e[c3, x] - e[c4, y]
                                dl = np.exp(-0.25 * de * de / SIG ** 2) if abs(de) < ETHR else 0.0
                                A[i, j, x, y] = o12[i, j] * o34[x, y] * dl
                for b1 in range(nb):
                    for b2 in range(nb):
                        for b3 in range(nb):
                            for b4 in range(nb):
                                s1 = s2 = 0.0
                                for b5 in range(nb):
                                    for b6 in range(nb):
                                        for b7 in range(nb):
                                            for b8 in range(nb):
                                                s1 += (R1[c3, b6, b5] * A[b1, b3, b5, b7] * R[c4, b7, b8]
                                                       * np.conj(A[b2, b4, b6, b8]))
                                                s2 += (R[c3, b6, b5] * A[b3, b1, b5, b7] * R1[c4, b7, b8]
                                                       * np.conj(A[b4, b2, b6, b8]))
                                W1[b1, b2, b3, b4] += s1
                                W2[b1, b2, b3, b4] += s2
        W1 *= pg / NKF
        W2 *= pg / NKF
        s = max(np.abs(P1).max(), 1e-300)
        assert np.abs(P1[p] - W1.reshape(-1)).max() <= 1e-13 * s, p
        assert np.abs(P2[p] - W2.reshape(-1)).max() <= 1e-13 * s, p
    # reverse pair at a general state, and a transposed state breaks the sums
    P1r, _ = oracle.ee_model_P(k2, k1, U, kv, e, f, vq, km, SIG, ETHR, NKF, True, dm=dm, dm1=dm1)
    assert np.abs(P2 - P1r.conj()).max() <= 1e-13 * np.abs(P1).max()
    if nb > 1:
        dmT = dm.reshape(nk, nb, nb).transpose(0, 2, 1).reshape(nk, nb * nb).copy()
        T1, _ = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, True, dm=dmT, dm1=_one_minus(dmT, nb))
        assert np.abs(T1 - P1).max() > 1e-6 * np.abs(P1).max()


def _jacobian_case(nb, antisym, lindblad, seed=3):
    rng = np.random.default_rng(seed)
    km, kv, e, f, U = _system(rng, nb=nb)
    nk = e.shape[0]
    pr = [(a, b) for a in range(nk) for b in range(a, nk)]
    k1 = np.array([a for a, _ in pr], np.uint64)
    k2 = np.array([b for _, b in pr], np.uint64)
    m = _debye()
    vq = _vq_pairs(m, kv, k1, k2)
    am = m if antisym else None
    prefac = 2 * np.pi / nk

    def G(rho):
        P1, P2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, lindblad, antisym_model=am, dm=rho,
                                   dm1=_one_minus(rho, nb))
        return oracle.evolve(k1, k2, P1, P2, e, rho, t=0.0, prefac=prefac, expt=False)

    rho0 = np.zeros((nk, nb * nb), complex)
    rho0[:, ::nb + 1] = f
    d = _herm_state(rng, np.zeros_like(f), 0.1)
    # G is a polynomial of degree 4 in rho: Richardson on two central differences is exact to rounding
    J1 = (G(rho0 + 1e-3 * d) - G(rho0 - 1e-3 * d)) / 2e-3
    J2 = (G(rho0 + 5e-4 * d) - G(rho0 - 5e-4 * d)) / 1e-3
    J = (4 * J2 - J1) / 3
    P1f, P2f = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, SIG, ETHR, NKF, lindblad, antisym_model=am)
    Lij, Lji = oracle.set_Lsc(k1, k2, P1f, P2f, f)
    P3456 = oracle.ee_model_P3P6(k1, k2, U, kv, e, f, km, SIG, ETHR, NKF, m, lindblad, antisym)

    def L(sign34):
        P3, P4, P5, P6 = P3456
        a, b = oracle.add_Lsc_ee(k1, k2, f, Lij, Lji, extra=1.0 if antisym else 0.0, Pee=(P1f, P2f),
                                 P3456=(sign34 * P3, sign34 * P4, P5, P6))
        return oracle.evolve_linear(k1, k2, a, b, e, rho0 + d, f, t=0.0, prefac=prefac, expt=False)

    err = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    return J, L, err, oracle.evolve_linear(k1, k2, Lij, Lji, e, rho0 + d, f, t=0.0, prefac=prefac, expt=False)


@pytest.mark.parametrize("nb", [1, 2, 3])
@pytest.mark.parametrize("lindblad", [True, False])
def test_linearized_ee_is_the_exact_jacobian(nb, lindblad):
    """Without exchange: Lsc(P_ee) + P3/P4 (added) + P5/P6 is the Jacobian;
    the P1/P2 form alone, and the reference's subtracted P3/P4, are not."""
    J, L, err, L_fixed = _jacobian_case(nb, False, lindblad)
    assert err(L(-1.0), J) < 1e-10, err(L(-1.0), J)
    assert err(L_fixed, J) > 1e-3
    if nb > 1 or lindblad:
        assert err(L(+1.0), J) > 1e-3  # the reference's sign: not the Jacobian


@pytest.mark.parametrize("nb", [1, 2, 3])
def test_linearized_ee_with_exchange_is_the_exact_jacobian(nb):
    """With exchange the reference skips P3/P4 and doubles the P1/P2 form's
    Lscij/Lscji (fac2): with P5/P6 that is the exact Jacobian."""
    J, L, err, L_fixed = _jacobian_case(nb, True, True)
    assert err(L(1.0), J) < 1e-10, err(L(1.0), J)
    assert err(L_fixed, J) > 1e-3


def test_rpa_cell_table_matches_the_pair_table():
    rng = np.random.default_rng(21)
    km, kv, e, f, U = _system(rng, nb=2, nkeep=14)
    nk = e.shape[0]
    k1, k2 = [x.ravel().astype(np.uint64) for x in np.meshgrid(np.arange(nk), np.arange(nk), indexing="ij")]
    from denmat_dynm_development_amd import scatt
    R = np.diag([7.0, 8.0, 9.0])
    m = scatt.eimp_model("RPA", 2e-6, 9.5e-4, abs(np.linalg.det(R)), scatt.reciprocal_metric(R), 3, n_ionized=1.0,
                         epsilon_background=3.0, degauss=4e-4, ethr=1.2e-3, kmesh=km, nk_full=30.0).as_dict()
    vq, qs, cells = oracle.eimp_rpa_vq(k1, k2, U, kv, e, f, m, cells=True)
    code = lambda x: int(np.dot(np.rint(np.array(km) * (x - np.floor(x))).astype(int) % np.array(km),
                                [km[1] * km[2], km[2], 1]))
    got = np.array([cells[code(kv[int(a)] - kv[int(b)])] for a, b in zip(k1, k2)])
    assert np.abs(got - vq).max() <= 1e-14 * np.abs(vq).max()
    assert np.count_nonzero(cells) == len(qs) - 1  # every q but q = 0
