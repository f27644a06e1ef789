"""The oracle's model e-e P at the equilibrium occupations (elecelec_model::
calc_P without exchange) against the e-e golden rule and the structural
identities of a Lindblad P."""
import numpy as np
import pytest
from scipy.stats import norm

from oracle import oracle


def _mesh_system(rng, km=(4, 3, 2), nkeep=20, nb=2, nw=3):
    cells = np.array([(a, b, c) for a in range(km[0]) for b in range(km[1]) for c in range(km[2])])
    keep = np.sort(rng.choice(len(cells), nkeep, replace=False))
    kv = cells[keep] / np.array(km)
    e = np.sort(rng.uniform(0, 2e-3, size=(nkeep, nb)), axis=1)
    f = 1 / (np.exp((e - 1e-3) / 9.5e-4) + 1)
    U = rng.normal(size=(nkeep, nw, nb)) + 1j * rng.normal(size=(nkeep, nw, nb))
    k1, k2 = np.meshgrid(np.arange(nkeep), np.arange(nkeep), indexing="ij")
    return km, kv, e, f, U, k1.ravel().astype(np.uint64), k2.ravel().astype(np.uint64)


@pytest.mark.parametrize("lindblad", [True, False])
def test_ee_P_diagonal_is_the_golden_rule(lindblad):
    """P1[(n n),(n' n')] of pair (k1, k2) = (1/nk_full) sum over k3 with
    k4 = k1 + k3 - k2 kept, and over (x, y), of |v <k1 n|k2 n'> <k3 x|k4 y>|^2
    (1 - f_k3x) f_k4y delta(e1n - e2n' + e3x - e4y), delta a normalised
    Gaussian cut at ethr; zero for k1 = k2."""
    rng = np.random.default_rng(4)
    km, kv, e, f, U, k1, k2 = _mesh_system(rng)
    nk, nb = e.shape
    vq = rng.uniform(0.5, 2.0, size=len(k1))
    sig, ethr, nkf = 4e-4, 1.2e-3, 30.0
    P1, P2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, sig, ethr, nkf, lindblad)
    code = lambda x: tuple(np.rint(np.array(km) * (x - np.floor(x))).astype(int) % np.array(km))
    cell = {}
    for i in range(nk):
        cell.setdefault(code(kv[i]), i)
    g = lambda x: np.where(np.abs(x) < ethr, norm.pdf(x, scale=sig), 0.0)
    for p in range(0, len(k1), 5):
        a, b = int(k1[p]), int(k2[p])
        diag = P1[p].reshape(nb, nb, nb, nb)[np.arange(nb), np.arange(nb)][:, np.arange(nb), np.arange(nb)]
        want = np.zeros((nb, nb))
        if a != b:
            o12 = vq[p] * (U[a].conj().T @ U[b])
            for c3 in range(nk):
                c4 = cell.get(code(kv[a] + kv[c3] - kv[b]))
                if c4 is None:
                    continue
                o34 = U[c3].conj().T @ U[c4]
                for n in range(nb):
                    for m in range(nb):
                        de = e[a, n] - e[b, m] + e[c3][:, None] - e[c4][None, :]
                        want[n, m] += np.sum(np.abs(o12[n, m] * o34) ** 2 * (1 - f[c3])[:, None] * f[c4][None, :] * g(de))
            want /= nkf
        assert np.allclose(diag.real, want, rtol=1e-11, atol=1e-14 * max(want.max(), 1e-300))
        assert np.abs(diag.imag).max() <= 1e-13 * max(np.abs(diag).max(), 1e-300)


def test_ee_P_reverse_pair_and_gram_form():
    """P2(k1, k2) = conj(P1(k2, k1)) (ElecElec_Model.h:188-192) and, for the
    Lindblad form, P1 reshaped as [(i1 i3),(i2 i4)] is Hermitian PSD (a sum of
    vec(A) vec(A)^H with weights (1 - f) f >= 0)."""
    rng = np.random.default_rng(6)
    km, kv, e, f, U, k1, k2 = _mesh_system(rng)
    nb = e.shape[1]
    vq = np.ones(len(k1))  # v(q) symmetric in q
    P1, P2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, 4e-4, 1.2e-3, 30.0, True)
    P1r, _ = oracle.ee_model_P(k2, k1, U, kv, e, f, vq, km, 4e-4, 1.2e-3, 30.0, True)
    s = np.abs(P1).max()
    assert s > 0 and np.abs(P2 - P1r.conj()).max() <= 1e-13 * s
    for p in range(len(k1)):
        G = P1[p].reshape(nb, nb, nb, nb).transpose(0, 2, 1, 3).reshape(nb * nb, nb * nb)
        if np.abs(G).max() == 0:
            continue
        assert np.abs(G - G.conj().T).max() <= 1e-13 * np.abs(G).max()
        ev = np.linalg.eigvalsh(0.5 * (G + G.conj().T))
        assert ev[0] >= -1e-12 * ev[-1]


def test_ee_exchange_diagonal_and_reverse_pair():
    """ee_antisymmetry: mee -> (mee - mee_ex) / 2 with mee_ex = v(k1-k4) <1|4><3|2>
    and the prefactor doubled, k1 = k2 pairs kept (k3 = k1 skipped): the P
    diagonal is 2 (1/nk_full) sum |(mee - mee_ex)/2|^2 (1-f3) f4 delta(de), and
    P2(k1, k2) = conj(P1(k2, k1)) still holds."""
    rng = np.random.default_rng(9)
    km, kv, e, f, U, k1, k2 = _mesh_system(rng, nb=2, nw=3)
    nk, nb = e.shape
    R = np.diag([7.0, 8.0, 9.0])
    from denmat_dynm_development_amd import scatt
    m = scatt.eimp_model("debye", 2e-6, 9.5e-4, abs(np.linalg.det(R)), scatt.reciprocal_metric(R), 3,
                         n_ionized=1.0, epsilon_background=3.0, degauss=4e-4, ethr=1.2e-3).as_dict()
    vq = np.array([oracle.eimp_vq(m, oracle.eimp_q2(m["GGT"], kv[a], kv[b])) for a, b in zip(k1, k2)])
    sig, ethr, nkf = 4e-4, 1.2e-3, 30.0
    P1, P2 = oracle.ee_model_P(k1, k2, U, kv, e, f, vq, km, sig, ethr, nkf, True, antisym_model=m)
    P1r, _ = oracle.ee_model_P(k2, k1, U, kv, e, f, vq, km, sig, ethr, nkf, True, antisym_model=m)
    assert np.abs(P2 - P1r.conj()).max() <= 1e-13 * np.abs(P1).max()
    code = lambda x: tuple(np.rint(np.array(km) * (x - np.floor(x))).astype(int) % np.array(km))
    cell = {}
    for i in range(nk):
        cell.setdefault(code(kv[i]), i)
    g = lambda x: np.where(np.abs(x) < ethr, norm.pdf(x, scale=sig), 0.0)
    vfun = lambda a, b: oracle.eimp_vq(m, oracle.eimp_q2(m["GGT"], kv[a], kv[b]))
    for p in range(0, len(k1), 7):
        a, b = int(k1[p]), int(k2[p])
        want = np.zeros((nb, nb))
        for c3 in range(nk):
            c4 = cell.get(code(kv[a] + kv[c3] - kv[b]))
            if c4 is None or (a == b == c3):
                continue
            o12 = vq[p] * (U[a].conj().T @ U[b])
            o34 = U[c3].conj().T @ U[c4]
            o14 = vfun(a, c4) * (U[a].conj().T @ U[c4])
            o32 = U[c3].conj().T @ U[b]
            for n in range(nb):
                for mm in range(nb):
                    mee = o12[n, mm] * o34 - o14[n][None, :] * o32[:, mm][:, None]
                    de = e[a, n] - e[b, mm] + e[c3][:, None] - e[c4][None, :]
                    want[n, mm] += 2 * np.sum(np.abs(mee / 2) ** 2 * (1 - f[c3])[:, None] * f[c4][None, :] * g(de))
        want /= nkf
        diag = P1[p].reshape(nb, nb, nb, nb)[np.arange(nb), np.arange(nb)][:, np.arange(nb), np.arange(nb)]
        assert np.allclose(diag.real, want, rtol=1e-11, atol=1e-14 * max(want.max(), 1e-300))
