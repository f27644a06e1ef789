"""The oracle's model ionized-impurity term (elecimp_model with the |q|-only
static screening of coulomb_model) pinned against the reference's own GaAs
e-ph + e-i run output (tests/golden/gaas_vq.npz, from its ldbd_vq.out) and
against the golden-rule and structural identities the P it builds must obey."""
import os

import numpy as np
import pytest
from scipy.stats import norm

from denmat_dynm_development_amd import ldbd, scatt
from oracle import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOHR2CM = 5.291772109038e-9  # comm/Units.h


def _wrapped_q2(kvec, GGT):
    d = kvec[:, None, :] - kvec[None, :, :]
    d = d - np.floor(0.5 + d)
    d[np.abs(d - 0.5) < 1e-6] = -0.5
    return np.einsum("ija,ab,ijb->ij", d, GGT, d)


def test_q_lengths_and_vq_against_the_reference_table():
    """Every |q|^2 of the reference's ldbd_vq.out (6807 rows, 8 digits) is the
    length of a wrapped k - k' of its k mesh under GGT = (2 pi)^2 R^-1 R^-T,
    the smallest nonzero one is its |qmin|^2, and v = 4 pi / (eps V_cell) /
    (q^2 + q_scr^2) with its own q_scr^2 column reproduces its v column."""
    g = np.load(os.path.join(GOLD, "gaas_vq.npz"))
    d = os.path.join(GOLD, "gaas_ldbd")
    info = ldbd.read_info(d)
    kv = ldbd.read_kvec(d, info)
    _, R = ldbd.read_R(d)
    GGT = scatt.reciprocal_metric(R)
    q2 = _wrapped_q2(kv, GGT)
    ii, jj = np.random.default_rng(0).integers(0, info.nk, size=(2, 50))
    for i, j in zip(ii, jj):  # the oracle's own wrap and metric
        assert oracle.eimp_q2(GGT, kv[i], kv[j]) == pytest.approx(q2[i, j], rel=1e-14, abs=1e-300)
    u = np.unique(q2.round(15))
    nz = u[u > 1e-20]
    assert nz[0] == pytest.approx(float(g["qmin2"]), rel=2e-4)  # printed as 2.815e-05
    tab = g["q2"]
    pos = np.clip(np.searchsorted(u, tab), 1, len(u) - 1)
    near = np.minimum(np.abs(u[pos] - tab), np.abs(u[pos - 1] - tab))
    assert np.all(near <= 6e-8 * np.maximum(tab, 1e-12) + 1e-15), near.max()
    V = abs(np.linalg.det(R))
    m = {"screening": 1, "prefac_vq": 4 * np.pi / 12.9 / V, "prefac_vq_bare": 4 * np.pi / V, "fac0": 0, "fac2": 0,
         "fac4": 0}
    for q, s, v in zip(tab[::7], g["qscr2"][::7], g["vq"][::7]):
        m["qscr2_debye"] = s
        assert oracle.eimp_vq(m, q) == pytest.approx(v, rel=6e-8)


def test_screening_lengths_against_the_reference_header():
    """n_free = 2e16 cm^-3, eps = 12.9, T of the run: the Debye q^2 and the
    Thomas-Fermi k_F^2, q_TF^2 (meff = 1) the reference printed."""
    g = np.load(os.path.join(GOLD, "gaas_vq.npz"))
    d = os.path.join(GOLD, "gaas_ldbd")
    info = ldbd.read_info(d)
    _, R = ldbd.read_R(d)
    n = float(g["nfree_cm3"]) * BOHR2CM ** 3
    GGT = scatt.reciprocal_metric(R)
    m = scatt.eimp_model("Bechstedt", n, info.temperature, abs(np.linalg.det(R)), GGT, 4, epsilon_background=12.9,
                         degauss=1e-4, ethr=1e-3)
    assert m.qscr2_debye == pytest.approx(float(g["qscr2_debye"]), rel=2e-6)
    qTF2 = 1 / m.fac2
    assert qTF2 == pytest.approx(float(g["qTF2"]), rel=2e-6)
    assert 3 / (4 * m.fac4 * qTF2) == pytest.approx(float(g["kF2"]), rel=2e-6)
    assert m.prefac_A == pytest.approx(np.sqrt(n * abs(np.linalg.det(R))), rel=1e-15)


def test_bechstedt_limits():
    """Bechstedt's model dielectric function: 1 at large q (bare 4 pi / V q^2)
    and eps_background as q -> 0."""
    m = scatt.eimp_model("Bechstedt", 1e-8, 1e-3, 300.0, np.eye(3), 2, epsilon_background=10.0, degauss=1e-4,
                         ethr=1e-3).as_dict()
    big, small = 1e6, 1e-14
    assert oracle.eimp_vq(m, big) * big / m["prefac_vq_bare"] == pytest.approx(1.0, rel=1e-6)
    assert oracle.eimp_vq(m, small) * small / m["prefac_vq_bare"] == pytest.approx(1 / 10.0, rel=1e-6)
    assert oracle.eimp_vq(m, 0.0) == 0.0  # Gamma skipped


def _system(rng, nk=6, nb=2, nw=4):
    e = np.sort(rng.uniform(0, 2e-3, size=(nk, nb)), axis=1)
    U = rng.normal(size=(nk, nw, nb)) + 1j * rng.normal(size=(nk, nw, nb))
    kv = rng.uniform(-0.5, 0.5, size=(nk, 3))
    k1, k2 = np.triu_indices(nk)
    return e, U, kv, k1.astype(np.uint64), k2.astype(np.uint64)


def _model(screening, scatt_kind, db, nw, T=9.5e-4):
    R = np.diag([10.0, 11.0, 12.0])
    return scatt.eimp_model(screening, 3e-6, T, abs(np.linalg.det(R)), scatt.reciprocal_metric(R), nw, Z=1.5,
                            n_ionized=2e-6, epsilon_background=5.0, degauss=3e-4, ethr=1.5e-3, scatt=scatt_kind,
                            detail_balance=db).as_dict()


@pytest.mark.parametrize("screening", ["unscreened", "debye", "Bechstedt"])
@pytest.mark.parametrize("kind", ["lindblad", "conventional"])
def test_model_P_diagonal_is_the_golden_rule(screening, kind):
    """P1[(n n),(n' n')] = |M_nn'|^2 delta(e_kn - e_k'n') for both scattering
    forms, M = prefac_A v(q) <k n|k' n'>, delta a normalised Gaussian of the e-i
    width cut at ethr; with detailed balance the emission/absorption pair
    (|M|^2/4 or /2) [delta(de + w) exp(-de/T) + delta(de - w)]."""
    rng = np.random.default_rng(3)
    e, U, kv, k1, k2 = _system(rng)
    for db in (False, True):
        m = _model(screening, kind, db, U.shape[1])
        P1, P2 = oracle.eimp_model_P(k1, k2, U, kv, e, m)
        nb = e.shape[1]
        for p, (a, b) in enumerate(zip(k1, k2)):
            o = U[a].conj().T @ U[b]
            vq = oracle.eimp_vq(m, oracle.eimp_q2(m["GGT"], kv[a], kv[b]))
            M2 = np.abs(m["prefac_A"] * vq * o) ** 2
            de = e[a][:, None] - e[b][None, :]
            g = lambda x: np.where(np.abs(x) < m["ethr"], norm.pdf(x, scale=m["degauss"]), 0.0)
            if not db:
                want = M2 * g(de)
            else:
                w, T = m["omegaL"], m["temperature"]
                want = (0.25 if kind == "lindblad" else 0.5) * M2 * (g(de + w) * np.exp(-de / T) + g(de - w))
            diag = P1[p].reshape(nb, nb, nb, nb)[np.arange(nb), np.arange(nb)][:, np.arange(nb), np.arange(nb)]
            assert np.allclose(diag.real, want, rtol=1e-12, atol=1e-300) and np.abs(diag.imag).max() <= 1e-14 * want.max()


@pytest.mark.parametrize("db", [False, True])
def test_model_P_reverse_pair_and_gram_form(db):
    """P2(k, k') = conj(P1(k', k)) (the reverse-direction identity evolve relies
    on, evolve_eph.cpp:101-106), and for the Lindblad form P1 reshaped as
    [(i1 i3),(i2 i4)] is a Hermitian positive semidefinite matrix of rank 1 (2
    with detailed balance): sum of vec(A) vec(A)^H."""
    rng = np.random.default_rng(5)
    e, U, kv, k1, k2 = _system(rng, nb=3, nw=5)
    m = _model("debye", "lindblad", db, U.shape[1])
    P1, P2 = oracle.eimp_model_P(k1, k2, U, kv, e, m)
    P1r, P2r = oracle.eimp_model_P(k2, k1, U, kv, e, m)
    scale = np.abs(P1).max()
    assert np.abs(P2 - P1r.conj()).max() <= 1e-13 * scale and np.abs(P2r - P1.conj()).max() <= 1e-13 * scale
    nb = 3
    for p in range(len(k1)):
        G = P1[p].reshape(nb, nb, nb, nb).transpose(0, 2, 1, 3).reshape(nb * nb, nb * nb)
        assert np.abs(G - G.conj().T).max() <= 1e-14 * max(np.abs(G).max(), 1e-300)
        ev = np.linalg.eigvalsh(0.5 * (G + G.conj().T))
        top = ev[-1]
        if top == 0:
            continue
        assert ev[0] >= -1e-12 * top
        assert np.sum(ev > 1e-9 * top) <= (2 if db else 1)


def _fermi(T, mu, e):
    return 1.0 / (np.exp((e - mu) / T) + 1.0)


def test_rpa_screening_at_q0_against_the_reference_run():
    """The static RPA q_scr^2 at q = 0 of the reference's GaAs e-ph + e-i run
    (first row of its ldbd_vq.out, 2.9821264e-06) from its own k points, e-ph
    conduction bands and mu (4.9841627e-02 Ha, printed to 8 digits), with
    eps = 12.9; at q = 0 the overlaps of orthonormal Wannier rotations are the
    identity.  Its q list has the table's 6807 entries."""
    g = np.load(os.path.join(GOLD, "gaas_vq.npz"))
    d = os.path.join(GOLD, "gaas_ldbd")
    info = ldbd.read_info(d)
    kv = ldbd.read_kvec(d, info)
    _, R = ldbd.read_R(d)
    e = ldbd.read_ek(d, info)[:, info.b_bot_eph:info.b_top_eph]
    nb = e.shape[1]
    f = _fermi(info.temperature, 4.9841627e-02, e)
    U = np.tile(np.eye(nb, dtype=np.complex128), (info.nk, 1, 1))
    V = abs(np.linalg.det(R))
    m = {"kmesh": list(info.kmesh), "temperature": info.temperature, "GGT": scatt.reciprocal_metric(R),
         "prefac_vq": 4 * np.pi / 12.9 / V, "nk_full": info.nk_full}
    k0 = np.zeros(1, np.uint64)
    vq, qs = oracle.eimp_rpa_vq(k0, k0, U, kv, e, f, m)
    assert len(qs) == len(g["q2"]) == 6807
    assert qs[0] == pytest.approx(float(g["qscr2"][0]), rel=1e-6)
    assert vq[0] == 0.0  # Gamma skipped


def test_rpa_screening_limits_and_brute_force():
    """q = 0 with orthonormal rotations: prefac_vq / nk_full sum f(1-f)/T, the
    Debye q^2 times sum f(1-f) / sum f; every q of a mesh with missing points
    (random, non-unitary rotations) against a brute-force double sum over
    (k, k') whose difference falls in the q cell; v(q) of each pair =
    prefac_vq / (|q|^2 + q_scr^2(k - k'))."""
    rng = np.random.default_rng(8)
    km = (4, 3, 2)
    cells = np.array([(a, b, c) for a in range(4) for b in range(3) for c in range(2)])
    keep = np.sort(rng.choice(len(cells), 18, replace=False))
    kv = cells[keep] / np.array(km) + rng.integers(-1, 2, size=(18, 3))  # off by whole reciprocal vectors
    nk, nb, nw, T = 18, 2, 3, 9.5e-4
    e = np.sort(rng.uniform(0, 3e-3, size=(nk, nb)), axis=1)
    f = _fermi(T, 1e-3, e)
    R = np.diag([7.0, 8.0, 9.0])
    m = {"kmesh": list(km), "temperature": T, "GGT": scatt.reciprocal_metric(R), "prefac_vq": 0.37, "nk_full": 24.0}
    Q, _ = np.linalg.qr(rng.normal(size=(nk, nw, nw)) + 1j * rng.normal(size=(nk, nw, nw)))
    k1, k2 = np.triu_indices(nk)
    k1, k2 = k1.astype(np.uint64), k2.astype(np.uint64)
    _, qs = oracle.eimp_rpa_vq(k1, k2, Q[:, :, :nb], kv, e, f, m)
    assert qs[0] == pytest.approx(0.37 / 24 * np.sum(f * (1 - f)) / T, rel=1e-13)
    U = rng.normal(size=(nk, nw, nb)) + 1j * rng.normal(size=(nk, nw, nb))
    vq, qs = oracle.eimp_rpa_vq(k1, k2, U, kv, e, f, m)

    def code(x):
        v = np.rint(np.array(km) * (x - np.floor(x))).astype(int) % np.array(km)
        return (v[..., 0] * km[1] + v[..., 1]) * km[2] + v[..., 2]
    diff = kv[:, None, :] - kv[None, :, :]
    C = code(diff)
    first = {}
    for i in range(nk):
        for j in range(nk):
            first.setdefault(C[i, j], len(first))
    assert len(first) == len(qs)
    brute = np.zeros(len(qs))
    for i in range(nk):
        for j in range(nk):
            o = U[i].conj().T @ U[j]
            de = e[i][:, None] - e[j][None, :]
            dd = np.where(np.abs(de) < 1e-8, 0.25 * (2 - f[i][:, None] - f[j][None, :]) * (f[i][:, None] + f[j][None, :]) / T,
                          (f[j][None, :] - f[i][:, None]) / np.where(np.abs(de) < 1e-8, 1, de))
            brute[first[C[i, j]]] += np.sum(dd * np.abs(o) ** 2)
    brute *= 0.37 / 24
    assert np.allclose(qs, brute, rtol=1e-12, atol=0)
    wrapped = diff - np.floor(0.5 + diff)
    for p in range(0, len(k1), 7):
        i, j = int(k1[p]), int(k2[p])
        q2 = wrapped[i, j] @ m["GGT"] @ wrapped[i, j]
        want = 0.0 if q2 < 1e-20 else 0.37 / (q2 + brute[first[C[i, j]]])
        assert vq[p] == pytest.approx(want, rel=1e-12, abs=0)
