"""Pins of the CPU oracle (oracle/oracle.c) against the mathematics.

The reference holds no golden vectors for this path (no P1/P2 or k-pair
files exist anywhere in it), so each oracle routine is checked against an
independent formulation chosen so that a dropped term, a wrong sign, index or
conjugation fails:
  * evolve     vs the Lindblad operator form (1-rho_k) A rho_k' A^+ - B (1-rho_k') B^+ rho_k
               built with matrix products from the jump operators A, B that
               generate P (write-up Eq. 9-10), reverse direction with A -> B^+,
               B -> A^+; invariants (Hermitian, particle conservation,
               rho = 0 / rho = I zeros, picture covariance);
  * compute_Pt vs t = 0, factorised phases, periodicity;
  * compute_P  vs the P1(k'k) = conj(P2(kk')) identity, Gram positivity and
               ImSigma from the P diagonal against a golden-rule sum;
  * coherent   vs the two-level precession closed form;
  * rk4        vs the stability polynomial on a linear system and O(h^4);
  * dm_Bpert_1st vs its closed form.
"""
import numpy as np
import pytest
import scipy.linalg

from helpers import FS, gaas_spin_fixture, random_P, random_energies, random_pairs, random_rho, rel_err
from oracle import oracle


def lindblad_brute(k1, k2, A, B, e, rho, t, prefac, expt=True):
    """dRho/dt from jump operators: A[p], B[p] are lists of nb x nb matrices."""
    nk, nb = e.shape
    R = rho.reshape(nk, nb, nb)
    I = np.eye(nb)
    C = np.zeros_like(R)
    for p in range(len(k1)):
        k, kp = int(k1[p]), int(k2[p])
        Uk = np.diag(np.exp(1j * e[k] * t)) if expt else I
        Ukp = np.diag(np.exp(1j * e[kp] * t)) if expt else I
        for a in A[p]:
            at = Uk @ a @ Ukp.conj().T
            C[k] += (I - R[k]) @ at @ R[kp] @ at.conj().T
            if k < kp:
                C[kp] -= at.conj().T @ (I - R[k]) @ at @ R[kp]
        for b in B[p]:
            bt = Uk @ b @ Ukp.conj().T
            C[k] -= bt @ (I - R[kp]) @ bt.conj().T @ R[k]
            if k < kp:
                C[kp] += (I - R[kp]) @ bt.conj().T @ R[k] @ bt
    D = prefac * 0.5 * (C + C.conj().transpose(0, 2, 1))
    return D.reshape(nk, nb * nb)


def ops_case(nk, nb, npart, seed, diag=True):
    rng = np.random.default_rng(seed)
    k1, k2 = random_pairs(nk, npart, rng, diag=diag)
    e = random_energies(nk, nb, rng)
    A = [[0.03 * (rng.normal(size=(nb, nb)) + 1j * rng.normal(size=(nb, nb))) for _ in range(2)] for _ in k1]
    B = [[0.03 * (rng.normal(size=(nb, nb)) + 1j * rng.normal(size=(nb, nb))) for _ in range(2)] for _ in k1]
    n4 = nb ** 4
    P1 = np.zeros((len(k1), n4), np.complex128)
    P2 = np.zeros_like(P1)
    for p in range(len(k1)):
        for a in A[p]:
            # P1[(i1 i2),(i3 i4)] = A[i1,i3] conj(A[i2,i4])
            P1[p] += np.einsum("ac,bd->abcd", a, a.conj()).reshape(n4)
        for b in B[p]:
            # P2[(i1 i2),(i3 i4)] = B[i3,i1] conj(B[i4,i2])
            P2[p] += np.einsum("ca,db->abcd", b, b.conj()).reshape(n4)
    rho = random_rho(nk, nb, rng)
    return k1, k2, e, A, B, P1, P2, rho


@pytest.mark.parametrize("nb", [1, 2, 3])
@pytest.mark.parametrize("t_fs", [0.0, 40.0])
def test_evolve_vs_lindblad_operator_form(nb, t_fs):
    k1, k2, e, A, B, P1, P2, rho = ops_case(7, nb, 3, 100 + nb)
    t = t_fs * FS
    got = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=0.7)
    ref = lindblad_brute(k1, k2, A, B, e, rho, t, 0.7)
    assert rel_err(got, ref) < 1e-13


def test_evolve_threads_match_serial():
    k1, k2, e, A, B, P1, P2, rho = ops_case(20, 2, 4, 7)
    a = oracle.evolve(k1, k2, P1, P2, e, rho, t=5 * FS, prefac=1.0, nthreads=1)
    b = oracle.evolve(k1, k2, P1, P2, e, rho, t=5 * FS, prefac=1.0, nthreads=4)
    assert rel_err(b, a) < 1e-14


def test_evolve_invariants():
    k1, k2, e, A, B, P1, P2, rho = ops_case(12, 2, 4, 3, diag=False)
    nk, nb = e.shape
    t = 17 * FS
    d = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=1.0).reshape(nk, nb, nb)
    # Hermitian, exactly: prefac/2 (C + C^+)
    assert np.array_equal(d, d.conj().transpose(0, 2, 1))
    # particle conservation: with the Gram-form P and the reverse trick sum_k Tr ddm_k = 0
    tr = np.trace(d, axis1=1, axis2=2).sum()
    scale = np.abs(np.trace(d, axis1=1, axis2=2)).sum()
    assert abs(tr) < 1e-13 * scale
    z = np.zeros_like(rho)
    assert np.abs(oracle.evolve(k1, k2, P1, P2, e, z, t=t)).max() == 0.0
    one = np.tile(np.eye(nb).reshape(1, nb * nb), (nk, 1)).astype(np.complex128)
    assert np.abs(oracle.evolve(k1, k2, P1, P2, e, one, t=t)).max() == 0.0
    # linear in P
    d2 = oracle.evolve(k1, k2, 2 * P1, 2 * P2, e, rho, t=t, prefac=1.0).reshape(nk, nb, nb)
    assert np.array_equal(d2, 2 * d)


def test_evolve_picture_covariance():
    """Interaction-picture result = rotated Schroedinger-picture result."""
    k1, k2, e, A, B, P1, P2, rho = ops_case(9, 2, 3, 5)
    nk, nb = e.shape
    t = 123 * FS
    U = np.exp(1j * e * t)  # U = diag(e^{i e t})
    R = rho.reshape(nk, nb, nb)
    rhoS = (U.conj()[:, :, None] * R * U[:, None, :]).reshape(nk, nb * nb)  # U^+ rho U
    dI = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=1.0, expt=True).reshape(nk, nb, nb)
    dS = oracle.evolve(k1, k2, P1, P2, e, rhoS, t=0.0, prefac=1.0, expt=False).reshape(nk, nb, nb)
    assert rel_err(dI, U[:, :, None] * dS * U.conj()[:, None, :]) < 1e-13


def test_compute_ddm_pairs_equals_evolve_at_t0():
    k1, k2, e, A, B, P1, P2, rho = ops_case(10, 2, 4, 6)
    a = oracle.compute_ddm_pairs(k1, k2, P1, P2, rho, prefac=0.3)
    b = oracle.evolve(k1, k2, P1, P2, e, rho, t=0.0, prefac=0.3, expt=False)
    assert rel_err(a, b) < 1e-14


def test_ddmdt_eq_subtraction():
    k1, k2, e, A, B, P1, P2, rho = ops_case(8, 2, 3, 8)
    nk, nb = e.shape
    eq = random_rho(nk, nb, np.random.default_rng(1), scale=1e-3)
    t = 50 * FS
    a = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, ddmdt_eq=eq)
    b = oracle.evolve(k1, k2, P1, P2, e, rho, t=t)
    ph = np.exp(1j * (e[:, :, None] - e[:, None, :]) * t).reshape(nk, nb * nb)
    assert rel_err(a, b - eq * ph) < 1e-14
    s = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, ddmdt_eq=eq, schrodinger=True)
    assert rel_err(s, b - eq) < 1e-14


def test_compute_Pt_phases():
    rng = np.random.default_rng(2)
    nb = 3
    ek, ekp = rng.uniform(0, 4e-3, nb), rng.uniform(0, 4e-3, nb)
    P = rng.normal(size=nb ** 4) + 1j * rng.normal(size=nb ** 4)
    assert np.array_equal(oracle.compute_Pt(nb, 0.0, ek, ekp, P, 0), P)
    t = 77 * FS
    u, up = np.exp(1j * ek * t), np.exp(1j * ekp * t)
    f1 = np.einsum("a,b,c,d->abcd", u, u.conj(), up.conj(), up).reshape(-1)
    f2 = np.einsum("a,b,c,d->abcd", up.conj(), up, u, u.conj()).reshape(-1)
    assert rel_err(oracle.compute_Pt(nb, t, ek, ekp, P, 0), P * f1) < 1e-14
    assert rel_err(oracle.compute_Pt(nb, t, ek, ekp, P, 1), P * f2) < 1e-14
    # all phase differences multiples of dE -> periodic in t with 2 pi / dE
    ek2 = np.array([0.0, 1e-3, 2e-3])
    T = 2 * np.pi / 1e-3
    a = oracle.compute_Pt(nb, 0.3 * T, ek2, ek2, P, 0)
    b = oracle.compute_Pt(nb, 1.3 * T, ek2, ek2, P, 0)
    assert rel_err(a, b) < 1e-12


def test_compute_P_reverse_identity_and_gram():
    """P1 of (k',k) equals conj(P2 of (k,k')) when g+_{kk'} = dagger(g-_{k'k})
    (lindbladInit.../ElectronPhonon.h:566-568 comment); P1 is a Gram form."""
    rng = np.random.default_rng(4)
    nb, nm = 2, 3
    Ek, Ekp = rng.uniform(0, 4e-3, nb), rng.uniform(0, 4e-3, nb)
    omega = rng.uniform(2e-4, 2e-3, nm)
    g_kkp = 1e-3 * (rng.normal(size=(nm, nb, nb)) + 1j * rng.normal(size=(nm, nb, nb)))  # g-_{kk'}
    g_kpk = 1e-3 * (rng.normal(size=(nm, nb, nb)) + 1j * rng.normal(size=(nm, nb, nb)))  # g-_{k'k}
    T, sig = 9.5e-4, 1.8e-4
    P1, P2 = oracle.compute_P(Ek, Ekp, omega, g_kkp.reshape(nm, -1),
                              g_kpk.conj().transpose(0, 2, 1).reshape(nm, -1), T, sig)
    Q1, Q2 = oracle.compute_P(Ekp, Ek, omega, g_kpk.reshape(nm, -1),
                              g_kkp.conj().transpose(0, 2, 1).reshape(nm, -1), T, sig)
    assert rel_err(Q1, P2.conj()) < 1e-14 and rel_err(Q2, P1.conj()) < 1e-14
    # Gram form: M[(i1 i3),(i2 i4)] = P1[(i1 i2),(i3 i4)] is Hermitian PSD
    M = P1.reshape(nb, nb, nb, nb).transpose(0, 2, 1, 3).reshape(nb * nb, nb * nb)
    assert np.allclose(M, M.conj().T, rtol=0, atol=1e-15 * np.abs(M).max())
    assert np.linalg.eigvalsh(M).min() > -1e-12 * np.abs(M).max()


def test_compute_P_imsigma_golden_rule():
    """ImSigma_k,b1 from the P diagonal (src_v4.5.7/ElectronPhonon_ana.cpp:788-795 form,
    pi/N_k [P1_(b1b1),(b2b2) F_k'b2 + (1-F_k'b2) P2_(b2b2),(b1b1)] summed over b2)
    equals the golden-rule sum pi/N_k sum |g|^2 delta [(F + n) or (n + 1 - F)] of
    lindbladInit.../ElectronPhonon.h:501-548 (no detailed balance)."""
    rng = np.random.default_rng(5)
    nb, nm = 2, 2
    Ek, Ekp = rng.uniform(0, 4e-3, nb), rng.uniform(0, 4e-3, nb)
    omega = rng.uniform(2e-4, 2e-3, nm)
    gm = 1e-3 * (rng.normal(size=(nm, nb, nb)) + 1j * rng.normal(size=(nm, nb, nb)))
    gp = 1e-3 * (rng.normal(size=(nm, nb, nb)) + 1j * rng.normal(size=(nm, nb, nb)))
    T, sig = 9.5e-4, 1.8e-4
    F = rng.uniform(0, 1, nb)
    P1, P2 = oracle.compute_P(Ek, Ekp, omega, gm.reshape(nm, -1), gp.reshape(nm, -1), T, sig)
    P1 = P1.reshape(nb, nb, nb, nb)
    P2 = P2.reshape(nb, nb, nb, nb)
    delta = lambda x: np.exp(-0.5 * x * x / sig ** 2) / (sig * np.sqrt(2 * np.pi))
    for b1 in range(nb):
        fromP = sum(P1[b1, b1, b2, b2].real * F[b2] + (1 - F[b2]) * P2[b2, b2, b1, b1].real for b2 in range(nb))
        gr = 0.0
        for a in range(nm):
            n = 1 / (np.exp(omega[a] / T) - 1)
            for b2 in range(nb):
                dE = Ek[b1] - Ekp[b2]
                gr += abs(gp[a, b1, b2]) ** 2 * delta(dE + omega[a]) * (F[b2] + n)
                gr += abs(gm[a, b1, b2]) ** 2 * delta(dE - omega[a]) * (n + 1 - F[b2])
        assert abs(fromP - gr) < 1e-12 * abs(gr)


def test_coherent_two_level_precession():
    """-i[H, rho] for H = (D/2) sigma_x, rho = (I + s sigma_z)/2: -(D s / 2) sigma_y."""
    D, s = 2e-4, 0.3
    sx = np.array([[0, 1], [1, 0]], complex)
    sy = np.array([[0, -1j], [1j, 0]])
    sz = np.diag([1.0, -1.0]).astype(complex)
    H = (D / 2 * sx).reshape(1, 4)
    rho = ((np.eye(2) + s * sz) / 2).reshape(1, 4)
    e = np.zeros((1, 2))
    out = oracle.coherent(H, e, rho, 0.0)
    assert rel_err(out, (-(D * s / 2) * sy).reshape(1, 4)) < 1e-15
    # traceless and Hermitian for random inputs, interaction picture phases
    rng = np.random.default_rng(3)
    Hr = random_rho(5, 3, rng, diag=0.0)
    r = random_rho(5, 3, rng)
    er = random_energies(5, 3, rng)
    o = oracle.coherent(Hr, er, r, 40 * FS).reshape(5, 3, 3)
    assert np.abs(np.trace(o, axis1=1, axis2=2)).max() < 1e-17
    assert np.allclose(o, o.conj().transpose(0, 2, 1), atol=1e-18)


def _superop(H):
    """Matrix of rho -> -i[H, rho] acting on row-major vec(rho)."""
    n = H.shape[0]
    I = np.eye(n)
    return -1j * (np.kron(H, I) - np.kron(I, H.T))


def test_rk4_stability_polynomial_and_order():
    """Coherent-only system (no pairs): one step is R(hL) y with
    R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24; global error O(h^4)."""
    D = 3e-3
    H = np.array([[0.2 * D, D / 2], [D / 2, -0.2 * D]], complex)
    rho0 = np.array([[0.6, 0.1 - 0.05j], [0.1 + 0.05j, 0.4]])
    e = np.zeros((1, 2))
    k = np.zeros(0, np.uint64)
    P = np.zeros((0, 16), complex)
    L = _superop(H)
    h = 40.0
    out, _ = oracle.rk4(k, k, P, P, e, rho0.reshape(1, 4), h, 1, H=H.reshape(1, 4))
    z = h * L
    R = np.eye(4) + z + z @ z / 2 + z @ z @ z / 6 + z @ z @ z @ z / 24
    assert rel_err(out.reshape(4), R @ rho0.reshape(4)) < 1e-14
    exact = lambda T: scipy.linalg.expm(T * L) @ rho0.reshape(4)
    errs = []
    for n in (20, 40):
        o, _ = oracle.rk4(k, k, P, P, e, rho0.reshape(1, 4), 2000.0 / n, n, H=H.reshape(1, 4))
        errs.append(np.linalg.norm(o.reshape(4) - exact(2000.0)))
    assert 14 < errs[0] / errs[1] < 18


def test_dm_Bpert_1st_closed_form():
    nb = 2
    e = np.array([[0.0, 1e-3], [5e-4, 5e-4]])  # second k degenerate
    T = 9.5e-4
    f = 1 / (np.exp((e + 1e-3) / T) + 1)
    rng = np.random.default_rng(6)
    s3 = np.stack([random_rho(2, nb, rng, scale=0.5, diag=0.0) for _ in range(3)], axis=1)
    B = np.array([0.0, 0.0, 1e-6])
    out = oracle.dm_Bpert_1st(e, f, s3, B, T, t0=0.0).reshape(2, 2, 2)
    H1 = (B[2] * s3[:, 2]).reshape(2, 2, 2)
    # non-degenerate k: (f_i - f_j)/(e_i - e_j) H1_ij off the diagonal, df/de on it
    k = 0
    assert np.isclose(out[k, 0, 1], (f[k, 0] - f[k, 1]) / (e[k, 0] - e[k, 1]) * H1[k, 0, 1], rtol=1e-14)
    fav = f[k, 0]
    assert np.isclose(out[k, 0, 0], f[k, 0] + fav * (fav - 1) / T * H1[k, 0, 0].real, rtol=1e-14)
    # degenerate k: favg (favg - 1)/T for every element
    k = 1
    fav = 0.5 * (f[k, 0] + f[k, 1])
    assert np.isclose(out[k, 0, 1], fav * (fav - 1) / T * H1[k, 0, 1], rtol=1e-14)
    assert np.allclose(out, out.conj().transpose(0, 2, 1), atol=0)


def test_spin_observable_brute_and_zero_at_f():
    rng = np.random.default_rng(7)
    nk, nb = 6, 3
    e = random_energies(nk, nb, rng)
    f = rng.uniform(0, 1, (nk, nb))
    s = random_rho(nk, nb, rng, scale=0.5, diag=0.0)
    rho = random_rho(nk, nb, rng)
    t = 60 * FS
    got = oracle.spin(s, e, f, rho, t, 1, 3, 0.25)
    S = s.reshape(nk, nb, nb)
    R = rho.reshape(nk, nb, nb) - np.einsum("ki,ij->kij", f, np.eye(nb))
    ph = np.exp(1j * (e[:, :, None] - e[:, None, :]) * t)
    St = S * ph  # s_t[i,b]
    ref = 0.25 * sum(np.real(St[k, i, b] * R[k, b, i]) for k in range(nk) for i in range(1, 3) for b in range(nb))
    assert abs(got - ref) < 1e-14 * abs(ref)
    feq = np.einsum("ki,ij->kij", f, np.eye(nb)).reshape(nk, nb * nb).astype(complex)
    assert oracle.spin(s, e, f, feq, t, 0, nb, 1.0) == 0.0


def test_spin_matches_reference_gaas_run():
    """S_z of the reference's own GaAs e-ph run (Examples/RealTime_Example_GaAs_20230105/
    ld-dmd/t300/mult8_gauss0.005/dBz/only-eph/n2e16): the final density matrix
    (restart/denmat_restart.bin) with that run's S_z matrices, band energies,
    T and mu reproduces the last line of its sz_elec_tot.out, with and without
    the interaction-picture phase.  The printed mu has 8 digits; its rounding
    moves S_z by ~1e-12 relative."""
    e, f, s, dm, i0, i1, t, prefac, sz, sz_amp = gaas_spin_fixture()
    v, a = oracle.spin(s, e, f, dm, t, i0, i1, prefac, with_amp=True)
    assert abs(v / sz - 1) < 5e-12 and abs(a / sz_amp - 1) < 5e-12
    # a planted mistake: the phase sign flipped must miss the reference value
    v_bad = oracle.spin(s, e, f, dm, -t, i0, i1, prefac)
    assert abs(v_bad / sz - 1) > 1e-6
