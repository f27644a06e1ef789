"""Parity of the HIP scattering operator (through the C-ABI) with the CPU oracle.

Tolerance: FP64, norm-wise relative error <= 1e-12 (the parity bar for ddm
against the reference's CPU compute_ddm); the kernel sums pairs in a different order than the reference loop and
applies the interaction-picture phase in factorised form, so results agree to
a few ulp, not bit for bit.
"""
import os
import sys

import numpy as np
import pytest

from helpers import FS, gaas_spin_fixture, random_P, random_energies, random_pairs, random_rho, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _mods():
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    return scatt, oracle


def _case(nk, nb, npart, seed, diag=False, structured=True):
    rng = np.random.default_rng(seed)
    k1, k2 = random_pairs(nk, npart, rng, diag=diag)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng, structured=structured)
    rho = random_rho(nk, nb, rng)
    return rng, k1, k2, e, P1, P2, rho


KERNELS = [1, 2, 3]  # DMD_KERNEL_ROWS (deterministic), _SINGLE_READ, _TILES


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("nb", [1, 2, 3, 4])
@pytest.mark.parametrize("t_fs", [0.0, 37.0, 2500.0])
def test_eval_matches_oracle(nb, t_fs, kernel):
    scatt, oracle = _mods()
    print(f"case eval nb={nb} t={t_fs}fs kernel={kernel}", flush=True)
    rng, k1, k2, e, P1, P2, rho = _case(40, nb, 5, 10 + nb, diag=True)
    t = t_fs * FS
    prefac = 2 * np.pi / 40
    ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=prefac)
    with scatt.ScatteringOperator(40, nb, k1, k2, e, nk_full=40) as op:
        op.upload_P(P1, P2)
        op.set_flags(kernel=kernel)
        got = op.eval(t, rho)
        again = op.eval(t, rho)  # the accumulator is re-zeroed by every evaluation
    assert rel_err(got, ref) < TOL
    assert rel_err(again, got) < 1e-14


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("nb", [1, 2, 4])
def test_eval_many_tiles_unsorted(kernel, nb):
    """nk spans several row tiles (512 rows at nb <= 2, 32 at nb = 4), pairs
    given in shuffled order (exercises the tile layout and the index lists)."""
    scatt, oracle = _mods()
    print(f"case tiles nb={nb} kernel={kernel}", flush=True)
    nk = 700 if nb < 4 else 150
    rng, k1, k2, e, P1, P2, rho = _case(nk, nb, 9, 40 + nb, diag=True)
    perm = rng.permutation(len(k1))
    k1, k2, P1, P2 = k1[perm], k2[perm], P1[perm], P2[perm]
    t = 211 * FS
    ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=2 * np.pi / nk)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        op.set_flags(kernel=kernel)
        got = op.eval(t, rho)
        D1, D2 = op.download_P()
    assert np.array_equal(D1, P1) and np.array_equal(D2, P2)
    assert rel_err(got, ref) < TOL


@pytest.mark.parametrize("kernel", KERNELS)
def test_eval_edge_cases(kernel):
    """rows without pairs, a k with only reverse pairs, diagonal-only pairs, one pair."""
    scatt, oracle = _mods()
    rng = np.random.default_rng(31)
    nb, nk = 2, 9
    for k1, k2 in ([[2], [7]], [[0, 3, 3, 8], [0, 3, 5, 8]], [[0, 1, 1, 4], [8, 2, 8, 8]]):
        k1 = np.array(k1, np.uint64)
        k2 = np.array(k2, np.uint64)
        e = random_energies(nk, nb, rng)
        P1, P2 = random_P(len(k1), nb, rng)
        rho = random_rho(nk, nb, rng)
        ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=9 * FS, prefac=0.5)
        with scatt.ScatteringOperator(nk, nb, k1, k2, e, nk_full=4 * np.pi) as op:
            op.upload_P(P1, P2)
            op.set_flags(kernel=kernel)
            got = op.eval(9 * FS, rho)
        assert rel_err(got, ref) < TOL
        empty = np.abs(ref).sum(axis=1) == 0
        assert np.all(got[empty] == 0)


@pytest.mark.parametrize("nb", [2, 4])
@pytest.mark.parametrize("kernel", KERNELS)
def test_eval_unstructured_explicit_dm1_eq_unsorted(nb, kernel):
    scatt, oracle = _mods()
    print(f"case eval-variants nb={nb}", flush=True)
    rng, k1, k2, e, P1, P2, rho = _case(33, nb, 4, 20 + nb, diag=True, structured=False)
    perm = rng.permutation(len(k1))
    k1, k2, P1, P2 = k1[perm], k2[perm], P1[perm], P2[perm]
    dm1 = oracle.one_minus(rho, nb) + 1e-3 * random_rho(33, nb, rng, diag=0.0)
    eq = random_rho(33, nb, rng, scale=1e-3, diag=1e-4)
    t = 123.0 * FS
    prefac = 2 * np.pi / 50.0
    with scatt.ScatteringOperator(33, nb, k1, k2, e, nk_full=50.0) as op:
        op.upload_P(P1, P2)
        for expt in (True, False):
            for use_eq in (False, True):
                op.set_flags(expt=expt, kernel=kernel)
                op.set_ddmdt_eq(eq if use_eq else None)
                ref = oracle.evolve(k1, k2, P1, P2, e, rho, dm1, t=t, prefac=prefac, expt=expt,
                                    ddmdt_eq=eq if use_eq else None)
                got = op.eval(t, rho, dm1)
                assert rel_err(got, ref) < TOL, (expt, use_eq)
        # Schroedinger picture: eq subtracted without phase, no P phases
        op.set_flags(expt=False, schrodinger=True, kernel=kernel)
        op.set_ddmdt_eq(eq)
        ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=prefac, expt=False, ddmdt_eq=eq, schrodinger=True)
        assert rel_err(op.eval(t, rho), ref) < TOL


def test_upload_add_download_roundtrip():
    scatt, oracle = _mods()
    rng, k1, k2, e, P1, P2, rho = _case(30, 2, 4, 5)
    perm = rng.permutation(len(k1))
    k1, k2, P1, P2 = k1[perm], k2[perm], P1[perm], P2[perm]
    Q1, Q2 = random_P(len(k1), 2, rng)
    with scatt.ScatteringOperator(30, 2, k1, k2, e) as op:
        op.upload_P(P1, P2, 0.5)
        op.add_P(Q1, Q2, 2.0)
        D1, D2 = op.download_P()
        np.testing.assert_allclose(D1, 0.5 * P1 + 2.0 * Q1, rtol=1e-15, atol=1e-18)
        np.testing.assert_allclose(D2, 0.5 * P2 + 2.0 * Q2, rtol=1e-15, atol=1e-18)
        d1, d2 = op.download_P(3, 5)
        np.testing.assert_array_equal(d1, D1[3:8])
        np.testing.assert_array_equal(d2, D2[3:8])


def test_compute_ddm_matches_initializer_form():
    scatt, oracle = _mods()
    rng, k1, k2, e, P1, P2, rho = _case(25, 2, 4, 7, diag=True)
    prefac = 2 * np.pi / 25
    base = random_rho(25, 2, rng, scale=1e-3, diag=0)
    ref = oracle.compute_ddm_pairs(k1, k2, P1, P2, rho, prefac=prefac) + base
    with scatt.ScatteringOperator(25, 2, k1, k2, e, nk_full=25) as op:
        op.upload_P(P1, P2)
        got = op.compute_ddm(rho, ddm=base.copy())
    assert rel_err(got, ref) < TOL


def test_compute_ddmdt_eq():
    scatt, oracle = _mods()
    rng, k1, k2, e, P1, P2, rho = _case(20, 2, 4, 8)
    f = rng.uniform(0, 1, (20, 2))
    dm = np.zeros((20, 4), np.complex128)
    dm[:, 0], dm[:, 3] = f[:, 0], f[:, 1]
    ref_eq = oracle.evolve(k1, k2, P1, P2, e, dm, t=0.0, prefac=2 * np.pi / 20)
    with scatt.ScatteringOperator(20, 2, k1, k2, e) as op:
        op.upload_P(P1, P2)
        op.compute_ddmdt_eq(f)
        t = 77 * FS
        got = op.eval(t, rho)
    ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=2 * np.pi / 20, ddmdt_eq=ref_eq)
    assert rel_err(got, ref) < TOL


def test_evolve_driver_band_window():
    scatt, oracle = _mods()
    nk, nb, nbd, b0 = 24, 2, 5, 2
    rng = np.random.default_rng(9)
    k1, k2 = random_pairs(nk, 4, rng)
    e_dm = random_energies(nk, nbd, rng)
    e = e_dm[:, b0:b0 + nb]
    P1, P2 = random_P(len(k1), nb, rng)
    dm = random_rho(nk, nbd, rng)
    t = 333 * FS
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        op.set_window(nbd, b0, e_dm)
        got = op.evolve_driver(t, dm)
    w = dm.reshape(nk, nbd, nbd)[:, b0:b0 + nb, b0:b0 + nb].reshape(nk, nb * nb)
    ref_w = oracle.evolve(k1, k2, P1, P2, e, w, t=t, prefac=2 * np.pi / nk)
    ref = np.zeros((nk, nbd, nbd), np.complex128)
    ref[:, b0:b0 + nb, b0:b0 + nb] = ref_w.reshape(nk, nb, nb)
    assert rel_err(got, ref.reshape(nk, nbd * nbd)) < TOL


def test_torch_device_tensors_in_place():
    import torch
    scatt, oracle = _mods()
    rng, k1, k2, e, P1, P2, rho = _case(30, 2, 5, 11)
    t = 10 * FS
    ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=2 * np.pi / 30)
    with scatt.ScatteringOperator(30, 2, k1, k2, e) as op:
        op.upload_P(torch.from_numpy(P1).cuda(), torch.from_numpy(P2).cuda())
        out = op.eval(t, torch.from_numpy(rho).cuda())
        assert out.is_cuda
    assert rel_err(out.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("with_H", [False, True])
def test_rk4_trajectory_matches_oracle(with_H):
    scatt, oracle = _mods()
    print(f"case rk4 with_H={with_H}", flush=True)
    nk, nb = 30, 2
    rng = np.random.default_rng(12)
    k1, k2 = random_pairs(nk, 5, rng)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    rho0 = random_rho(nk, nb, rng)
    s = random_rho(nk, nb, rng, scale=0.5, diag=0.0)
    f = rng.uniform(0.2, 0.4, (nk, nb))
    H = random_rho(nk, nb, rng, scale=2e-4, diag=0.0) if with_H else None
    dt, n, rep = 2 * FS, 24, 4
    ref, ref_obs = oracle.rk4(k1, k2, P1, P2, e, rho0, dt, n, prefac=2 * np.pi / nk, H=H, report_every=rep,
                              s_obs=s, f_dm=f, i0=0, i1=nb, obs_prefac=1.7)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        if H is not None:
            op.set_coherent(H)
        op.set_observable(s, f, 0, nb, 1.7)
        got, obs = op.integrate_rk4(rho0, 0.0, dt, n, report_every=rep)
        m0 = op.measure(0.0, rho0)
    assert rel_err(got, ref) < 1e-12
    assert len(obs) == len(ref_obs)
    np.testing.assert_allclose(obs, ref_obs, rtol=1e-10, atol=1e-14 * np.abs(ref_obs).max())
    assert abs(m0 - ref_obs[0]) <= 1e-12 * abs(ref_obs[0])


def test_rk4_graph_replays_match_eager(monkeypatch):
    """Graph replays equal the eager stream: bit for bit with the deterministic
    row kernel, to rounding with the single-read kernel (atomic order)."""
    scatt, oracle = _mods()
    nk, nb = 26, 2
    rng = np.random.default_rng(13)
    k1, k2 = random_pairs(nk, 5, rng)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    rho0 = random_rho(nk, nb, rng)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        for kernel, tol in ((1, 0.0), (2, 1e-13), (3, 1e-13)):
            op.set_flags(kernel=kernel)
            monkeypatch.setenv("DMD_GRAPH", "0")
            eager, _ = op.integrate_rk4(rho0, 0.0, FS, 12)
            monkeypatch.setenv("DMD_GRAPH", "1")
            for _ in range(3):  # several replays, eager work in between
                g, _ = op.integrate_rk4(rho0, 0.0, FS, 12)
                assert rel_err(g, eager) <= tol
                op.eval(0.0, rho0)


def test_window_rk4_with_coherent_matches_oracle():
    scatt, oracle = _mods()
    nk, nb, nbd, b0 = 18, 2, 4, 1
    rng = np.random.default_rng(14)
    k1, k2 = random_pairs(nk, 4, rng)
    e_dm = random_energies(nk, nbd, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    P1, P2 = 0.1 * P1, 0.1 * P2  # keeps dt * rate well inside RK4's stability region
    rho0 = random_rho(nk, nbd, rng)
    H = random_rho(nk, nbd, rng, scale=1e-4, diag=0.0)
    s = random_rho(nk, nbd, rng, scale=0.5, diag=0.0)
    f = rng.uniform(0.1, 0.3, (nk, nbd))
    dt, n = 3 * FS, 10
    ref, ref_obs = oracle.rk4(k1, k2, P1, P2, e_dm, rho0, dt, n, b0=b0, nb=nb, prefac=2 * np.pi / nk, H=H,
                              report_every=5, s_obs=s, f_dm=f, i0=1, i1=3, obs_prefac=0.5)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e_dm[:, b0:b0 + nb]) as op:
        op.upload_P(P1, P2)
        op.set_window(nbd, b0, e_dm)
        op.set_coherent(H)
        op.set_observable(s, f, 1, 3, 0.5)
        got, obs = op.integrate_rk4(rho0, 0.0, dt, n, report_every=5)
    assert rel_err(got, ref) < 1e-12
    np.testing.assert_allclose(obs, ref_obs, rtol=1e-10, atol=1e-14 * np.abs(ref_obs).max())


def test_generate_on_device_matches_host_generator():
    scatt, oracle = _mods()
    p = scatt.synth_params(0)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    P1, P2 = scatt.synth_P(p, e, k1, k2)
    with scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e) as op:
        op.generate_P(p)
        D1, D2 = op.download_P()
    # same formula on host and device; transcendental functions may differ by an ulp
    assert rel_err(D1, P1) < 1e-13 and rel_err(D2, P2) < 1e-13


def test_kernel_switch_relayouts_P_in_place():
    """Switching between the tile kernel (SoA granules) and the row kernels
    (pair-contiguous P) transposes P on the device; results and the
    downloaded P stay the same through several switches."""
    scatt, oracle = _mods()
    rng, k1, k2, e, P1, P2, rho = _case(600, 2, 7, 77, diag=True)
    t = 55 * FS
    ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=2 * np.pi / 600)
    with scatt.ScatteringOperator(600, 2, k1, k2, e) as op:
        op.upload_P(P1, P2)
        for kernel in (3, 1, 2, 3, 0, 1, 3):
            op.set_flags(kernel=kernel)
            assert rel_err(op.eval(t, rho), ref) < TOL, kernel
            D1, D2 = op.download_P()
            assert np.array_equal(D1, P1) and np.array_equal(D2, P2), kernel


def test_generate_P_in_both_layouts():
    scatt, oracle = _mods()
    p = scatt.synth_params(0)
    p.nk, p.partners = 600, 6
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    out = []
    for kernel in (3, 1):
        with scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e) as op:
            op.set_flags(kernel=kernel)
            op.generate_P(p)
            out.append(op.download_P())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    P1, P2 = scatt.synth_P(p, e, k1, k2)
    assert rel_err(out[0][0], P1) < 1e-13 and rel_err(out[0][1], P2) < 1e-13


@pytest.mark.parametrize("kernel", KERNELS)
def test_no_pairs_gives_minus_eq_only(kernel):
    """An operator without stored pairs: ddmdt is zero, or -ddmdt_eq (rotated)."""
    scatt, oracle = _mods()
    rng = np.random.default_rng(3)
    nk, nb = 300, 2
    k = np.zeros(0, np.uint64)
    e = random_energies(nk, nb, rng)
    rho = random_rho(nk, nb, rng)
    eq = random_rho(nk, nb, rng, scale=1e-3)
    P = np.zeros((0, nb ** 4), np.complex128)
    with scatt.ScatteringOperator(nk, nb, k, k, e) as op:
        op.upload_P(P, P)
        op.set_flags(kernel=kernel)
        assert np.abs(op.eval(7 * FS, rho)).max() == 0.0
        op.set_ddmdt_eq(eq)
        got = op.eval(7 * FS, rho)
    ref = oracle.evolve(k, k, P, P, e, rho, t=7 * FS, prefac=2 * np.pi / nk, ddmdt_eq=eq)
    assert rel_err(got, ref) < TOL


def test_spin_observable_on_reference_gaas_state():
    """dmd_scatt_measure on the reference GaAs run's final state reproduces the
    last S_z of its sz_elec_tot.out (the CPU twin of this check pins the oracle)."""
    scatt, oracle = _mods()
    e, f, s, dm, i0, i1, t, prefac, sz, sz_amp = gaas_spin_fixture()
    nk, nbd = e.shape
    k1 = np.zeros(0, np.uint64)
    with scatt.ScatteringOperator(nk, i1 - i0, k1, k1, e[:, i0:i1]) as op:
        op.upload_P(np.zeros((0, (i1 - i0) ** 4), np.complex128), np.zeros((0, (i1 - i0) ** 4), np.complex128))
        op.set_window(nbd, i0, e)
        op.set_observable(s, f, i0, i1, prefac)
        v = op.measure(t, dm)
    assert abs(v / sz - 1) < 5e-12
