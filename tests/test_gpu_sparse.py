"""Sparse P on the GPU (alg_sparseP, alg_Pin_is_sparse; ElectronPhonon.cpp:
115-129, 235-246): the sparse files densified into every layout with the
reader's scale, threshold and last-value rule, bit for bit; evaluation against
the oracle's sparse evolve; the sparsification of stored P in every layout and
precision; the writer's files (row-major elements above the threshold, no
overwriting, hole names) and their round trip; the sparse set-up against the
dense one."""
import numpy as np
import pytest

from helpers import FS, quantise_P, random_P, random_energies, random_pairs, random_rho, rel_err, sparse_lists, \
    sparsify, write_ldbd_folder, write_sP_files

pytestmark = pytest.mark.gpu


def _mods():
    from denmat_dynm_development_amd import _lib, ldbd, scatt
    from oracle import oracle
    return _lib, ldbd, scatt, oracle


def _mag(P, rng):
    """P with magnitudes spread over three decades, so thresholds bite."""
    return P * 10.0 ** rng.uniform(-3, 0, P.shape)


def _keep(P, thr):
    return np.where(np.sqrt(P.real * P.real + P.imag * P.imag) > thr, P, 0)


def _case(nb, seed, nk=None):
    rng = np.random.default_rng(seed)
    nk = nk or (120 if nb <= 2 else 40)
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    return rng, nk, k1, k2, e, sparsify(_mag(P1, rng), rng, 0.5), sparsify(_mag(P2, rng), rng, 0.3)


@pytest.mark.parametrize("nb", [1, 2, 3, 4])
def test_load_sP_densifies_scales_thresholds_and_keeps_the_last(nb, tmp_path):
    _lib, ldbd, scatt, oracle = _mods()
    print(f"case load_sP nb={nb}", flush=True)
    rng, nk, k1, k2, e, P1, P2 = _case(nb, 110 + nb)
    d = tmp_path / "ldbd_data"
    write_ldbd_folder(d, 0, nb, 0, (0, nb), (0, nb), e, k1, k2)
    write_sP_files(d, 1, sparse_lists(P1, rng, dup_frac=0.3))
    write_sP_files(d, 2, sparse_lists(P2, rng, dup_frac=0.3))
    scale = 0.7
    nz = np.abs(scale * np.concatenate([P1[P1 != 0], P2[P2 != 0]]))
    thr = float(np.quantile(nz, 0.3))
    E1, E2 = _keep(scale * P1, thr), _keep(scale * P2, thr)
    assert 0 < np.count_nonzero(E1) < np.count_nonzero(P1)
    prefac = 2 * np.pi / nk
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.load_sP(d, scale=scale, thr=thr)
        A1, A2 = op.download_P()
        assert np.array_equal(A1, E1) and np.array_equal(A2, E2)
        for t_fs in (0.0, 29.0):
            rho = random_rho(nk, nb, rng)
            ref = oracle.evolve_sparse(k1, k2, sparse_lists(E1, rng), sparse_lists(E2, rng), e, rho, t=t_fs * FS,
                                       prefac=prefac)
            assert rel_err(op.eval(t_fs * FS, rho), ref) < 1e-12
    if nb == 4:
        for prec in (_lib.DMD_PREC_C64, _lib.DMD_PREC_BF16):
            with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=prec) as op:
                op.load_sP(d, scale=scale, thr=thr)
                A1, A2 = op.download_P()
                Q1, Q2 = quantise_P(E1, E2, prec)
                assert np.array_equal(A1, Q1) and np.array_equal(A2, Q2), prec


def test_load_sP_rejects_bad_files(tmp_path):
    _lib, ldbd, scatt, oracle = _mods()
    print("case load_sP errors", flush=True)
    rng, nk, k1, k2, e, P1, P2 = _case(2, 120)
    d = tmp_path / "ldbd_data"
    write_ldbd_folder(d, 0, 2, 0, (0, 2), (0, 2), e, k1, k2)
    s1, s2 = sparse_lists(P1), sparse_lists(P2)
    with scatt.ScatteringOperator(nk, 2, k1, k2, e) as op:
        op.upload_P(P1, P2)
        write_sP_files(d, 2, s2)
        for bad, what in (((s1[0], s1[1][:-1], s1[2], s1[3]), "size"),
                          ((s1[0][:-1], s1[1], s1[2], s1[3]), "size"),
                          ((s1[0], s1[1], s1[2] + 3, s1[3]), "index")):
            write_sP_files(d, 1, bad)
            with pytest.raises(_lib.DMDError, match=what):
                op.load_sP(d)
        # the size checks leave the uploaded P alone; the index failure comes
        # inside the chunk loop, after which no half-loaded P is used
        with pytest.raises(_lib.DMDError, match="P has not been"):
            op.eval(0.0, random_rho(nk, 2, rng))
        write_sP_files(d, 1, s1)
        op.load_sP(d)
        A1, _ = op.download_P()
        assert np.array_equal(A1, P1)


@pytest.mark.parametrize("nb,prec", [(1, 0), (2, 0), (3, 0), (4, 0), (4, 1), (4, 2)])
def test_threshold_P_in_every_layout(nb, prec):
    _lib, ldbd, scatt, oracle = _mods()
    print(f"case threshold nb={nb} precision={prec}", flush=True)
    rng = np.random.default_rng(130 + 10 * nb + prec)
    nk = 150 if nb <= 2 else 50
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    P1, P2 = _mag(P1, rng), _mag(P2, rng)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=prec) as op:
        op.upload_P(P1, P2)
        B1, B2 = op.download_P()
        thr = float(np.quantile(np.abs(B1), 0.4))
        op.threshold_P(thr)
        A1, A2 = op.download_P()
        assert np.array_equal(A1, _keep(B1, thr)) and np.array_equal(A2, _keep(B2, thr))
        assert 0 < np.count_nonzero(A1) < A1.size
        if prec == 0:
            rho = random_rho(nk, nb, rng)
            ref = oracle.evolve(k1, k2, _keep(P1, thr), _keep(P2, thr), e, rho, t=11 * FS, prefac=2 * np.pi / nk)
            assert rel_err(op.eval(11 * FS, rho), ref) < 1e-12


@pytest.mark.parametrize("kind", [0, 1])
def test_write_sP_files_and_round_trip(kind, tmp_path):
    _lib, ldbd, scatt, oracle = _mods()
    print(f"case write_sP kind={kind}", flush=True)
    nb = 2 if kind == 1 else 4
    rng, nk, k1, k2, e, P1, P2 = _case(nb, 140 + kind)
    d = tmp_path / "ldbd_data"
    if kind == 1:
        write_ldbd_folder(d, 1, nb + 1, nb + 1, (0, nb + 1), (1, nb + 1), np.c_[e[:, :1] - 1e-3, e], k1, k2)
    else:
        write_ldbd_folder(d, 0, nb, 0, (0, nb), (0, nb), e, k1, k2)
    info = ldbd.read_info(d)
    thr = float(np.quantile(np.abs(P1[P1 != 0]), 0.25))
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        op.write_sP(d, info, thr=thr)
        for which, P in ((1, P1), (2, P2)):
            got = ldbd.read_sP(d, which, info)
            want = sparse_lists(_keep(P, thr))  # row-major order, values as stored
            for a, b in zip(got, want):
                assert np.array_equal(a, b), which
        assert (d / ("sP1_lindblad_hole_i.bin" if kind == 1 else "sP1_lindblad_i.bin")).exists()
        # an existing _s file is kept
        before = (d / ("sP1_lindblad_hole_s.bin" if kind == 1 else "sP1_lindblad_s.bin")).read_bytes()
        op.upload_P(P2, P1)
        op.write_sP(d, info, thr=0.0)
        assert (d / ("sP1_lindblad_hole_s.bin" if kind == 1 else "sP1_lindblad_s.bin")).read_bytes() == before
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.load_sP(d, info)
        A1, A2 = op.download_P()
        assert np.array_equal(A1, _keep(P1, thr)) and np.array_equal(A2, _keep(P2, thr))


def test_from_ldbd_sparse_matches_dense(tmp_path):
    _lib, ldbd, scatt, oracle = _mods()
    print("case from_ldbd sparse", flush=True)
    p = scatt.synth_params(0)
    d = tmp_path / "ldbd_data"
    d.mkdir()
    scatt.write_ldbd(str(d), p)
    info = ldbd.read_info(d)
    R1, R2 = ldbd.read_P(d, info)
    write_sP_files(d, 1, sparse_lists(R1))
    write_sP_files(d, 2, sparse_lists(R2))
    rng = np.random.default_rng(150)
    with scatt.ScatteringOperator.from_ldbd(d, scale_eph=0.9) as dense, \
            scatt.ScatteringOperator.from_ldbd(d, scale_eph=0.9, Pin_is_sparse=True) as sp:
        D1, D2 = dense.download_P()
        S1, S2 = sp.download_P()
        assert np.array_equal(D1, S1) and np.array_equal(D2, S2)
        rho = random_rho(dense.nk, dense.nb, rng)
        assert rel_err(sp.eval(13 * FS, rho), dense.eval(13 * FS, rho)) < 1e-14
