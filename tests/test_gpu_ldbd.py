"""Handles built from ldbd_data folders (dmd_scatt_create_from_ldbd and the
streaming P loader dmd_scatt_load_P) against handles built from the same
arrays in memory (bit for bit, deterministic row kernel), against the files'
own P (bit for bit through several 64 MiB chunks), and against the oracle for
the band window, scale_eph and the e-i fold-in of the _D1 files."""
import time

import numpy as np
import pytest

from helpers import FS, random_P, random_pairs, random_rho, rel_err, write_ldbd_folder

pytestmark = pytest.mark.gpu


def _mods():
    from denmat_dynm_development_amd import _lib, ldbd, scatt
    from oracle import oracle
    return _lib, ldbd, scatt, oracle


def test_from_ldbd_matches_the_in_memory_handle(tmp_path):
    _lib, ldbd, scatt, oracle = _mods()
    print("case from_ldbd preset 0", flush=True)
    p = scatt.synth_params(0)
    scatt.write_ldbd(str(tmp_path), p)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    P1, P2 = scatt.synth_P(p, e, k1, k2)
    rho = random_rho(p.nk, p.nb, np.random.default_rng(1))
    t = 37 * FS
    with scatt.ScatteringOperator.from_ldbd(tmp_path) as a, \
            scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk) as b:
        b.upload_P(P1, P2)
        for op in (a, b):
            op.set_flags(kernel=_lib.DMD_KERNEL_ROWS)
        ra, rb = a.eval(t, rho), b.eval(t, rho)
        assert np.array_equal(ra, rb)
        Q1, Q2 = a.download_P()
        assert np.array_equal(Q1, P1) and np.array_equal(Q2, P2)
    ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t, prefac=2 * np.pi / p.nk)
    assert rel_err(ra, ref) < 1e-12


def test_load_P_streams_chunks_bit_exact(tmp_path):
    """About 1 M stored pairs of nb = 2 (256 MB per P file): four 64 MiB chunks
    through the pinned double buffers."""
    _lib, ldbd, scatt, oracle = _mods()
    print("case load_P preset 1", flush=True)
    p = scatt.synth_params(1)
    scatt.write_ldbd(str(tmp_path), p)
    info = ldbd.read_info(tmp_path)
    t0 = time.perf_counter()
    op = scatt.ScatteringOperator.from_ldbd(tmp_path)
    el = time.perf_counter() - t0
    try:
        nbytes = 2 * info.nkpair * p.nb ** 4 * 16
        print(f"create_from_ldbd: {info.nkpair} pairs, {nbytes / 1e9:.2f} GB of P in {el:.2f} s "
              f"({nbytes / el / 1e9:.2f} GB/s, page cache)")
        chunk = (64 << 20) // (p.nb ** 4 * 16)
        assert info.nkpair > 3 * chunk
        P1, P2 = ldbd.read_P(tmp_path, info)
        Q1, Q2 = op.download_P()
        assert np.array_equal(Q1, P1) and np.array_equal(Q2, P2)
        for first in (0, chunk - 2, 2 * chunk - 1, info.nkpair - 3):
            R1, R2 = op.download_P(first, 3)
            assert np.array_equal(R1, P1[first:first + 3]) and np.array_equal(R2, P2[first:first + 3])
        # the in-memory upload of the same arrays gives the same device layout
        k1, k2 = ldbd.read_kpairs(tmp_path, info)
        e = ldbd.read_ek(tmp_path, info)
        rho = random_rho(p.nk, p.nb, np.random.default_rng(2))
        with scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=info.nk_full) as b:
            b.upload_P(P1, P2)
            for o in (op, b):
                o.set_flags(kernel=_lib.DMD_KERNEL_ROWS)
            assert np.array_equal(op.eval(11 * FS, rho), b.eval(11 * FS, rho))
    finally:
        op.close()


@pytest.mark.parametrize("kind", [1, 2])
def test_window_scale_and_eimp_from_files(tmp_path, kind):
    """nb = 6 file bands, density matrix on [1, 5), e-ph on [2, 4) (valence-only
    files for kind 1, with the _hole names); P x scale_eph from the files and
    the _D1 e-i P added with its fraction, through evolve_driver."""
    _lib, ldbd, scatt, oracle = _mods()
    print(f"case window/eimp kind={kind}", flush=True)
    rng = np.random.default_rng(10 + kind)
    nk, nbf, nv = 40, 6, (5 if kind == 1 else 1)
    dmw, ephw = (1, 5), (2, 4)
    e = np.sort(rng.uniform(-0.02, 0.06, (nk, nbf)), axis=1)
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    P1, P2 = random_P(len(k1), 2, rng)
    Q1, Q2 = random_P(len(k1), 2, rng)
    write_ldbd_folder(str(tmp_path), kind, nbf, nv, dmw, ephw, e, k1, k2, P1, P2, nk_full=1000.0,
                      extra_P=[(1, Q1, Q2)])
    se, frac = 0.8, 0.3
    nbd, b0 = dmw[1] - dmw[0], ephw[0] - dmw[0]
    e_dm, e_eph = e[:, dmw[0]:dmw[1]], e[:, ephw[0]:ephw[1]]
    A = random_rho(nk, nbd, rng).reshape(nk, nbd, nbd)
    dm = A.reshape(nk, nbd * nbd)
    t = 53 * FS
    with scatt.ScatteringOperator.from_ldbd(tmp_path, scale_eph=se) as a, \
            scatt.ScatteringOperator(nk, 2, k1, k2, e_eph, nk_full=1000.0) as b:
        assert (a.nb, a.nb_dm, a.b0) == (2, nbd, b0)
        a.load_P(tmp_path, impurity=1, scale=frac, add=True)
        b.set_window(nbd, b0, e_dm)
        b.upload_P(P1, P2, se)
        b.add_P(Q1, Q2, frac)
        for op in (a, b):
            op.set_flags(kernel=_lib.DMD_KERNEL_ROWS)
        ra, rb = a.evolve_driver(t, dm), b.evolve_driver(t, dm)
        assert np.array_equal(ra, rb)
    cut = A[:, b0:b0 + 2, b0:b0 + 2].reshape(nk, 4)
    inner = oracle.evolve(k1, k2, se * P1 + frac * Q1, se * P2 + frac * Q2, e_eph, cut, t=t,
                          prefac=2 * np.pi / 1000.0)
    ref = np.zeros((nk, nbd, nbd), np.complex128)
    ref[:, b0:b0 + 2, b0:b0 + 2] = inner.reshape(nk, 2, 2)
    assert rel_err(ra, ref.reshape(nk, -1)) < 1e-12


@pytest.mark.parametrize("prec", [1, 2])
def test_reduced_precision_handles_from_files(tmp_path, prec):
    """nb = 4 files into complex64 / BF16 handles: the stored P equals the
    in-memory upload's, and the evaluation agrees."""
    _lib, ldbd, scatt, oracle = _mods()
    print(f"case reduced from files prec={prec}", flush=True)
    p = scatt.synth_params(3)
    p.nk, p.partners, p.partners_max = 48, 6, 0
    scatt.write_ldbd(str(tmp_path), p)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    P1, P2 = scatt.synth_P(p, e, k1, k2)
    rho = random_rho(p.nk, p.nb, np.random.default_rng(3))
    with scatt.ScatteringOperator.from_ldbd(tmp_path, precision=prec) as a, \
            scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk, precision=prec) as b:
        b.upload_P(P1, P2)
        A1, A2 = a.download_P()
        B1, B2 = b.download_P()
        assert np.array_equal(A1, B1) and np.array_equal(A2, B2)
        assert rel_err(a.eval(9 * FS, rho), b.eval(9 * FS, rho)) < 1e-13


def test_handles_without_eph_P(tmp_path):
    """from_ldbd(eph_enable=False): no P is read and P counts as unset until the
    e-i term sets it; from_ldbd(scatt_enable=False): no stored pairs, ddm = 0."""
    _lib, ldbd, scatt, oracle = _mods()
    print("case handles without e-ph P", flush=True)
    rng = np.random.default_rng(30)
    nk, nbf = 40, 2
    e = np.sort(rng.uniform(0.0, 0.01, (nk, nbf)), axis=1)
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    Q1, Q2 = random_P(len(k1), nbf, rng)
    write_ldbd_folder(str(tmp_path), 0, nbf, 0, (0, nbf), (0, nbf), e, k1, k2, nk_full=1000.0, extra_P=[(1, Q1, Q2)])
    rho = random_rho(nk, nbf, rng)
    with scatt.ScatteringOperator.from_ldbd(tmp_path, eph_enable=False) as op:
        with pytest.raises(_lib.DMDError, match="P has not been"):
            op.eval(0.0, rho)
        op.load_P(tmp_path, impurity=1, scale=0.4, add=False)
        A1, A2 = op.download_P()
        assert np.array_equal(A1, 0.4 * Q1) and np.array_equal(A2, 0.4 * Q2)
    with scatt.ScatteringOperator.from_ldbd(tmp_path, scatt_enable=False) as op:
        assert op.info()["n_pairs"] == 0
        assert np.abs(op.eval(0.0, rho)).max() == 0.0
