"""Host-side checks (no GPU): the C-ABI library loads and exports every symbol
declared in include/dmd_scatt.h, argument validation, the synthetic
generator against the oracle's restatement of the initializer's compute_P,
and the ldbd_data writer's layouts."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "dmd_scatt.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dmd_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from denmat_dynm_development_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python prototypes cover the header exactly
    assert sorted(_lib.PROTOTYPES) == syms


def test_create_rejects_bad_arguments_before_touching_the_device():
    from denmat_dynm_development_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    e = np.zeros((4, 2))
    k1 = np.array([0, 2], np.uint64)
    k2 = np.array([1, 1], np.uint64)  # 2 > 1: not k1st <= k2nd
    rc = L.dmd_scatt_create(4, 2, 2, k1.ctypes.data, k2.ctypes.data, e.ctypes.data, 4.0, 0, ctypes.byref(h))
    assert rc == _lib.DMD_ERR_ARG and "k1st <= k2nd" in _lib.last_error()
    rc = L.dmd_scatt_create(4, 5, 0, None, None, e.ctypes.data, 4.0, 0, ctypes.byref(h))
    assert rc == _lib.DMD_ERR_UNSUPPORTED
    k2 = np.array([1, 9], np.uint64)
    rc = L.dmd_scatt_create(4, 2, 2, k1.ctypes.data, k2.ctypes.data, e.ctypes.data, 4.0, 0, ctypes.byref(h))
    assert rc == _lib.DMD_ERR_ARG


def test_synth_kpairs_sorted_unique_and_sized():
    from denmat_dynm_development_amd import scatt
    p = scatt.synth_params(0)
    k1, k2 = scatt.synth_kpairs(p)
    assert np.all(k1 < k2) and np.all(k2 < p.nk)
    key = k1 * p.nk + k2
    assert np.all(np.diff(key) > 0)  # sorted, no duplicates
    # each k draws `partners` partners: n_pairs close to nk * partners (a few collide)
    assert 0.85 * p.nk * p.partners < len(k1) <= p.nk * p.partners
    a1, a2 = scatt.synth_kpairs(p)
    assert np.array_equal(a1, k1) and np.array_equal(a2, k2)
    p.partners_min, p.partners_max, p.powerlaw_alpha = 2, 20, 1.5
    b1, b2 = scatt.synth_kpairs(p)
    assert np.all(b1 < b2) and len(b1) > 0


def test_generator_matches_oracle_compute_P():
    """The product's P (counter-based couplings + the initializer formula) equals
    the oracle's restatement of compute_P fed the same couplings."""
    from denmat_dynm_development_amd import scatt
    for preset in (0,):
        p = scatt.synth_params(preset)
        for nb in (2, 4):
            p.nb = nb
            e = scatt.synth_energies(p)
            k1, k2 = scatt.synth_kpairs(p)
            sel = np.arange(0, len(k1), max(1, len(k1) // 7))
            P1, P2 = scatt.synth_P(p, e, k1[sel], k2[sel])
            for j, i in enumerate(sel):
                omega, gm, gp = scatt.synth_couplings(p, k1[i], k2[i])
                Q1, Q2 = oracle.compute_P(e[k1[i]], e[k2[i]], omega, gm, gp, p.temperature, p.degauss)
                assert np.abs(P1[j] - Q1).max() <= 1e-13 * np.abs(Q1).max()
                assert np.abs(P2[j] - Q2).max() <= 1e-13 * np.abs(Q2).max()


def test_generator_reverse_pair_identity():
    """P1 of the swapped pair equals conj(P2): g+_{kk'} = dagger(g-_{k'k})."""
    from denmat_dynm_development_amd import scatt
    p = scatt.synth_params(0)
    e = scatt.synth_energies(p)
    k1 = np.array([3, 10], np.uint64)
    k2 = np.array([10, 3], np.uint64)
    P1, P2 = scatt.synth_P(p, e, k1, k2)
    assert np.abs(P1[1] - P2[0].conj()).max() <= 1e-14 * np.abs(P1[1]).max()
    assert np.abs(P2[1] - P1[0].conj()).max() <= 1e-14 * np.abs(P2[1]).max()


def test_eimp_fold_in_is_linear_in_fraction():
    from denmat_dynm_development_amd import scatt
    p = scatt.synth_params(0)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    k1, k2 = k1[:20], k2[:20]
    A1, A2 = scatt.synth_P(p, e, k1, k2)
    p.sigma_imp, p.imp_fraction = 5e-4, 0.25
    B1, B2 = scatt.synth_P(p, e, k1, k2)
    p.imp_fraction = 0.5
    C1, C2 = scatt.synth_P(p, e, k1, k2)
    # the e-i term rides on the e-ph P; compare on the scale of P itself
    for C, B, A in ((C1, B1, A1), (C2, B2, A2)):
        d = np.abs((C - A) - 2 * (B - A)).max()
        assert d <= 1e-13 * np.abs(A).max()
        assert np.abs(B - A).max() > 1e-6 * np.abs(A).max()  # the term is really there


def test_ldbd_writer_layout(tmp_path):
    from denmat_dynm_development_amd import scatt
    p = scatt.synth_params(0)
    scatt.write_ldbd(str(tmp_path), p)
    k1, k2 = scatt.synth_kpairs(p)
    np_, nk, nb = len(k1), p.nk, p.nb
    sizes = {"ldbd_ek.bin": nk * nb * 8, "ldbd_kvec.bin": nk * 3 * 8, "ldbd_smat.bin": nk * 3 * nb * nb * 16,
             "ldbd_kpair_k1st.bin": np_ * 8, "ldbd_kpair_k2nd.bin": np_ * 8,
             "ldbd_P1_lindblad.bin": np_ * nb ** 4 * 16, "ldbd_P2_lindblad.bin": np_ * nb ** 4 * 16}
    for f, sz in sizes.items():
        assert os.path.getsize(tmp_path / f) == sz, f
    lines = open(tmp_path / "ldbd_size.dat").read().splitlines()
    assert lines[0].startswith("There are scatterings only for conduction electrons")
    assert [int(x) for x in lines[1].split("#")[0].split()][:6] == [nb, 0, 0, nb, 0, nb]
    assert int(lines[2].split()[1]) == nk and int(lines[3].split()[0]) == np_
    assert np.array_equal(np.fromfile(tmp_path / "ldbd_kpair_k2nd.bin", np.uint64), k2)
    e = np.fromfile(tmp_path / "ldbd_ek.bin").reshape(nk, nb)
    assert np.array_equal(e, scatt.synth_energies(p))
    P1 = np.fromfile(tmp_path / "ldbd_P1_lindblad.bin", np.complex128).reshape(np_, nb ** 4)
    Q1, _ = scatt.synth_P(p, e, k1[:3], k2[:3])
    assert np.array_equal(P1[:3], Q1)
