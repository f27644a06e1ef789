"""The ldbd_data and restart readers of the C-ABI (host code, no GPU): against
the writer's files, against the reference GaAs example's own ldbd_data
(tests/golden/gaas_ldbd, copied unchanged by make_gaas_ldbd.py), against the
ldbd_size.dat layouts of the initializer's saveSize for every band case, and
their error paths."""
import os

import numpy as np
import pytest

from helpers import write_ldbd_folder

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _mods():
    from denmat_dynm_development_amd import _lib, ldbd, scatt
    return _lib, ldbd, scatt


def test_reader_round_trip_of_the_writer(tmp_path):
    _lib, ldbd, scatt = _mods()
    p = scatt.synth_params(0)
    scatt.write_ldbd(str(tmp_path), p)
    info = ldbd.read_info(tmp_path)
    assert (info.kind, info.nb, info.nv, info.b_bot_dm, info.b_top_dm, info.b_bot_eph, info.b_top_eph) == \
        (0, p.nb, 0, 0, p.nb, 0, p.nb)
    assert info.nk == p.nk and info.nk_full == p.nk
    assert info.temperature == float("%21.14le" % p.temperature)  # the file keeps 15 digits
    assert info.n_modes == p.n_modes and info.degauss == float("%g" % p.degauss)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    assert info.nkpair == len(k1)
    assert np.array_equal(ldbd.read_ek(tmp_path, info), e)
    a1, a2 = ldbd.read_kpairs(tmp_path, info)
    assert np.array_equal(a1, k1) and np.array_equal(a2, k2)
    b1, b2 = ldbd.read_kpairs(tmp_path, info, first=7, count=30)
    assert np.array_equal(b1, k1[7:37]) and np.array_equal(b2, k2[7:37])
    P1, P2 = scatt.synth_P(p, e, k1, k2)
    R1, R2 = ldbd.read_P(tmp_path, info)
    assert np.array_equal(R1, P1) and np.array_equal(R2, P2)
    R1, R2 = ldbd.read_P(tmp_path, info, first=100, count=11)
    assert np.array_equal(R1, P1[100:111]) and np.array_equal(R2, P2[100:111])
    assert np.array_equal(ldbd.read_smat(tmp_path, info), scatt.synth_spin(p))
    assert ldbd.read_kvec(tmp_path, info).shape == (p.nk, 3)
    # scissor: added to the conduction bands nv.. only (electron.cpp:177-179); nv = 0 here
    assert np.array_equal(ldbd.read_ek(tmp_path, info, scissor=0.01), e + 0.01)


def test_reader_on_the_reference_gaas_ldbd_data():
    """ldbd_size.dat, ldbd_ek.bin, ldbd_kvec.bin, ldbd_imsig.bin, ldbd_R.dat and
    time_restart.dat of the reference's GaAs run (only-eph/n2e16)."""
    _lib, ldbd, scatt = _mods()
    d = os.path.join(GOLD, "gaas_ldbd")
    info = ldbd.read_info(d)
    # the values printed in the file (an older 8-integer line 2, no omegaMax line)
    assert (info.kind, info.n_line2, info.nb, info.nv) == (0, 8, 10, 6)
    assert (info.b_bot_dm, info.b_top_dm, info.b_bot_eph, info.b_top_eph, info.nb_wannier) == (2, 8, 6, 8, 16)
    assert info.nk_full == 7.077888e6 and info.nk == 941 and tuple(info.kmesh) == (192, 192, 192)
    assert info.nkpair == 443211 and (info.mode_start, info.mode_stop, info.n_modes) == (0, 6, 6)
    assert info.temperature == 9.50043414701698e-04
    assert info.mu_min == info.mu_max == info.mu == 5.25515330209184e-02
    assert (info.degauss, info.ndegauss) == (0.000183747, 1000.0)
    assert (info.e_bot_eph, info.e_top_eph) == (2.6275762e-02, 5.9201837e-02)
    assert info.omega_max == 0.0
    z = np.load(os.path.join(GOLD, "gaas_spin_restart.npz"))
    e = ldbd.read_ek(d, info)
    assert np.array_equal(e[:, info.b_bot_dm:info.b_top_dm], z["e_dm"])
    # band indices: a gap between the nv valence bands and the rest; the e-ph
    # bands inside the e-ph energy window the file states
    assert e[:, :info.nv].max() < e[:, info.nv:].min()
    ee = e[:, info.b_bot_eph:info.b_top_eph]
    assert info.e_bot_eph <= ee.min() and ee.max() <= info.e_top_eph
    kv = ldbd.read_kvec(d, info)
    assert np.abs(kv * 192 - np.round(kv * 192)).max() < 1e-9  # k on the 192^3 mesh
    im = ldbd.read_imsig(d, info)
    assert im.shape == (941, 2) and im.min() > 0
    dim, R = ldbd.read_R(d)
    assert dim == 3 and np.array_equal(R, z["R"])
    t = ldbd.read_restart_time(d)
    assert abs(t / float(z["t"]) - 1) < 1e-7  # "%14.7le" against sz_elec_tot.out's time


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_size_layouts_of_every_band_case(tmp_path, kind):
    """The three line-2 / line-4 layouts of saveSize (conduction only, valence
    only with _hole files, both) and a mode-skip line."""
    _lib, ldbd, scatt = _mods()
    rng = np.random.default_rng(kind)
    nk, nb, nv = 12, 7, 3
    e = np.sort(rng.uniform(-0.05, 0.1, (nk, nb)), axis=1)
    k1 = np.array([0, 0, 1, 3, 5], np.uint64)
    k2 = np.array([0, 4, 9, 3, 11], np.uint64)
    dm, eph = ((2, 7), (3, 5)) if kind != 1 else ((0, 5), (1, 3))
    write_ldbd_folder(str(tmp_path), kind, nb, nv, dm, eph, e, k1, k2, nk_full=96, kmesh=(4, 4, 6),
                      modes=(0, 9, 3, 5))
    info = ldbd.read_info(tmp_path)
    assert info.kind == kind and info.n_line2 == (11 if kind == 1 else 9)
    assert (info.nb, info.nv, info.b_bot_dm, info.b_top_dm) == (nb, nv, *dm)
    assert (info.b_bot_eph, info.b_top_eph, info.nb_wannier) == (*eph, nb + 4)
    assert info.nkpair == 5 and info.nk == nk and info.nk_full == 96 and tuple(info.kmesh) == (4, 4, 6)
    assert info.n_modes == 9 - (5 - 3)
    assert info.omega_max == 1.9e-3
    a1, a2 = ldbd.read_kpairs(tmp_path, info)  # the _hole files for kind 1
    assert np.array_equal(a1, k1) and np.array_equal(a2, k2)
    assert np.array_equal(ldbd.read_ek(tmp_path, info, scissor=0.5)[:, nv:], e[:, nv:] + 0.5)


def test_reader_errors(tmp_path):
    _lib, ldbd, scatt = _mods()
    p = scatt.synth_params(0)
    scatt.write_ldbd(str(tmp_path), p)
    info = ldbd.read_info(tmp_path)
    with open(tmp_path / "ldbd_ek.bin", "ab") as f:  # one double too many
        f.write(b"\0" * 8)
    with pytest.raises(_lib.DMDError, match="expected size") as ex:
        ldbd.read_ek(tmp_path, info)
    assert ex.value.code == _lib.DMD_ERR_IO
    os.remove(tmp_path / "ldbd_P2_lindblad.bin")
    with pytest.raises(_lib.DMDError, match="cannot open"):
        ldbd.read_P(tmp_path, info)
    with pytest.raises(_lib.DMDError):
        ldbd.read_kpairs(tmp_path, info, first=info.nkpair - 2, count=3)  # past the end
    with pytest.raises(_lib.DMDError, match="_D2"):
        ldbd.read_P(tmp_path, info, impurity=2)
    # a pair naming k >= nk, and a band window outside the density-matrix window
    bad = tmp_path / "bad"
    e = np.zeros((4, 3))
    write_ldbd_folder(str(bad), 0, 3, 0, (0, 3), (0, 2), e, np.array([0], np.uint64), np.array([4], np.uint64))
    with pytest.raises(_lib.DMDError, match="beyond nk"):
        ldbd.read_kpairs(bad, ldbd.read_info(bad))
    write_ldbd_folder(str(bad), 0, 3, 0, (1, 3), (0, 2), e, np.array([0], np.uint64), np.array([1], np.uint64))
    with pytest.raises(_lib.DMDError, match="bStart_dm <= bStart_eph"):
        ldbd.read_info(bad)


def test_restart_round_trip(tmp_path):
    _lib, ldbd, scatt = _mods()
    rng = np.random.default_rng(5)
    dm = rng.normal(size=(9, 16)) + 1j * rng.normal(size=(9, 16))
    t = 123456.789012345
    ldbd.write_restart(tmp_path, dm, t)
    assert open(tmp_path / "time_restart.dat").read() == "%14.7le" % t  # write_dm_tofile's format
    back, tb = ldbd.read_restart(tmp_path, 9, 4)
    assert np.array_equal(back, dm) and tb == float("%14.7le" % t)
    with pytest.raises(_lib.DMDError, match="expected size"):
        ldbd.read_restart(tmp_path, 10, 4)
