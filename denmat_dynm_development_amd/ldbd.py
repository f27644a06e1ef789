"""ldbd_data and restart files through the C-ABI readers of libdmd_scatt.so.

The files are the ones the reference's initializer writes and its dynamics
reads (src_v4.5.7/electron.cpp:100-314, ElectronPhonon.cpp:132-253,
ElectronImpurity.cpp:3-48, lattice.h:48-56, DenMat.cpp:301-334); the readers
check every file's size as the reference's check_file_size does and raise
DMDError (code DMD_ERR_IO) where the reference would exit.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import LdbdInfo, call


def _d(dirname):
    return str(dirname).encode()


def read_info(dirname):
    """LdbdInfo of <dirname>/ldbd_size.dat."""
    info = LdbdInfo()
    call("dmd_ldbd_read_info", _d(dirname), ctypes.byref(info))
    return info


def _info(dirname, info):
    return read_info(dirname) if info is None else info


def read_ek(dirname, info=None, scissor=0.0):
    """ldbd_ek.bin as (nk, nb) float64; scissor (Hartree) added to bands nv.."""
    info = _info(dirname, info)
    e = np.empty((info.nk, info.nb), np.float64)
    call("dmd_ldbd_read_ek", _d(dirname), ctypes.byref(info), float(scissor), e.ctypes.data)
    return e


def read_kvec(dirname, info=None):
    info = _info(dirname, info)
    k = np.empty((info.nk, 3), np.float64)
    call("dmd_ldbd_read_kvec", _d(dirname), ctypes.byref(info), k.ctypes.data)
    return k


def read_smat(dirname, info=None):
    """ldbd_smat.bin as (nk, 3, nb_dm * nb_dm) complex128."""
    info = _info(dirname, info)
    s = np.empty((info.nk, 3, info.nb_dm * info.nb_dm), np.complex128)
    call("dmd_ldbd_read_smat", _d(dirname), ctypes.byref(info), s.ctypes.data)
    return s


def read_lmat(dirname, info=None):
    """ldbd_lmat.bin (needL) as (nk, 3, nb_dm * nb_dm) complex128."""
    info = _info(dirname, info)
    m = np.empty((info.nk, 3, info.nb_dm * info.nb_dm), np.complex128)
    call("dmd_ldbd_read_lmat", _d(dirname), ctypes.byref(info), m.ctypes.data)
    return m


def read_Bso(dirname, info=None):
    """ldbd_Bso.bin (alg_read_Bso) as (nk, 3) float64."""
    info = _info(dirname, info)
    b = np.empty((info.nk, 3), np.float64)
    call("dmd_ldbd_read_Bso", _d(dirname), ctypes.byref(info), b.ctypes.data)
    return b


def read_HEzmat(dirname, info=None):
    """ldbd_HEzmat.bin (scale_Ez) as (nk, nb_dm * nb_dm) complex128."""
    info = _info(dirname, info)
    m = np.empty((info.nk, info.nb_dm * info.nb_dm), np.complex128)
    call("dmd_ldbd_read_HEzmat", _d(dirname), ctypes.byref(info), m.ctypes.data)
    return m


def read_Umat(dirname, info=None):
    """ldbd_Umat.bin (electron.cpp:388-410) as (nk, nb_wannier, nb_eph) complex128:
    the e-ph columns of the full matrices, or the reduced file as it is."""
    info = _info(dirname, info)
    U = np.empty((info.nk, info.nb_wannier, info.nb_eph), np.complex128)
    call("dmd_ldbd_read_Umat", _d(dirname), ctypes.byref(info), U.ctypes.data)
    return U


def read_imsig(dirname, info=None):
    info = _info(dirname, info)
    x = np.empty((info.nk, info.nb_eph), np.float64)
    call("dmd_ldbd_read_imsig", _d(dirname), ctypes.byref(info), x.ctypes.data)
    return x


def read_R(dirname):
    """(dim, R) of ldbd_R.dat; the rows of R are the lattice vectors."""
    dim = ctypes.c_int()
    R = np.empty((3, 3), np.float64)
    call("dmd_ldbd_read_R", _d(dirname), ctypes.byref(dim), R.ctypes.data)
    return dim.value, R


def read_kpairs(dirname, info=None, first=0, count=None):
    info = _info(dirname, info)
    count = info.nkpair - first if count is None else count
    k1 = np.empty(count, np.uint64)
    k2 = np.empty(count, np.uint64)
    call("dmd_ldbd_read_kpairs", _d(dirname), ctypes.byref(info), int(first), int(count), k1.ctypes.data,
         k2.ctypes.data)
    return k1, k2


def read_P(dirname, info=None, scatt="lindblad", impurity=0, first=0, count=None):
    """P1, P2 of pairs [first, first+count) as (count, nb_eph**4) complex128."""
    info = _info(dirname, info)
    count = info.nkpair - first if count is None else count
    n4 = info.nb_eph ** 4
    P1 = np.empty((count, n4), np.complex128)
    P2 = np.empty((count, n4), np.complex128)
    call("dmd_ldbd_read_P", _d(dirname), ctypes.byref(info), scatt.encode(), int(impurity), int(first), int(count),
         P1.ctypes.data, P2.ctypes.data)
    return P1, P2


def sP_path(dirname, which, part, scatt="lindblad", hole=False):
    """ldbd_data/sP<which>_<scatt>[_hole]_<part>.bin (ElectronPhonon.cpp:240-242)."""
    return os.path.join(str(dirname), f"sP{which}_{scatt}{'_hole' if hole else ''}_{part}.bin")


def read_sP(dirname, which, info=None, scatt="lindblad"):
    """(ns, s, i, j) of the sparse P<which> files: int32 counts per stored
    pair, complex128 values, int32 row indices i1*nb+i2 and column indices
    i3*nb+i4 (sparse_mat::write_to_files, comm/sparse_matrix.h:37-42), with
    the sizes checked against each other and the stored pairs."""
    info = _info(dirname, info)
    hole = info.kind == 1
    ns = np.fromfile(sP_path(dirname, which, "ns", scatt, hole), dtype=np.int32)
    if len(ns) != info.nkpair:
        raise ValueError(f"sP{which} _ns: {len(ns)} counts for {info.nkpair} stored pairs")
    tot = int(ns.astype(np.int64).sum())
    sv = np.fromfile(sP_path(dirname, which, "s", scatt, hole), dtype=np.complex128)
    iv = np.fromfile(sP_path(dirname, which, "i", scatt, hole), dtype=np.int32)
    jv = np.fromfile(sP_path(dirname, which, "j", scatt, hole), dtype=np.int32)
    if not (len(sv) == len(iv) == len(jv) == tot):
        raise ValueError(f"sP{which}: {tot} elements counted, files hold {len(sv)}, {len(iv)}, {len(jv)}")
    return ns, sv, iv, jv


def read_restart(dirname, nk, nb):
    """(dm as (nk, nb * nb) complex128, t) from denmat_restart.bin and time_restart.dat."""
    dm = np.empty((nk, nb * nb), np.complex128)
    t = ctypes.c_double()
    call("dmd_restart_read", _d(dirname), int(nk), int(nb), dm.ctypes.data, ctypes.byref(t))
    return dm, t.value


def read_restart_time(dirname):
    """t of time_restart.dat (Hartree atomic time)."""
    t = ctypes.c_double()
    call("dmd_restart_read", _d(dirname), 0, 0, None, ctypes.byref(t))
    return t.value


def write_restart(dirname, dm, t):
    dm = np.ascontiguousarray(dm, np.complex128)
    nk = dm.shape[0]
    nb = int(round(np.sqrt(dm.size // nk)))
    call("dmd_restart_write", _d(dirname), int(nk), int(nb), dm.ctypes.data, float(t))


__all__ = ["LdbdInfo", "read_info", "read_ek", "read_kvec", "read_smat", "read_lmat", "read_Bso", "read_HEzmat",
           "read_Umat", "read_imsig", "read_R", "read_kpairs", "read_P", "sP_path", "read_sP", "read_restart", "read_restart_time",
           "write_restart", "_lib"]
