"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library
never loads it.

Arrays are numpy complex128 (interleaved re/im, the reference's complex
layout, src_v4.5.7/comm/scalar.h:67-69) or float64 / uint64.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_u64 = ctypes.POINTER(ctypes.c_uint64)
_i64 = ctypes.c_int64


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.oracle_compute_Pt.argtypes = [ctypes.c_int, ctypes.c_double, _d, _d, _d, _d, ctypes.c_int]
        L.oracle_evolve.argtypes = [_i64, ctypes.c_int, _i64, _u64, _u64, _d, _d, _d, _d, _d,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, _d, ctypes.c_int,
                                    ctypes.c_int, _d]
        L.oracle_compute_ddm_pairs.argtypes = [_i64, ctypes.c_int, _i64, _u64, _u64, _d, _d, _d, _d,
                                               ctypes.c_double, _d]
        L.oracle_compute_P.argtypes = [ctypes.c_int, _d, _d, ctypes.c_int, _d, _d, _d, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, _d, _d]
        L.oracle_compute_P_eimp.argtypes = [ctypes.c_int, _d, _d, _d, _d]
        L.oracle_coherent.argtypes = [_i64, ctypes.c_int, _d, _d, _d, ctypes.c_double, ctypes.c_int, _d]
        L.oracle_eimp_q2.argtypes = [_d, _d, _d]
        L.oracle_eimp_q2.restype = ctypes.c_double
        L.oracle_eimp_vq.argtypes = [ctypes.c_int] + [ctypes.c_double] * 7
        L.oracle_eimp_vq.restype = ctypes.c_double
        L.oracle_eimp_model_P.argtypes = ([_i64, ctypes.c_int, ctypes.c_int, _i64, _u64, _u64, _d, _d, _d, _d] +
                                          [ctypes.c_int] * 3 + [ctypes.c_double] * 11 + [_d, _d, _d])
        L.oracle_ee_model_P.argtypes = [_i64, ctypes.c_int, ctypes.c_int, _i64, _u64, _u64, _d, _d,
                                        ctypes.POINTER(ctypes.c_int), _d, _d, _d, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, _d, _d, ctypes.c_int, _d, ctypes.c_int] + \
            [ctypes.c_double] * 6 + [_d, _d, _d, _d, _d]
        L.oracle_ee_model_P3P6.argtypes = [_i64, ctypes.c_int, ctypes.c_int, _i64, _u64, _u64, _d, _d,
                                           ctypes.POINTER(ctypes.c_int), _d, _d, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int, _d, _d, ctypes.c_int] + \
            [ctypes.c_double] * 6 + [_d, _d, _d, _d]
        L.oracle_add_Lsc_ee.argtypes = [_i64, ctypes.c_int, _i64, _u64, _u64, _d, ctypes.c_double, _d, _d, _d, _d,
                                        _d, _d, _d, _d]
        L.oracle_eimp_rpa_vq.argtypes = [_i64, ctypes.c_int, ctypes.c_int, _d, ctypes.POINTER(ctypes.c_int), _d, _d,
                                         _d, ctypes.c_double, _d, ctypes.c_double, ctypes.c_double, _i64, _u64, _u64,
                                         _d, _d, ctypes.POINTER(ctypes.c_int64), _d]
        L.oracle_spin.argtypes = [_i64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, _d, _d, _d,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_double, _d]
        L.oracle_spin.restype = ctypes.c_double
        L.oracle_dm_Bpert_1st.argtypes = [_i64, ctypes.c_int, _d, _d, _d, _d, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, _d]
        L.oracle_rk4.argtypes = [_i64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64, _u64, _u64, _d, _d,
                                 _d, _d, _d, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, _i64, ctypes.c_int, _d, _d, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_int, _d, _d]
        L.oracle_set_Lsc.argtypes = [_i64, ctypes.c_int, _i64, _u64, _u64, _d, _d, _d, _d, _d]
        L.oracle_ee_dP_term.argtypes = [_i64, ctypes.c_int, _i64, _u64, _u64, _d, _d, _d, _d, ctypes.c_double,
                                        ctypes.c_double, _d]
        L.oracle_mt19937_64.argtypes = [ctypes.c_uint64, _i64, _u64]
        L.oracle_normal_draws.argtypes = [ctypes.c_uint64, _i64, _d]
        L.oracle_gfac_normal.argtypes = [_i64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _d]
        L.oracle_evolve_linear.argtypes = [_i64, ctypes.c_int, _i64, _u64, _u64, _d, _d, _d, _d, _d,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int, _d]
        _i32p = ctypes.POINTER(ctypes.c_int32)
        L.oracle_evolve_sparse.argtypes = [_i64, ctypes.c_int, _i64, _u64, _u64, _i32p, _d, _i32p, _i32p,
                                           _i32p, _d, _i32p, _i32p, _d, _d, _d, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, _d]
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_d if a.dtype != np.uint64 else _u64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def compute_Pt(nb, t, ek, ekp, P, minus):
    P = _c128(P)
    out = np.empty_like(P)
    lib().oracle_compute_Pt(nb, t, _p(_f64(ek)), _p(_f64(ekp)), _p(P), _p(out), int(minus))
    return out


def evolve(k1, k2, P1, P2, e, dm, dm1=None, t=0.0, prefac=1.0, expt=True, ddmdt_eq=None,
           schrodinger=False, nthreads=1):
    """electronphonon::evolve (src_v4.5.7/evolve_eph.cpp:34-166).

    P1, P2: (npairs, nb^4) complex; e: (nk, nb); dm, dm1: (nk, nb^2) complex.
    dm1 defaults to 1 - dm as set_oneminusdm forms it (DenMat.cpp:249-261).
    """
    e = _f64(e)
    nk, nb = e.shape
    dm = _c128(dm).reshape(nk, nb * nb)
    if dm1 is None:
        dm1 = one_minus(dm, nb)
    dm1 = _c128(dm1).reshape(nk, nb * nb)
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    P1 = _c128(P1)
    P2 = _c128(P2)
    out = np.zeros((nk, nb * nb), np.complex128)
    eq = None if ddmdt_eq is None else _c128(ddmdt_eq)
    rc = lib().oracle_evolve(nk, nb, len(k1), _p(k1), _p(k2), _p(P1), _p(P2), _p(e), _p(dm), _p(dm1),
                             float(t), float(prefac), int(bool(expt)), _p(eq), int(bool(schrodinger)),
                             int(nthreads), _p(out))
    if rc != 0:
        raise ValueError(f"oracle_evolve failed: {rc}")
    return out


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def evolve_sparse(k1, k2, sP1, sP2, e, dm, dm1=None, t=0.0, prefac=1.0, expt=True):
    """The sparse branch of electronphonon::evolve (src_v4.5.7/evolve_eph.cpp:69-86,
    194-203; comm/sparse_matrix.cpp:6-40).  sP1, sP2: (ns, s, i, j) as the
    files sP{1,2}_*_{ns,s,i,j}.bin hold them: per-pair counts, then complex
    values, row indices i1*nb+i2 and column indices i3*nb+i4 in file order."""
    e = _f64(e)
    nk, nb = e.shape
    dm = _c128(dm).reshape(nk, nb * nb)
    dm1 = one_minus(dm, nb) if dm1 is None else _c128(dm1).reshape(nk, nb * nb)
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    keep, args = [], []
    for ns, sv, iv, jv in (sP1, sP2):
        a_ns, p_ns = _i32(ns)
        a_s = _c128(sv)
        a_i, p_i = _i32(iv)
        a_j, p_j = _i32(jv)
        if len(a_ns) != len(k1) or int(a_ns.sum()) != len(a_s) or len(a_i) != len(a_s) or len(a_j) != len(a_s):
            raise ValueError("sparse lists do not match the stored pairs")
        keep += [a_ns, a_s, a_i, a_j]
        args += [p_ns, _p(a_s), p_i, p_j]
    out = np.zeros((nk, nb * nb), np.complex128)
    rc = lib().oracle_evolve_sparse(nk, nb, len(k1), _p(k1), _p(k2), *args, _p(e), _p(dm), _p(dm1), float(t),
                                    float(prefac), int(bool(expt)), _p(out))
    if rc != 0:
        raise ValueError(f"oracle_evolve_sparse failed: {rc}")
    return out


def set_Lsc(k1, k2, P1, P2, f_eq):
    """electronphonon::set_Lsc, e-ph part (src_v4.5.7/ElectronPhonon_linear.cpp:3-36,
    147-207): (Lscij, Lscji) per stored pair, Lscii folded into diagonal pairs."""
    f_eq = _f64(f_eq)
    nk, nb = f_eq.shape
    P1 = _c128(P1)
    P2 = _c128(P2)
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    Lij = np.zeros_like(P1)
    Lji = np.zeros_like(P2)
    rc = lib().oracle_set_Lsc(nk, nb, len(k1), _p(k1), _p(k2), _p(P1), _p(P2), _p(f_eq), _p(Lij), _p(Lji))
    if rc != 0:
        raise ValueError(f"oracle_set_Lsc failed: {rc}")
    return Lij, Lji


def evolve_linear(k1, k2, Lij, Lji, e, dm, f_eq, t=0.0, prefac=1.0, expt=True):
    """electronphonon::evolve_linear (src_v4.5.7/evolve_eph_linear.cpp:3-53)."""
    e = _f64(e)
    nk, nb = e.shape
    dm = _c128(dm).reshape(nk, nb * nb)
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    out = np.zeros((nk, nb * nb), np.complex128)
    rc = lib().oracle_evolve_linear(nk, nb, len(k1), _p(k1), _p(k2), _p(_c128(Lij)), _p(_c128(Lji)), _p(e), _p(dm),
                                    _p(_f64(f_eq)), float(t), float(prefac), int(bool(expt)), _p(out))
    if rc != 0:
        raise ValueError(f"oracle_evolve_linear failed: {rc}")
    return out


def compute_ddm_pairs(k1, k2, P1, P2, dm, dm1=None, prefac=1.0, nk=None, nb=None):
    dm = _c128(dm)
    nk = dm.shape[0] if nk is None else nk
    nb = int(round(np.sqrt(dm.shape[1]))) if nb is None else nb
    if dm1 is None:
        dm1 = one_minus(dm, nb)
    dm1 = _c128(dm1)
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    out = np.zeros((nk, nb * nb), np.complex128)
    rc = lib().oracle_compute_ddm_pairs(nk, nb, len(k1), _p(k1), _p(k2), _p(_c128(P1)), _p(_c128(P2)),
                                        _p(dm), _p(dm1), float(prefac), _p(out))
    if rc != 0:
        raise ValueError(f"oracle_compute_ddm_pairs failed: {rc}")
    return out


def compute_P(Ek, Ekp, omega, gm, gp, T, sigma, detail_balance=False):
    nb = len(Ek)
    nm = len(omega)
    P1 = np.zeros(nb ** 4, np.complex128)
    P2 = np.zeros(nb ** 4, np.complex128)
    lib().oracle_compute_P(nb, _p(_f64(Ek)), _p(_f64(Ekp)), nm, _p(_f64(omega)), _p(_c128(gm)),
                           _p(_c128(gp)), float(T), float(sigma), int(detail_balance), _p(P1), _p(P2))
    return P1, P2


def eimp_q2(GGT, ka, kb):
    """|wrap(ka - kb)|^2 under GGT (comm/kmap.h:6-15, matrix3.h:119-122)."""
    return lib().oracle_eimp_q2(_p(_f64(GGT)), _p(_f64(ka)), _p(_f64(kb)))


def eimp_vq(m, q2):
    """coulomb_model::vq (Coulomb_Model.h:632-662) for a model dict m."""
    return lib().oracle_eimp_vq(int(m["screening"]), float(q2), m["prefac_vq"], m["prefac_vq_bare"],
                                m["qscr2_debye"], m["fac0"], m["fac2"], m["fac4"])


def eimp_rpa_vq(k1, k2, U, kvec, e, f, m, cells=False):
    """(v(q) per stored pair, q_scr^2 per q of the mesh's q list) of the static RPA
    screening (Coulomb_Model.h:690-744,820-828); with cells=True also v(q) of
    every mesh cell (0 for cells that are no k-point difference, and at q = 0)."""
    e = _f64(e)
    nk, nb = e.shape
    U = _c128(U)
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    km = (ctypes.c_int * 3)(*[int(x) for x in m["kmesh"]])
    vq = np.zeros(len(k1))
    qs = np.zeros(int(np.prod(m["kmesh"])))
    nq = ctypes.c_int64(0)
    lib().oracle_eimp_rpa_vq(nk, nb, U.shape[1], _p(_f64(kvec)), km, _p(U), _p(e), _p(_f64(f)),
                             float(m["temperature"]), _p(_f64(m["GGT"])), float(m["prefac_vq"]), float(m["nk_full"]),
                             len(k1), _p(k1), _p(k2), _p(vq), _p(qs), ctypes.byref(nq),
                             _p(qs_cell := np.zeros(int(np.prod(m["kmesh"])))))
    if cells:
        return vq, qs[:nq.value], qs_cell
    return vq, qs[:nq.value]


def eimp_model_P(k1, k2, U, kvec, e, m, f=None):
    """elecimp_model::calc_P of every stored pair (ElecImp_Model.h:67-137,185-213).
    U: (nk, nb_wannier, nb) complex; kvec (nk, 3); e (nk, nb); m a dict with the
    fields of dmd_eimp_model; f (nk, nb) at mu for the RPA screening."""
    e = _f64(e)
    nk, nb = e.shape
    U = _c128(U)
    nw = U.shape[1]
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    P1 = np.zeros((len(k1), nb ** 4), np.complex128)
    P2 = np.zeros_like(P1)
    vq = None
    if int(m["screening"]) == 3:
        vq, _ = eimp_rpa_vq(k1, k2, U, kvec, e, f, m)
    rc = lib().oracle_eimp_model_P(nk, nb, nw, len(k1), _p(k1), _p(k2), _p(U), _p(_f64(kvec)), _p(e),
                                   _p(_f64(m["GGT"])), int(m["screening"]), int(m["lindblad"]),
                                   int(m["detail_balance"]), m["prefac_A"], m["prefac_vq"], m["prefac_vq_bare"],
                                   m["qscr2_debye"], m["fac0"], m["fac2"], m["fac4"], m["degauss"], m["ethr"],
                                   m["omegaL"], m["temperature"], _p(vq) if vq is not None else None, _p(P1),
                                   _p(P2))
    if rc:
        raise ValueError("oracle_eimp_model_P: k index out of range")
    return P1, P2


_NO_SCREENING = {"GGT": np.eye(3), "screening": 0, "prefac_vq": 0.0, "prefac_vq_bare": 0.0,
                 "qscr2_debye": 0.0, "fac0": 0.0, "fac2": 0.0, "fac4": 0.0}


def ee_model_P(k1, k2, U, kvec, e, f, vq_pair, kmesh, degauss, ethr, nk_full, lindblad=True, antisym_model=None,
               vq_cell=None, dm=None, dm1=None, with_dP=False):
    """elecelec_model::calc_P (ElecElec_Model.h:185-262); vq_pair: v(k1 - k2)
    of each stored pair.  antisym_model: a screening dict (eimp_model fields)
    to include the exchange term (ee_antisymmetry), whose v(k1 - k4) is
    evaluated from it, or from vq_cell (v per mesh cell, the static RPA) when
    given.  dm, dm1 (nk x nb^2, Schroedinger picture): P at that state
    (calc_A1rho) instead of at f; with with_dP (alg_linearize_dPee) P stays at
    f (= f_eq) and (P1, P2, dP1, dP2) is returned, dP from calc_A1rho_dP."""
    m = antisym_model or _NO_SCREENING
    e = _f64(e)
    nk, nb = e.shape
    U = _c128(U)
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    km = (ctypes.c_int * 3)(*[int(x) for x in kmesh])
    P1 = np.zeros((len(k1), nb ** 4), np.complex128)
    P2 = np.zeros_like(P1)
    dP1 = np.zeros_like(P1) if with_dP else None
    dP2 = np.zeros_like(P1) if with_dP else None
    if (dm is None) != (dm1 is None) or (with_dP and dm is None):
        raise ValueError("dm and dm1 go together (and with_dP needs them)")
    lib().oracle_ee_model_P(nk, nb, U.shape[1], len(k1), _p(k1), _p(k2), _p(U), _p(_f64(kvec)), km, _p(e),
                            _p(_f64(f)), _p(_f64(vq_pair)), int(lindblad), float(degauss), float(ethr),
                            float(nk_full), _p(P1), _p(P2), int(antisym_model is not None), _p(_f64(m["GGT"])),
                            int(m["screening"]), float(m["prefac_vq"]), float(m["prefac_vq_bare"]),
                            float(m["qscr2_debye"]), float(m["fac0"]), float(m["fac2"]), float(m["fac4"]),
                            _p(None if vq_cell is None else _f64(vq_cell)),
                            _p(None if dm is None else _c128(dm)), _p(None if dm1 is None else _c128(dm1)),
                            _p(dP1), _p(dP2))
    if with_dP:
        return P1, P2, dP1, dP2
    return P1, P2


def ee_dP_term(k1, k2, dP1, dP2, e, f_eq, t=0.0, prefac=1.0):
    """The alg_linearize_dPee part of electronphonon::evolve
    (src_v4.5.7/evolve_eph.cpp:107-131,138,144-151): the row sums of dP at
    f_eq, 1 - f_eq, the row term with cis((e_i - e_j) t) and its Hermitian
    assembly times prefac/2.  Returns (nk, nb^2) complex."""
    e = _f64(e)
    nk, nb = e.shape
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    out = np.zeros((nk, nb * nb), np.complex128)
    rc = lib().oracle_ee_dP_term(nk, nb, len(k1), _p(k1), _p(k2), _p(_c128(dP1)), _p(_c128(dP2)), _p(e),
                                 _p(_f64(f_eq)), float(t), float(prefac), _p(out))
    if rc != 0:
        raise ValueError(f"oracle_ee_dP_term failed: {rc}")
    return out


def mt19937_64(seed, n):
    """The first n outputs of a 64-bit Mersenne Twister seeded with seed (the
    engine behind src_v4.5.7/comm/Random.cpp:24)."""
    out = np.zeros(n, np.uint64)
    lib().oracle_mt19937_64(int(seed), int(n), _p(out))
    return out


def normal_draws(seed, n):
    """n standard normal draws of the polar method over that engine
    (src_v4.5.7/comm/Random.cpp:25,41-45 with sigma 1, mean 0, no cap)."""
    out = np.zeros(n, np.float64)
    lib().oracle_normal_draws(int(seed), int(n), _p(out))
    return out


def gfac_normal(n, mean, sigma, cap):
    """The k-resolved g factors of gfac_normal_dist: random_normal_array(gfack,
    nk, gfac_mean, gfac_sigma, gfac_cap) (src_v4.5.7/electron.cpp:7-18,
    comm/myarray.cpp:299-321), generator seeded with n."""
    out = np.zeros(n, np.float64)
    lib().oracle_gfac_normal(int(n), float(mean), float(sigma), float(cap), _p(out))
    return out


def ee_model_P3P6(k1, k2, U, kvec, e, f, kmesh, degauss, ethr, nk_full, model, lindblad=True, antisym=False,
                  vq_cell=None):
    """elecelec_model::calc_P3P4 and calc_P5P6 at f for every stored pair
    (scattering_model/ElecElec_Model_linear.cpp:3-111); model: a screening dict
    (eimp_model fields) for the closed-form v(q), or vq_cell (v per mesh cell,
    static RPA).  P3 and P4 are zero with antisym, as the reference skips them.
    Returns (P3, P4, P5, P6)."""
    e = _f64(e)
    nk, nb = e.shape
    U = _c128(U)
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    km = (ctypes.c_int * 3)(*[int(x) for x in kmesh])
    out = [np.zeros((len(k1), nb ** 4), np.complex128) for _ in range(4)]
    m = model
    rc = lib().oracle_ee_model_P3P6(nk, nb, U.shape[1], len(k1), _p(k1), _p(k2), _p(U), _p(_f64(kvec)), km, _p(e),
                                    _p(_f64(f)), int(lindblad), float(degauss), float(ethr), float(nk_full),
                                    int(antisym), _p(_f64(m["GGT"])), _p(None if vq_cell is None else _f64(vq_cell)),
                                    int(m["screening"]), float(m["prefac_vq"]), float(m["prefac_vq_bare"]),
                                    float(m["qscr2_debye"]), float(m["fac0"]), float(m["fac2"]), float(m["fac4"]),
                                    _p(None if antisym else out[0]), _p(None if antisym else out[1]), _p(out[2]),
                                    _p(out[3]))
    if rc != 0:
        raise ValueError(f"oracle_ee_model_P3P6 failed: {rc}")
    return tuple(out)


def add_Lsc_ee(k1, k2, f_eq, Lij, Lji, extra=0.0, Pee=None, P3456=None):
    """The e-e terms of set_Lsc beyond the P1/P2 form (ElectronPhonon_linear.cpp:
    183-253): extra x add_Lscij_from_P1_P2 of Pee = (P1, P2) (extra = 1 under
    ee_antisymmetry, whose fac2 doubles the e-e Lscij/Lscji), and the P3/P4,
    P5/P6 terms.  Adds into copies of (Lij, Lji) and returns them."""
    f_eq = _f64(f_eq)
    nk, nb = f_eq.shape
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    Lij, Lji = _c128(Lij).copy(), _c128(Lji).copy()
    P1, P2 = (None, None) if Pee is None else (_c128(Pee[0]), _c128(Pee[1]))
    P3, P4, P5, P6 = (None,) * 4 if P3456 is None else tuple(_c128(x) for x in P3456)
    if P3 is not None and not np.any(P3) and not np.any(P4):
        P3 = P4 = None  # skipped under exchange
    rc = lib().oracle_add_Lsc_ee(nk, nb, len(k1), _p(k1), _p(k2), _p(f_eq), float(extra), _p(P1), _p(P2), _p(P3),
                                 _p(P4), _p(P5), _p(P6), _p(Lij), _p(Lji))
    if rc != 0:
        raise ValueError(f"oracle_add_Lsc_ee failed: {rc}")
    return Lij, Lji


def compute_P_eimp(A1, A2):
    A1 = _c128(A1)
    nb = int(round(np.sqrt(A1.size)))
    P1 = np.zeros(nb ** 4, np.complex128)
    P2 = np.zeros(nb ** 4, np.complex128)
    lib().oracle_compute_P_eimp(nb, _p(A1), _p(_c128(A2)), _p(P1), _p(P2))
    return P1, P2


def coherent(H, e, dm, t, interaction=True):
    e = _f64(e)
    nk, nb = e.shape
    out = np.zeros((nk, nb * nb), np.complex128)
    lib().oracle_coherent(nk, nb, _p(_c128(H)), _p(e), _p(_c128(dm)), float(t), int(interaction), _p(out))
    return out


def spin(s, e, f, dm, t, i0, i1, prefac, interaction=True, with_amp=False):
    e = _f64(e)
    nk, nb = e.shape
    amp = np.zeros(1)
    v = lib().oracle_spin(nk, nb, i0, i1, _p(_c128(s)), _p(e), _p(_f64(f)), _p(_c128(dm)), float(t),
                          int(interaction), float(prefac), _p(amp))
    return (v, float(amp[0])) if with_amp else v


def dm_Bpert_1st(e, f, s3, Bpert, T, t0=0.0, interaction=True):
    e = _f64(e)
    nk, nb = e.shape
    out = np.zeros((nk, nb * nb), np.complex128)
    lib().oracle_dm_Bpert_1st(nk, nb, _p(e), _p(_f64(f)), _p(_c128(s3)), _p(_f64(Bpert)), float(T),
                              float(t0), int(interaction), _p(out))
    return out


def rk4(k1, k2, P1, P2, e_dm, dm0, dt, nsteps, b0=0, nb=None, prefac=1.0, expt=True, t0=0.0, H=None,
        ddmdt_eq=None, report_every=0, s_obs=None, f_dm=None, i0=0, i1=None, obs_prefac=1.0,
        nthreads=1):
    e_dm = _f64(e_dm)
    nk, nbd = e_dm.shape
    nb = nbd if nb is None else nb
    i1 = nbd if i1 is None else i1
    dm = _c128(dm0).copy().reshape(nk, nbd * nbd)
    nobs = (nsteps // report_every + 1) if (report_every and s_obs is not None) else 0
    obs = np.zeros(max(nobs, 1))
    k1 = np.ascontiguousarray(k1, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2, dtype=np.uint64)
    rc = lib().oracle_rk4(nk, nbd, b0, nb, len(k1), _p(k1), _p(k2), _p(_c128(P1)), _p(_c128(P2)), _p(e_dm),
                          _p(None if H is None else _c128(H)), _p(None if ddmdt_eq is None else _c128(ddmdt_eq)),
                          float(prefac), int(expt), float(t0), float(dt), int(nsteps), int(report_every),
                          _p(None if s_obs is None else _c128(s_obs)), _p(None if f_dm is None else _f64(f_dm)),
                          i0, i1, float(obs_prefac), int(nthreads), _p(dm), _p(obs))
    if rc != 0:
        raise ValueError(f"oracle_rk4 failed: {rc}")
    return dm, obs[:nobs]


def one_minus(dm, nb):
    """1 - rho exactly as set_oneminusdm forms it (src_v4.5.7/DenMat.cpp:249-261)."""
    dm = _c128(dm)
    out = -dm.reshape(-1, nb, nb)
    idx = np.arange(nb)
    out[:, idx, idx] = 1.0 - dm.reshape(-1, nb, nb)[:, idx, idx]
    return out.reshape(dm.shape)


def max_threads():
    return lib().oracle_max_threads()
