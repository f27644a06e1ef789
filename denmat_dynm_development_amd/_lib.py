"""Loader and prototypes of libdmd_scatt.so (the HIP library behind the C-ABI
declared in include/dmd_scatt.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
denmat_dynm_development_amd/csrc``).  There is no fallback: when the shared
object is missing every entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdmd_scatt.so")

DMD_OK = 0
DMD_ERR_ARG = -1
DMD_ERR_HIP = -2
DMD_ERR_STATE = -3
DMD_ERR_UNSUPPORTED = -4
DMD_ERR_IO = -5
DMD_ERR_NOMEM = -6

DMD_PREC_FP64 = 0
DMD_PREC_C64 = 1
DMD_PREC_BF16 = 2
DMD_SCR_UNSCREENED = 0
DMD_SCR_DEBYE = 1
DMD_SCR_BECHSTEDT = 2
DMD_SCR_RPA = 3
DMD_SCR_LINDHARD = 4

DMD_KERNEL_AUTO = 0
DMD_KERNEL_ROWS = 1
DMD_KERNEL_SINGLE_READ = 2
DMD_KERNEL_TILES = 3


class DMDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class SynthParams(ctypes.Structure):
    """dmd_synth_params (include/dmd_scatt.h)."""
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("nb", ctypes.c_int),
        ("n_modes", ctypes.c_int),
        ("nk", ctypes.c_int64),
        ("partners", ctypes.c_int64),
        ("partners_min", ctypes.c_int64),
        ("partners_max", ctypes.c_int64),
        ("powerlaw_alpha", ctypes.c_double),
        ("e_max", ctypes.c_double),
        ("split_max", ctypes.c_double),
        ("gap_min", ctypes.c_double),
        ("gap_max", ctypes.c_double),
        ("omega_min", ctypes.c_double),
        ("omega_max", ctypes.c_double),
        ("sigma_g", ctypes.c_double),
        ("degauss", ctypes.c_double),
        ("temperature", ctypes.c_double),
        ("sigma_imp", ctypes.c_double),
        ("imp_fraction", ctypes.c_double),
        ("branch_mask", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LdbdInfo(ctypes.Structure):
    """dmd_ldbd_info (include/dmd_scatt.h): the contents of ldbd_size.dat."""
    _fields_ = [
        ("kind", ctypes.c_int),
        ("n_line2", ctypes.c_int),
        ("nb", ctypes.c_int),
        ("nv", ctypes.c_int),
        ("b_bot_dm", ctypes.c_int),
        ("b_top_dm", ctypes.c_int),
        ("b_bot_eph", ctypes.c_int),
        ("b_top_eph", ctypes.c_int),
        ("nb_wannier", ctypes.c_int),
        ("nk_full", ctypes.c_double),
        ("nk", ctypes.c_int64),
        ("kmesh", ctypes.c_int * 3),
        ("nkpair", ctypes.c_int64),
        ("mode_start", ctypes.c_int),
        ("mode_stop", ctypes.c_int),
        ("n_modes", ctypes.c_int),
        ("temperature", ctypes.c_double),
        ("mu_min", ctypes.c_double),
        ("mu_max", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("degauss", ctypes.c_double),
        ("ndegauss", ctypes.c_double),
        ("e_bot_probe", ctypes.c_double),
        ("e_top_probe", ctypes.c_double),
        ("e_bot_dm", ctypes.c_double),
        ("e_top_dm", ctypes.c_double),
        ("e_bot_eph", ctypes.c_double),
        ("e_top_eph", ctypes.c_double),
        ("omega_max", ctypes.c_double),
        ("b_skip_wannier", ctypes.c_int),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["kmesh"] = tuple(self.kmesh)
        return d

    @property
    def nb_dm(self):
        return self.b_top_dm - self.b_bot_dm

    @property
    def nb_eph(self):
        return self.b_top_eph - self.b_bot_eph


class EimpModel(ctypes.Structure):
    """dmd_eimp_model (include/dmd_scatt.h): the model ionized-impurity matrix
    element and its static screening."""
    _fields_ = [("screening", ctypes.c_int), ("lindblad", ctypes.c_int), ("detail_balance", ctypes.c_int),
                ("nb_wannier", ctypes.c_int), ("prefac_A", ctypes.c_double), ("prefac_vq", ctypes.c_double),
                ("prefac_vq_bare", ctypes.c_double), ("qscr2_debye", ctypes.c_double), ("fac0", ctypes.c_double),
                ("fac2", ctypes.c_double), ("fac4", ctypes.c_double), ("degauss", ctypes.c_double),
                ("ethr", ctypes.c_double), ("omegaL", ctypes.c_double), ("temperature", ctypes.c_double),
                ("GGT", ctypes.c_double * 9), ("kmesh", ctypes.c_int * 3), ("nk_full", ctypes.c_double),
                ("exchange", ctypes.c_int)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["GGT"] = list(self.GGT)
        d["kmesh"] = list(self.kmesh)
        return d


class OdeStats(ctypes.Structure):
    """dmd_ode_stats (include/dmd_scatt.h)."""
    _fields_ = [("steps", ctypes.c_int64), ("rejected", ctypes.c_int64), ("rhs_evals", ctypes.c_int64),
                ("h_next", ctypes.c_double), ("last_err_ratio", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pd = ctypes.POINTER(ctypes.c_double)
_pSP = ctypes.POINTER(SynthParams)
_pLI = ctypes.POINTER(LdbdInfo)
_cs = ctypes.c_char_p

# name -> argtypes (all return int unless listed in _RESTYPE)
PROTOTYPES = {
    "dmd_scatt_last_error": [],
    "dmd_scatt_version": [],
    "dmd_scatt_create": [_i64, _i, _i64, _vp, _vp, _vp, _d, _i, ctypes.POINTER(_vp)],
    "dmd_scatt_destroy": [_vp],
    "dmd_scatt_upload_P": [_vp, _vp, _vp, _d],
    "dmd_scatt_add_P": [_vp, _vp, _vp, _d],
    "dmd_scatt_download_P": [_vp, _i64, _i64, _vp, _vp],
    "dmd_scatt_copy_P": [_vp, _vp, _d, _i],
    "dmd_scatt_add_eimp_model": [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(EimpModel), _d, _i],
    "dmd_scatt_add_ee_model": [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(EimpModel), _d, _i],
    "dmd_scatt_add_ee_model_state": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(EimpModel), _d, _i],
    "dmd_scatt_add_ee_dP": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(EimpModel), _d, _i],
    "dmd_scatt_set_ee_dP": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(EimpModel)],
    "dmd_scatt_get_ee_dP": [_vp, _vp],
    "dmd_scatt_set_rhs_hook": [_vp, _vp, _vp],
    "dmd_scatt_ee_vq_cells": [_vp, _vp, _vp, _vp, ctypes.POINTER(EimpModel), _vp],
    "dmd_scatt_add_Lsc_ee": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(EimpModel), _d, _d, _d, _i],
    "dmd_scatt_generate_P": [_vp, _pSP],
    "dmd_scatt_set_flags": [_vp, _i, _i, _i],
    "dmd_scatt_set_ddmdt_eq": [_vp, _vp],
    "dmd_scatt_compute_ddmdt_eq": [_vp, _vp],
    "dmd_scatt_set_split": [_vp, _vp],
    "dmd_scatt_set_Lsc": [_vp, _vp],
    "dmd_scatt_eval": [_vp, _d, _vp, _vp, _vp],
    "dmd_scatt_eval_noalloc": [_vp, _d, _vp, _vp],
    "dmd_scatt_eval_batched": [_vp, _d, _i, _vp, _vp],
    "dmd_scatt_compute_ddm": [_vp, _vp, _vp, _vp],
    "dmd_scatt_set_window": [_vp, _i, _i, _vp],
    "dmd_scatt_evolve_driver": [_vp, _d, _vp, _vp, _vp],
    "dmd_scatt_set_coherent": [_vp, _vp],
    "dmd_scatt_set_observable": [_vp, _vp, _vp, _i, _i, _d],
    "dmd_scatt_measure": [_vp, _d, _vp, _pd],
    "dmd_scatt_integrate_rk4": [_vp, _d, _d, _i64, _i, _vp, _vp],
    "dmd_scatt_integrate_rk4_batched": [_vp, _i, _d, _d, _i64, _i, _vp, _vp],
    "dmd_scatt_rkf45_apply": [_vp, _pd, _d, _pd, _d, _d, _d, _d, _vp, ctypes.POINTER(OdeStats)],
    "dmd_scatt_integrate_euler": [_vp, _d, _d, _i64, _vp],
    "dmd_scatt_compute_rhs": [_vp, _d, _vp, _vp],
    "dmd_scatt_vjp": [_vp, _d, _vp, _vp, _vp],
    "dmd_scatt_combine_P": [_vp, _i, _vp, _vp],
    "dmd_scatt_adjoint_rk4": [_vp, _d, _d, _i64, _i, _i, _vp, _vp, _i, _vp, _vp, _vp],
    "dmd_scatt_info": [_vp, _pi64, _pi64, _pd],
    "dmd_scatt_time_eval": [_vp, _d, _i, _pd, _pd],
    "dmd_scatt_time_eval_batched": [_vp, _d, _i, _pd],
    "dmd_synth_preset": [_i, _pSP],
    "dmd_synth_energies": [_pSP, _vp],
    "dmd_synth_kpairs": [_pSP, _vp, _vp, _pi64],
    "dmd_synth_P": [_pSP, _vp, _i64, _vp, _vp, _vp, _vp, _i],
    "dmd_synth_couplings": [_pSP, ctypes.c_uint64, ctypes.c_uint64, _vp, _vp, _vp],
    "dmd_synth_spin": [_pSP, _vp],
    "dmd_ldbd_write_synth": [ctypes.c_char_p, _pSP, _i],
    "dmd_ldbd_read_info": [_cs, _pLI],
    "dmd_ldbd_read_ek": [_cs, _pLI, _d, _vp],
    "dmd_ldbd_read_kvec": [_cs, _pLI, _vp],
    "dmd_ldbd_read_smat": [_cs, _pLI, _vp],
    "dmd_ldbd_read_lmat": [_cs, _pLI, _vp],
    "dmd_ldbd_read_Bso": [_cs, _pLI, _vp],
    "dmd_ldbd_read_HEzmat": [_cs, _pLI, _vp],
    "dmd_ldbd_read_Umat": [_cs, _pLI, _vp],
    "dmd_ldbd_read_imsig": [_cs, _pLI, _vp],
    "dmd_ldbd_read_R": [_cs, ctypes.POINTER(_i), _vp],
    "dmd_ldbd_read_kpairs": [_cs, _pLI, _i64, _i64, _vp, _vp],
    "dmd_ldbd_read_P": [_cs, _pLI, _cs, _i, _i64, _i64, _vp, _vp],
    "dmd_scatt_load_P": [_vp, _cs, _pLI, _cs, _i, _d, _i],
    "dmd_scatt_create_from_ldbd": [_cs, _cs, _d, _d, _i, ctypes.POINTER(_vp)],
    "dmd_scatt_load_sP": [_vp, _cs, _pLI, _cs, _d, _d],
    "dmd_scatt_threshold_P": [_vp, _d],
    "dmd_scatt_compress_P": [_vp, _pi64],
    "dmd_scatt_mask_pairs": [_vp, _vp],
    "dmd_scatt_set_semiclassical": [_vp, _i],
    "dmd_scatt_set_phenom_relax": [_vp, _d, _i, _i, _vp],
    "dmd_ldbd_write_sP": [_vp, _cs, _pLI, _cs, _d],
    "dmd_scatt_create_from_ldbd_sparse": [_cs, _cs, _d, _d, _i, ctypes.POINTER(_vp)],
    "dmd_scatt_create_from_ldbd_noP": [_cs, _d, _i, _i, ctypes.POINTER(_vp)],
    "dmd_scatt_upload_P_range": [_vp, _i64, _i64, _vp, _vp, _d, _i],
    "dmd_restart_read": [_cs, _i64, _i, _vp, _pd],
    "dmd_restart_write": [_cs, _i64, _i, _vp, _d],
}
_RESTYPE = {"dmd_scatt_last_error": ctypes.c_char_p}

_lib = None


def load():
    """Loads libdmd_scatt.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(there is no CPU fallback for the scattering operator)")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in PROTOTYPES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, ctypes.c_int)
    _lib = L
    return L


def last_error():
    return load().dmd_scatt_last_error().decode(errors="replace")


def check(rc, what=""):
    if rc != DMD_OK:
        raise DMDError(rc, f"{what}: {last_error()}")
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args), name)
