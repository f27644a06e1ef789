"""Host-side mirror of the reference's scattering operator interface.

``ElectronPhonon`` follows the members of the reference's ``electronphonon``
class that the dynamics driver touches (src_v4.5.7/ElectronPhonon.h:150-165,
src_v4.5.7/DensityMatrix_Dynamics.h:59-63,268): ``evolve``,
``evolve_driver``, ``compute_ddmdt_eq``, plus the loaders ``set_eph`` (P and
k-pairs, ElectronPhonon.cpp:3-13,182-233) and the e-i fold-in
``add_scatt_contrib`` (ElectronPhonon.cpp:44-113).  Every call goes through
the C-ABI of libdmd_scatt.so to the HIP kernels; there is no CPU path.

Arrays are numpy complex128/float64 (host) or torch tensors on the GPU; the
library detects device pointers and then works in place on them.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import SynthParams, call

try:  # torch is optional: only used when tensors are passed
    import torch
except Exception:  # pragma: no cover
    torch = None


def _is_tensor(a):
    return torch is not None and isinstance(a, torch.Tensor)


def _ptr(a, dtype):
    """(pointer, keepalive) for numpy arrays or torch tensors."""
    if a is None:
        return None, None
    if _is_tensor(a):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        want = torch.complex128 if dtype == np.complex128 else (torch.float64 if dtype == np.float64 else torch.uint64)
        if a.dtype != want:
            raise TypeError(f"tensor dtype {a.dtype}, expected {want}")
        return ctypes.c_void_p(a.data_ptr()), a
    arr = np.ascontiguousarray(a, dtype=dtype)
    return ctypes.c_void_p(arr.ctypes.data), arr


def _empty_like_state(ref, shape):
    if _is_tensor(ref):
        return torch.empty(shape, dtype=torch.complex128, device=ref.device)
    return np.empty(shape, np.complex128)


class ScatteringOperator:
    """One dmd_scatt handle: k-pairs + P resident in HBM."""

    def __init__(self, nk, nb, k1st, k2nd, e, nk_full=None, precision=_lib.DMD_PREC_FP64):
        self.nk, self.nb = int(nk), int(nb)
        self.nb_dm, self.b0 = self.nb, 0
        k1 = np.ascontiguousarray(k1st, dtype=np.uint64)
        k2 = np.ascontiguousarray(k2nd, dtype=np.uint64)
        self.n_pairs = len(k1)
        e = np.ascontiguousarray(e, dtype=np.float64).reshape(self.nk, self.nb)
        h = ctypes.c_void_p()
        call("dmd_scatt_create", self.nk, self.nb, self.n_pairs, k1.ctypes.data, k2.ctypes.data, e.ctypes.data,
             float(nk_full if nk_full is not None else nk), int(precision), ctypes.byref(h))
        self._h = h

    @classmethod
    def from_ldbd(cls, dirname, scatt="lindblad", scale_eph=1.0, scissor=0.0, precision=_lib.DMD_PREC_FP64,
                  Pin_is_sparse=False, eph_enable=True, scatt_enable=True):
        """The handle of an ldbd_data folder, set up as the reference's
        electronphonon sets itself up (ElectronPhonon.h:53-93): e-ph window of
        e_dm, stored pairs, P x scale_eph streamed from the files (the sparse
        sP files with Pin_is_sparse, ElectronPhonon.cpp:235-246), and the band
        window when the density matrix has more bands.  eph_enable=False
        reads no P (the e-i fold-in, load_P(impurity=1, add=False), sets it);
        scatt_enable=False keeps no stored pairs (coherent dynamics only)."""
        from . import ldbd
        info = ldbd.read_info(dirname)
        self = cls.__new__(cls)
        self.nk, self.nb = int(info.nk), int(info.nb_eph)
        self.nb_dm, self.b0 = int(info.nb_dm), int(info.b_bot_eph - info.b_bot_dm)
        self.n_pairs = int(info.nkpair)
        self.ldbd_info = info
        h = ctypes.c_void_p()
        if not scatt_enable or not eph_enable:
            call("dmd_scatt_create_from_ldbd_noP", str(dirname).encode(), float(scissor), int(precision),
                 int(not scatt_enable), ctypes.byref(h))
            if not scatt_enable:
                self.n_pairs = 0
        else:
            call("dmd_scatt_create_from_ldbd_sparse" if Pin_is_sparse else "dmd_scatt_create_from_ldbd",
                 str(dirname).encode(), scatt.encode(), float(scale_eph), float(scissor), int(precision),
                 ctypes.byref(h))
        self._h = h
        return self

    # -- lifetime ---------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _lib.load().dmd_scatt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- P ------------------------------------------------------------------------
    def upload_P(self, P1, P2, scale=1.0):
        p1, k1 = _ptr(P1, np.complex128)
        p2, k2 = _ptr(P2, np.complex128)
        call("dmd_scatt_upload_P", self._h, p1, p2, float(scale))

    def add_P(self, P1, P2, scale=1.0):
        p1, k1 = _ptr(P1, np.complex128)
        p2, k2 = _ptr(P2, np.complex128)
        call("dmd_scatt_add_P", self._h, p1, p2, float(scale))

    def upload_P_range(self, first, P1, P2, scale=1.0, add=False):
        """P of stored pairs [first, first + len(P1)) in file order."""
        p1, k1 = _ptr(P1, np.complex128)
        p2, k2 = _ptr(P2, np.complex128)
        count = (P1.numel() if _is_tensor(P1) else np.asarray(P1).size) // self.nb ** 4
        call("dmd_scatt_upload_P_range", self._h, int(first), int(count), p1, p2, float(scale), int(bool(add)))

    def load_P(self, dirname, info=None, scatt="lindblad", impurity=0, scale=1.0, add=False):
        """P (or P += with add) from the ldbd_P files, streamed into HBM
        (read_ldbd_eph, ElectronPhonon.cpp:209-233; impurity = iD + 1 for the
        e-i files of read_ldbd_imp_P, ElectronImpurity.cpp:3-35)."""
        from . import ldbd
        info = ldbd.read_info(dirname) if info is None else info
        call("dmd_scatt_load_P", self._h, str(dirname).encode(), ctypes.byref(info), scatt.encode(), int(impurity),
             float(scale), int(bool(add)))

    def load_sP(self, dirname, info=None, scatt="lindblad", scale=1.0, thr=1e-40):
        """P from the sparse files sP{1,2}_<scatt>[_hole]_{ns,s,i,j}.bin, each
        value x scale kept when |value x scale| > thr (ElectronPhonon.cpp:
        235-246), placed dense in the handle's layout."""
        from . import ldbd
        info = ldbd.read_info(dirname) if info is None else info
        call("dmd_scatt_load_sP", self._h, str(dirname).encode(), ctypes.byref(info), scatt.encode(), float(scale),
             float(thr))

    def compress_P(self):
        """Keep only the nonzero elements of P (dmd_scatt_compress_P); returns
        the number of stored values (padding included)."""
        n = ctypes.c_int64(0)
        call("dmd_scatt_compress_P", self._h, ctypes.byref(n))
        return n.value

    def threshold_P(self, thr):
        """alg_sparseP's sparsification (set_sparseP, ElectronPhonon.cpp:115-129):
        every stored element with |P| <= thr becomes 0."""
        call("dmd_scatt_threshold_P", self._h, float(thr))

    def set_semiclassical(self, on=True):
        """alg_semiclassical (DensityMatrix_Dynamics.h:249,277): the integrators
        and compute_rhs keep only the populations."""
        call("dmd_scatt_set_semiclassical", self._h, int(bool(on)))

    def set_phenom_relax(self, tau, b_start, b_end, f_eq):
        """alg_phenom_relax: -(rho - diag f_eq) / tau on density-matrix bands
        [b_start, b_end) in the integrators (tau <= 0: off)."""
        p, keep = _ptr(f_eq, np.float64)
        call("dmd_scatt_set_phenom_relax", self._h, float(tau), int(b_start), int(b_end), p)

    def mask_pairs(self, keep):
        """The k-pair filter of evolve (evolve_eph.cpp:52-58): P of the stored
        pairs with keep == 0 (file order) becomes 0."""
        k = np.ascontiguousarray(keep, dtype=np.uint8)
        if k.size != self.n_pairs:
            raise ValueError("keep needs one entry per stored pair")
        call("dmd_scatt_mask_pairs", self._h, k.ctypes.data)

    def write_sP(self, dirname, info=None, scatt="lindblad", thr=1e-40):
        """The stored P as sparse files (write_smat, ElectronPhonon.cpp:126-128);
        a matrix whose _s file exists is left alone."""
        from . import ldbd
        info = ldbd.read_info(dirname) if info is None else info
        call("dmd_ldbd_write_sP", self._h, str(dirname).encode(), ctypes.byref(info), scatt.encode(), float(thr))

    def add_ee_model(self, k1st, k2nd, U, kvec, f, model, scale=1.0, add=True):
        """The model e-e P at the equilibrium occupations f (nk, nb), generated
        on the device and added (add) or set: add_scatt_contrib("ee") with
        elecelec_model::calc_P at dm = null, no exchange (ElecElec_Model.h:
        185-262), as eeMode = Pee_fixed_at_eq keeps it; scale = scale_ee.
        model: eimp_model(...) with degauss = degauss_ee and the k mesh."""
        k1, k2, U, kv, f = _ee_args(self, k1st, k2nd, U, kvec, f, model)
        call("dmd_scatt_add_ee_model", self._h, k1.ctypes.data, k2.ctypes.data, U.ctypes.data, kv.ctypes.data,
             f.ctypes.data, ctypes.byref(model), float(scale), int(bool(add)))

    def ee_vq_cells(self, U, kvec, f, model):
        """The static RPA v(q) of every k-mesh cell that is a difference of two k
        points (dmd_scatt_ee_vq_cells); f (nk, nb) at mu.  Returns an array of
        kmesh[0] * kmesh[1] * kmesh[2] values."""
        U = np.ascontiguousarray(U, dtype=np.complex128)
        kv = np.ascontiguousarray(kvec, dtype=np.float64)
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros(int(np.prod(list(model.kmesh))))
        call("dmd_scatt_ee_vq_cells", self._h, U.ctypes.data, kv.ctypes.data, f.ctypes.data, ctypes.byref(model),
             out.ctypes.data)
        return out

    def add_ee_model_state(self, k1st, k2nd, U, kvec, f, model, dm=None, dm1=None, vq_cell=None, scale=1.0,
                           add=True):
        """The e-e P at the density matrix dm (nk, nb^2, e-ph window,
        Schroedinger picture) with dm1 = 1 - dm as set_oneminusdm forms it
        (default), or at f when dm is None: the Pee_update rebuild
        (dmd_scatt_add_ee_model_state).  vq_cell: ee_vq_cells(...) for the
        static RPA."""
        k1, k2, U, kv, f = _ee_args(self, k1st, k2nd, U, kvec, f, model)
        pdm = pdm1 = None
        if dm is not None:
            dm = np.ascontiguousarray(dm, dtype=np.complex128).reshape(self.nk, self.nb * self.nb)
            if dm1 is None:
                dm1 = -dm.copy()
                dm1[:, ::self.nb + 1] += 1.0
            dm1 = np.ascontiguousarray(dm1, dtype=np.complex128).reshape(self.nk, self.nb * self.nb)
            pdm, pdm1 = dm.ctypes.data, dm1.ctypes.data
        vc = None if vq_cell is None else np.ascontiguousarray(vq_cell, dtype=np.float64)
        call("dmd_scatt_add_ee_model_state", self._h, k1.ctypes.data, k2.ctypes.data, U.ctypes.data, kv.ctypes.data,
             f.ctypes.data, pdm, pdm1, None if vc is None else vc.ctypes.data, ctypes.byref(model), float(scale),
             int(bool(add)))

    def _dP_state(self, dm, dm1):
        dm = np.ascontiguousarray(dm, dtype=np.complex128).reshape(self.nk, self.nb * self.nb)
        if dm1 is None:
            dm1 = -dm.copy()
            dm1[:, ::self.nb + 1] += 1.0
        return dm, np.ascontiguousarray(dm1, dtype=np.complex128).reshape(self.nk, self.nb * self.nb)

    def add_ee_dP(self, k1st, k2nd, U, kvec, f_eq, model, dm, dm1=None, vq_cell=None, scale=1.0, add=True):
        """P (+)= scale x dP, the alg_linearize_dPee change of the e-e P at f_eq
        towards the state dm (nk, nb^2, Schroedinger picture; dm1 = 1 - dm by
        default): the dP1ee, dP2ee tensors (dmd_scatt_add_ee_dP)."""
        k1, k2, U, kv, f = _ee_args(self, k1st, k2nd, U, kvec, f_eq, model)
        dm, dm1 = self._dP_state(dm, dm1)
        vc = None if vq_cell is None else np.ascontiguousarray(vq_cell, dtype=np.float64)
        call("dmd_scatt_add_ee_dP", self._h, k1.ctypes.data, k2.ctypes.data, U.ctypes.data, kv.ctypes.data,
             f.ctypes.data, dm.ctypes.data, dm1.ctypes.data, None if vc is None else vc.ctypes.data,
             ctypes.byref(model), float(scale), int(bool(add)))

    def set_ee_dP(self, k1st, k2nd, U, kvec, f_eq, model, dm, dm1=None, vq_cell=None):
        """Builds dP at the state dm and keeps its row term, which every later
        evaluation adds with its phase (dmd_scatt_set_ee_dP); dm None clears it."""
        if dm is None:
            call("dmd_scatt_set_ee_dP", self._h, None, None, None, None, None, None, None, None, None)
            return
        k1, k2, U, kv, f = _ee_args(self, k1st, k2nd, U, kvec, f_eq, model)
        dm, dm1 = self._dP_state(dm, dm1)
        vc = None if vq_cell is None else np.ascontiguousarray(vq_cell, dtype=np.float64)
        call("dmd_scatt_set_ee_dP", self._h, k1.ctypes.data, k2.ctypes.data, U.ctypes.data, kv.ctypes.data,
             f.ctypes.data, dm.ctypes.data, dm1.ctypes.data, None if vc is None else vc.ctypes.data,
             ctypes.byref(model))

    def get_ee_dP(self):
        """The kept row term prefac/2 (D + D^+), (nk, nb^2) complex."""
        out = np.zeros((self.nk, self.nb * self.nb), np.complex128)
        call("dmd_scatt_get_ee_dP", self._h, out.ctypes.data)
        return out

    def add_Lsc_ee(self, k1st, k2nd, U, kvec, f_eq, model, scale=1.0, extra=0.0, coef34=1.0, dP_terms=True,
                   vq_cell=None):
        """The e-e terms of add_Lsc_contrib("ee") beyond the P1/P2 form (after
        set_Lsc): extra x Lscij/Lscji of P_ee(f_eq) (1 under ee_antisymmetry),
        and with dP_terms the P3/P4 (coef34 = 1: the reference's sign, -1: the
        exact Jacobian) and P5/P6 terms (dmd_scatt_add_Lsc_ee)."""
        k1, k2, U, kv, f = _ee_args(self, k1st, k2nd, U, kvec, f_eq, model)
        vc = None if vq_cell is None else np.ascontiguousarray(vq_cell, dtype=np.float64)
        call("dmd_scatt_add_Lsc_ee", self._h, k1.ctypes.data, k2.ctypes.data, U.ctypes.data, kv.ctypes.data,
             f.ctypes.data, None if vc is None else vc.ctypes.data, ctypes.byref(model), float(scale), float(extra),
             float(coef34), int(bool(dP_terms)))

    def download_P(self, first=0, count=None):
        count = self.n_pairs - first if count is None else count
        n4 = self.nb ** 4
        P1 = np.empty((count, n4), np.complex128)
        P2 = np.empty((count, n4), np.complex128)
        call("dmd_scatt_download_P", self._h, int(first), int(count), P1.ctypes.data, P2.ctypes.data)
        return P1, P2

    def copy_P_from(self, src, scale=1.0, add=False):
        """P = scale x src's stored P (same pairs), converted on the device
        between precisions (dmd_scatt_copy_P)."""
        call("dmd_scatt_copy_P", self._h, src._h, float(scale), int(bool(add)))

    def generate_P(self, params):
        call("dmd_scatt_generate_P", self._h, ctypes.byref(params))

    def add_eimp_model(self, k1st, k2nd, U, kvec, model, scale=1.0, add=True, f=None):
        """Model ionized-impurity P of every stored pair generated on the device
        and added (add) or set: elecimp_model::calc_P through
        add_scatt_contrib("eimp") (ElecImp_Model.h:185-213, ElectronPhonon.cpp:
        44-113), scale = scale_ei.  U: (nk, nb_wannier, nb) complex (the e-ph
        columns of ldbd_Umat.bin), kvec: (nk, 3), model: an EimpModel
        (eimp_model()); f: (nk, nb) occupations at mu, for the RPA screening."""
        k1 = np.ascontiguousarray(k1st, dtype=np.uint64)
        k2 = np.ascontiguousarray(k2nd, dtype=np.uint64)
        U = np.ascontiguousarray(U, dtype=np.complex128)
        kv = np.ascontiguousarray(kvec, dtype=np.float64)
        if len(k1) != self.n_pairs or len(k2) != self.n_pairs:
            raise ValueError("k1st/k2nd need one entry per stored pair")
        if U.shape != (self.nk, model.nb_wannier, self.nb) or kv.shape != (self.nk, 3):
            raise ValueError(f"U must be (nk, nb_wannier, nb) = {(self.nk, model.nb_wannier, self.nb)} and kvec (nk, 3)")
        fp = None
        if f is not None:
            f = np.ascontiguousarray(f, dtype=np.float64)
            if f.shape != (self.nk, self.nb):
                raise ValueError("f must be (nk, nb)")
            fp = f.ctypes.data
        call("dmd_scatt_add_eimp_model", self._h, k1.ctypes.data, k2.ctypes.data, U.ctypes.data, kv.ctypes.data, fp,
             ctypes.byref(model), float(scale), int(bool(add)))

    # -- flags / eq ----------------------------------------------------------------
    def set_flags(self, expt=True, schrodinger=False, kernel=_lib.DMD_KERNEL_AUTO):
        call("dmd_scatt_set_flags", self._h, int(bool(expt)), int(bool(schrodinger)), int(kernel))

    def set_split(self, f):
        """Occupations f (nk x nb) of the split rho = diag(f) + drho used by the
        reduced-precision forms (None: f = 0)."""
        if f is None:
            call("dmd_scatt_set_split", self._h, None)
            return
        p, keep = _ptr(f, np.float64)
        call("dmd_scatt_set_split", self._h, p)

    def set_Lsc(self, f_eq):
        """Linearize at f_eq (nk x nb): set_Lsc, ElectronPhonon_linear.cpp:3-36.
        The stored P becomes Lscij | Lscji; evaluations are evolve_linear."""
        p, keep = _ptr(f_eq, np.float64)
        call("dmd_scatt_set_Lsc", self._h, p)

    def set_ddmdt_eq(self, ddmdt_eq):
        p, keep = _ptr(ddmdt_eq, np.complex128)
        call("dmd_scatt_set_ddmdt_eq", self._h, p)

    def compute_ddmdt_eq(self, f):
        p, keep = _ptr(f, np.float64)
        call("dmd_scatt_compute_ddmdt_eq", self._h, p)

    # -- hot path ---------------------------------------------------------------
    def eval(self, t, rho, rho1=None, out=None):
        """dRho/dt|scatt at time t for rho (nk x nb^2), evolve_eph.cpp:34-166."""
        shape = (self.nk, self.nb * self.nb)
        pr, kr = _ptr(rho, np.complex128)
        p1, k1 = _ptr(rho1, np.complex128)
        if out is None:
            out = _empty_like_state(rho, shape)
        po, ko = _ptr(out, np.complex128)
        call("dmd_scatt_eval", self._h, float(t), pr, p1, po)
        if not _is_tensor(out) and ko is not out:
            out[...] = ko.reshape(out.shape)
        return out

    def eval_noalloc(self, t, rho, out=None):
        """dmd_scatt_eval_noalloc: eval with no device allocation; rho and out
        are CUDA tensors (nk x nb^2 complex128)."""
        if not _is_tensor(rho) or not rho.is_cuda:
            raise TypeError("eval_noalloc takes device tensors")
        if out is None:
            out = torch.empty((self.nk, self.nb * self.nb), dtype=torch.complex128, device=rho.device)
        pr, kr = _ptr(rho, np.complex128)
        po, ko = _ptr(out, np.complex128)
        call("dmd_scatt_eval_noalloc", self._h, float(t), pr, po)
        return out

    def eval_batched(self, t, rho_B, out=None):
        B = rho_B.shape[0]
        if out is None:
            out = _empty_like_state(rho_B, (B, self.nk, self.nb * self.nb))
        pr, kr = _ptr(rho_B, np.complex128)
        po, ko = _ptr(out, np.complex128)
        call("dmd_scatt_eval_batched", self._h, float(t), int(B), pr, po)
        return out

    def compute_ddm(self, rho, rho1=None, ddm=None):
        if ddm is None:
            ddm = np.zeros((self.nk, self.nb * self.nb), np.complex128)
        pr, kr = _ptr(rho, np.complex128)
        p1, k1 = _ptr(rho1, np.complex128)
        pd, kd = _ptr(ddm, np.complex128)
        call("dmd_scatt_compute_ddm", self._h, pr, p1, pd)
        if not _is_tensor(ddm) and kd is not ddm:
            ddm[...] = kd.reshape(ddm.shape)
        return ddm

    # -- driver pieces --------------------------------------------------------------
    def set_window(self, nb_dm, b0, e_dm):
        p, keep = _ptr(e_dm, np.float64)
        call("dmd_scatt_set_window", self._h, int(nb_dm), int(b0), p)
        self.nb_dm, self.b0 = int(nb_dm), int(b0)

    def evolve_driver(self, t, dm_expand, dm1_expand=None, out=None):
        shape = (self.nk, self.nb_dm * self.nb_dm)
        pr, kr = _ptr(dm_expand, np.complex128)
        p1, k1 = _ptr(dm1_expand, np.complex128)
        if out is None:
            out = _empty_like_state(dm_expand, shape)
        po, ko = _ptr(out, np.complex128)
        call("dmd_scatt_evolve_driver", self._h, float(t), pr, p1, po)
        return out

    def set_coherent(self, H):
        p, keep = _ptr(H, np.complex128)
        call("dmd_scatt_set_coherent", self._h, p)

    def set_observable(self, s, f, i0, i1, prefac):
        ps, ks = _ptr(s, np.complex128)
        pf, kf = _ptr(f, np.float64)
        call("dmd_scatt_set_observable", self._h, ps, pf, int(i0), int(i1), float(prefac))

    def measure(self, t, rho):
        pr, kr = _ptr(rho, np.complex128)
        v = ctypes.c_double()
        call("dmd_scatt_measure", self._h, float(t), pr, ctypes.byref(v))
        return v.value

    def integrate_rk4(self, rho0, t0, dt, n_steps, report_every=0):
        """Fixed-step RK4 on the device; returns (rho(t0 + n dt), observable samples)."""
        rho = np.array(rho0, dtype=np.complex128, copy=True).reshape(self.nk, self.nb_dm * self.nb_dm) \
            if not _is_tensor(rho0) else rho0.clone()
        nobs = (n_steps // report_every + 1) if report_every > 0 else 0
        obs = np.zeros(max(nobs, 1))
        pr, kr = _ptr(rho, np.complex128)
        call("dmd_scatt_integrate_rk4", self._h, float(t0), float(dt), int(n_steps), int(report_every), pr,
             obs.ctypes.data)
        return rho, obs[:nobs]

    def rkf45_apply(self, rho, t, t1, h_step, hmin=0.0, hmax=None, eps_abs=1e-8, eps_rel=0.0):
        """One gsl_odeiv2_driver_apply of evolve_gsl (rkf45, standard control)
        from t to t1; returns (rho(t1), t1, next step suggestion, stats)."""
        y = np.array(rho, dtype=np.complex128, copy=True).reshape(self.nk, self.nb_dm * self.nb_dm) \
            if not _is_tensor(rho) else rho.clone()
        tt = ctypes.c_double(float(t))
        hh = ctypes.c_double(float(h_step))
        st = _lib.OdeStats()
        py, ky = _ptr(y, np.complex128)
        call("dmd_scatt_rkf45_apply", self._h, ctypes.byref(tt), float(t1), ctypes.byref(hh), float(hmin),
             float(hmax if hmax is not None else t1 - t), float(eps_abs), float(eps_rel), py, ctypes.byref(st))
        return y, tt.value, hh.value, st.as_dict()

    def compute_rhs(self, t, rho, out=None):
        """The whole right-hand side at t (dm_dynamics::compute)."""
        shape = (self.nk, self.nb_dm * self.nb_dm)
        if out is None:
            out = _empty_like_state(rho, shape)
        pr, kr = _ptr(rho, np.complex128)
        po, ko = _ptr(out, np.complex128)
        call("dmd_scatt_compute_rhs", self._h, float(t), pr, po)
        return out

    def integrate_euler(self, rho, t0, dt, n_steps):
        """evolve_euler: n_steps of rho += dt (ddmdt + ddmdt^+)/2 from t0."""
        y = np.array(rho, dtype=np.complex128, copy=True).reshape(self.nk, self.nb_dm * self.nb_dm) \
            if not _is_tensor(rho) else rho.clone()
        py, ky = _ptr(y, np.complex128)
        call("dmd_scatt_integrate_euler", self._h, float(t0), float(dt), int(n_steps), py)
        return y

    def set_rhs_hook(self, fn):
        """fn(t, rho) before every right-hand side the integrators evaluate
        (dmd_scatt_set_rhs_hook; update_scatt_inside in compute): rho is a
        host copy (nk, nb_dm^2) of the stage state; fn may change P (e.g. an
        e-e rebuild at that state); an exception stops the integrator.  None
        removes the hook."""
        if fn is None:
            call("dmd_scatt_set_rhs_hook", self._h, None, None)
            self._hook = None
            return
        shape = (self.nk, self.nb_dm * self.nb_dm)

        def tramp(ctx, t, rho):
            try:
                arr = np.ctypeslib.as_array(rho, shape=(shape[0] * shape[1] * 2,)).view(np.complex128).reshape(shape)
                fn(t, arr.copy())
                return 0
            except Exception as e:  # reported through the integrator's error
                self._hook_error = e
                return 1

        cb = _RHS_HOOK(tramp)
        call("dmd_scatt_set_rhs_hook", self._h, ctypes.cast(cb, ctypes.c_void_p), None)
        self._hook = cb  # the C side keeps only the pointer

    def integrate_rk4_batched(self, rho0_B, t0, dt, n_steps, report_every=0):
        B = rho0_B.shape[0]
        rho = np.array(rho0_B, dtype=np.complex128, copy=True).reshape(B, self.nk, self.nb_dm * self.nb_dm)
        nobs = (n_steps // report_every + 1) if report_every > 0 else 0
        obs = np.zeros((B, max(nobs, 1)))
        call("dmd_scatt_integrate_rk4_batched", self._h, int(B), float(t0), float(dt), int(n_steps),
             int(report_every), rho.ctypes.data, obs.ctypes.data)
        return rho, obs[:, :nobs]

    # -- introspection --------------------------------------------------------------
    def info(self):
        n = ctypes.c_int64()
        b = ctypes.c_int64()
        ms = ctypes.c_double()
        call("dmd_scatt_info", self._h, ctypes.byref(n), ctypes.byref(b), ctypes.byref(ms))
        return {"n_pairs": n.value, "p_bytes": b.value, "last_kernel_ms": ms.value}

    def time_eval(self, t, n):
        k = ctypes.c_double()
        e = ctypes.c_double()
        call("dmd_scatt_time_eval", self._h, float(t), int(n), ctypes.byref(k), ctypes.byref(e))
        return k.value, e.value

    # -- adjoint ---------------------------------------------------------------
    def vjp(self, t, rho, v, out=None):
        """(d ddmdt / d rho)^T v at the Hermitian state rho (FP64, nb <= 2)."""
        if out is None:
            out = _empty_like_state(rho, (self.nk, self.nb * self.nb))
        pr, kr = _ptr(rho, np.complex128)
        pv, kv = _ptr(v, np.complex128)
        po, ko = _ptr(out, np.complex128)
        call("dmd_scatt_vjp", self._h, float(t), pr, pv, po)
        return out

    def combine_P(self, comps, weights):
        """P = sum_c weights[c] P of comps[c] (FP64 handles of the same pairs)."""
        arr = (ctypes.c_void_p * len(comps))(*[c._h.value for c in comps])
        w = np.ascontiguousarray(weights, np.float64)
        call("dmd_scatt_combine_P", self._h, len(comps), ctypes.cast(arr, ctypes.c_void_p), w.ctypes.data)

    def adjoint_rk4(self, rho0, t0, dt, n_steps, report_every, checkpoint_every, dL_dobs, comps=()):
        """(dL/dtheta over comps, dL/drho0) of a loss with gradient dL_dobs on
        the observable reports of integrate_rk4(rho0, t0, dt, n_steps, report_every)."""
        r0 = np.ascontiguousarray(rho0, np.complex128).reshape(self.nk, self.nb * self.nb)
        dobs = None if dL_dobs is None else np.ascontiguousarray(dL_dobs, np.float64)
        arr = (ctypes.c_void_p * max(1, len(comps)))(*[c._h.value for c in comps])
        dth = np.zeros(max(1, len(comps)))
        drho = np.zeros_like(r0)
        call("dmd_scatt_adjoint_rk4", self._h, float(t0), float(dt), int(n_steps), int(report_every),
             int(checkpoint_every), r0.ctypes.data, None if dobs is None else dobs.ctypes.data, len(comps),
             ctypes.cast(arr, ctypes.c_void_p), dth.ctypes.data, drho.ctypes.data)
        return dth[:len(comps)], drho

    def time_eval_batched(self, t, n):
        """ms per batched evaluation of the states of the last eval_batched."""
        k = ctypes.c_double()
        call("dmd_scatt_time_eval_batched", self._h, float(t), int(n), ctypes.byref(k))
        return k.value


class ElectronPhonon(ScatteringOperator):
    """Drop-in shaped like the reference's electronphonon for the driver.

    evolve(t, dm, dm1)          -> ddmdt        (evolve_eph.cpp:34-166)
    evolve_driver(t, dm, dm1)   -> ddmdt_expand (evolve_eph.cpp:21-32)
    compute_ddmdt_eq(f)                          (evolve_eph.cpp:3-19)
    set_eph(P1, P2, scale_eph)                   (ElectronPhonon.cpp:209-233)
    add_scatt_contrib(P1, P2, factor)            (ElectronPhonon.cpp:44-113, "eimp")
    """

    def __init__(self, nk, nb, k1st, k2nd, e, nk_full, alg_expt=True, alg_ddmdteq=False,
                 picture="interaction", precision=_lib.DMD_PREC_FP64):
        super().__init__(nk, nb, k1st, k2nd, e, nk_full, precision)
        self.prefac_eph = 2 * np.pi / nk_full
        self.alg_ddmdteq = alg_ddmdteq
        self.set_flags(expt=alg_expt, schrodinger=(picture == "schrodinger"))

    def set_eph(self, P1, P2, scale_eph=1.0):
        self.upload_P(P1, P2, scale_eph)

    def add_scatt_contrib(self, P1add, P2add, factor):
        self.add_P(P1add, P2add, factor)

    def evolve(self, t, dm, dm1=None):
        return self.eval(t, dm, dm1)


_RHS_HOOK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double))


def _ee_args(op, k1st, k2nd, U, kvec, f, model):
    k1 = np.ascontiguousarray(k1st, dtype=np.uint64)
    k2 = np.ascontiguousarray(k2nd, dtype=np.uint64)
    U = np.ascontiguousarray(U, dtype=np.complex128)
    kv = np.ascontiguousarray(kvec, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    if len(k1) != op.n_pairs or len(k2) != op.n_pairs:
        raise ValueError("k1st/k2nd need one entry per stored pair")
    if U.shape != (op.nk, model.nb_wannier, op.nb) or kv.shape != (op.nk, 3) or f.shape != (op.nk, op.nb):
        raise ValueError("U must be (nk, nb_wannier, nb), kvec (nk, 3) and f (nk, nb)")
    return k1, k2, U, kv, f


SCREENING = {"unscreened": _lib.DMD_SCR_UNSCREENED, "debye": _lib.DMD_SCR_DEBYE,
             "Bechstedt": _lib.DMD_SCR_BECHSTEDT, "RPA": _lib.DMD_SCR_RPA}


def eimp_model(screening, n_free, temperature, cell_size, GGT, nb_wannier, Z=1.0, n_ionized=None,
               epsilon_background=1.0, meff=1.0, degauss=None, ethr=None, scatt="lindblad",
               detail_balance=False, kmesh=(0, 0, 0), nk_full=0.0, exchange=False):
    """The dmd_eimp_model of one ionized impurity (atomic units throughout):
    elecimp_model's prefactors (ElecImp_Model.h:22-40: prefac_A = Z sqrt(n_i
    V_cell), omegaL = 1e-4 T) and coulomb_model's static screening
    (Coulomb_Model.h:419-425,554-575: prefac_vq = 4 pi / (eps V_cell); Debye
    q^2 = 4 pi n_free / (eps T); Thomas-Fermi q^2 = 6 pi n_free / (eps E_F)
    with k_F = (3 pi^2 n_free)^(1/3), E_F = k_F^2 / 2 meff for Bechstedt).
    n_free and n_ionized are in bohr^-3; degauss is the delta width of the e-i
    term, ethr the |de| cut-off (degauss x ndegauss of the e-ph set).  RPA sums
    over the k mesh (kmesh, nk_full; Coulomb_Model.h:690-744) at the
    occupations given to add_eimp_model."""
    if screening not in SCREENING:
        raise ValueError(f"screening {screening!r}: one of {sorted(SCREENING)} (the static forms)")
    if screening in ("debye", "Bechstedt") and not n_free > 0:
        raise ValueError("nfreetot must be postive")
    n_ionized = n_free if n_ionized is None else n_ionized
    if not n_ionized > 0:
        raise ValueError("eip.ni_ionized[iD] must be postive")
    if epsilon_background < 1:
        raise ValueError("epsilon_background < 1 is not right")
    m = _lib.EimpModel()
    m.screening = SCREENING[screening]
    m.lindblad = 1 if scatt == "lindblad" else 0
    m.detail_balance = int(bool(detail_balance))
    m.nb_wannier = int(nb_wannier)
    m.prefac_A = Z * np.sqrt(n_ionized * cell_size)
    m.prefac_vq = 4 * np.pi / epsilon_background / cell_size
    m.prefac_vq_bare = 4 * np.pi / cell_size
    n = max(n_free, 0.0)
    kF = (3 * np.pi * np.pi * n) ** (1. / 3.)
    EF = kF * kF / 2. / meff
    qTF2 = 6 * np.pi * n / epsilon_background / EF if n > 0 else 0.0
    m.qscr2_debye = 4 * np.pi * n / epsilon_background / temperature
    if screening == "Bechstedt":
        m.fac0 = 1. / (epsilon_background - 1) if epsilon_background > 1 else np.inf
        m.fac2 = 1. / qTF2
        m.fac4 = 3. / 4. / (kF * kF) / qTF2
    m.degauss = float(degauss)
    m.ethr = float(ethr)
    m.omegaL = 1e-4 * temperature
    m.temperature = float(temperature)
    for i, v in enumerate(np.asarray(GGT, dtype=np.float64).reshape(9)):
        m.GGT[i] = v
    for i in range(3):
        m.kmesh[i] = int(kmesh[i])
    m.nk_full = float(nk_full)
    m.exchange = int(bool(exchange))  # e-e only: ee_antisymmetry
    return m


def reciprocal_metric(R):
    """GGT = G G^T with G = 2 pi R^-1 (the rows of R are the lattice vectors;
    lattice.h:40-41)."""
    Ri = np.linalg.inv(np.asarray(R, dtype=np.float64))
    return 4 * np.pi * np.pi * Ri @ Ri.T


def synth_params(preset):
    """dmd_synth_params of a named workload preset (0 tiny, 1..5 benchmark shapes)."""
    p = SynthParams()
    call("dmd_synth_preset", int(preset), ctypes.byref(p))
    return p


def synth_energies(p):
    e = np.empty((p.nk, p.nb), np.float64)
    call("dmd_synth_energies", ctypes.byref(p), e.ctypes.data)
    return e


def synth_kpairs(p):
    n = ctypes.c_int64(0)
    call("dmd_synth_kpairs", ctypes.byref(p), None, None, ctypes.byref(n))
    k1 = np.empty(n.value, np.uint64)
    k2 = np.empty(n.value, np.uint64)
    call("dmd_synth_kpairs", ctypes.byref(p), k1.ctypes.data, k2.ctypes.data, ctypes.byref(n))
    return k1, k2


def synth_P(p, e, k1, k2, nthreads=0):
    n = len(k1)
    n4 = p.nb ** 4
    P1 = np.empty((n, n4), np.complex128)
    P2 = np.empty((n, n4), np.complex128)
    e = np.ascontiguousarray(e, np.float64)
    k1 = np.ascontiguousarray(k1, np.uint64)
    k2 = np.ascontiguousarray(k2, np.uint64)
    call("dmd_synth_P", ctypes.byref(p), e.ctypes.data, n, k1.ctypes.data, k2.ctypes.data, P1.ctypes.data,
         P2.ctypes.data, int(nthreads))
    return P1, P2


def synth_couplings(p, k, kp):
    n2 = p.nb * p.nb
    omega = np.empty(p.n_modes)
    gm = np.empty((p.n_modes, n2), np.complex128)
    gp = np.empty((p.n_modes, n2), np.complex128)
    call("dmd_synth_couplings", ctypes.byref(p), int(k), int(kp), omega.ctypes.data, gm.ctypes.data, gp.ctypes.data)
    return omega, gm, gp


def synth_spin(p):
    s3 = np.empty((p.nk, 3, p.nb * p.nb), np.complex128)
    call("dmd_synth_spin", ctypes.byref(p), s3.ctypes.data)
    return s3


def synth_initial_state(p, e, s3, axis=2, amp=1e-3, mu=-0.1 / 27.21138505):
    """rho0 = f + (df/de) b S_axis, the first-order state of
    electron::compute_dm_Bpert_1st (src_v4.5.7/electron.cpp:62-98, degenerate
    limit f(f-1)/T) with b set so that Tr(S_axis drho) / sum Tr f = amp; f is
    Fermi-Dirac at mu (default CBM - 0.1 eV).  Returns (rho0, f)."""
    nk, nb = e.shape
    T = p.temperature
    f = 1.0 / (np.exp((e - mu) / T) + 1.0)
    sa = s3[:, axis].reshape(nk, nb, nb)
    de = e[:, :, None] - e[:, None, :]
    fd = f[:, :, None] - f[:, None, :]
    deg = np.abs(de) < 1e-8
    favg = 0.5 * (f[:, :, None] + f[:, None, :])
    dfde = np.where(deg, favg * (favg - 1) / T, fd / np.where(deg, 1.0, de))
    d1 = dfde * sa  # drho for b = 1
    s_per_b = np.real(np.einsum("kij,kji->", sa, d1))
    b = amp * f.sum() / s_per_b
    rho = (np.einsum("ki,ij->kij", f, np.eye(nb)) + b * d1).reshape(nk, nb * nb)
    return rho.astype(np.complex128), f


def synth_batch_states(p, e, s3, B):
    """B initial states of the batched workload: spin along z, x, y in turn,
    imbalance amplitudes log-uniform in [1e-4, 1e-2] (trajectory 0: z, 1e-3).
    Returns (B x nk x nb^2 states, f)."""
    states = []
    f = None
    for b in range(B):
        axis = (2, 0, 1)[b % 3]
        amp = 1e-3 if b == 0 else 10.0 ** (-4 + 2 * ((b * 0.6180339887) % 1.0))
        rho, f = synth_initial_state(p, e, s3, axis, amp)
        states.append(rho)
    return np.stack(states), f


def write_ldbd(dirname, p, nthreads=0):
    call("dmd_ldbd_write_synth", str(dirname).encode(), ctypes.byref(p), int(nthreads))
