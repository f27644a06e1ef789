"""dmd_scatt_eval_noalloc (the allocation-free evaluation, scratch in one
static device array per GPU) against dmd_scatt_eval on the same handle: two
calls in a row, the second on a smaller operator, for every contraction
kernel; and the refusals (host pointers, an operator larger than the
array)."""
import numpy as np
import pytest

from helpers import FS, random_P, random_energies, random_pairs, random_rho, rel_err

pytestmark = pytest.mark.gpu


def _mods():
    import torch
    from denmat_dynm_development_amd import _lib, scatt
    return torch, _lib, scatt


def _case(nk, nb, seed):
    rng = np.random.default_rng(seed)
    k1, k2 = random_pairs(nk, 7, rng, diag=True)
    perm = rng.permutation(len(k1))
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    rho = random_rho(nk, nb, rng)
    return k1[perm], k2[perm], e, P1[perm], P2[perm], rho


# (nb, kernel, precision): the tile kernel (nb = 2), the single-read and row
# kernels, the nb = 4 panels in FP64 and complex64
CASES = [(2, 0, 0), (2, 2, 0), (2, 1, 0), (1, 0, 0), (3, 0, 0), (4, 0, 0), (4, 0, 1)]


@pytest.mark.parametrize("nb,kernel,prec", CASES)
def test_noalloc_matches_eval_twice_then_smaller(nb, kernel, prec):
    torch, _lib, scatt = _mods()
    print(f"case noalloc nb={nb} kernel={kernel} prec={prec}", flush=True)
    for call_no, nk in enumerate((900, 230)):
        k1, k2, e, P1, P2, rho = _case(nk, nb, 10 * nb + kernel + call_no)
        with scatt.ScatteringOperator(nk, nb, k1, k2, e, precision=prec) as op:
            op.upload_P(P1, P2)
            op.set_flags(expt=True, kernel=kernel)
            if prec:
                op.set_split(np.full((nk, nb), 0.3))
            for t_fs in (0.0, 53.0):
                rd = torch.from_numpy(rho).cuda()
                ref = op.eval(t_fs * FS, rd).cpu().numpy()  # also brings the reduced form's sums up to date
                got = op.eval_noalloc(t_fs * FS, rd).cpu().numpy()
                got2 = op.eval_noalloc(t_fs * FS, rd).cpu().numpy()
                err = rel_err(got, ref)
                print(f"  call {call_no} nk={nk} t={t_fs} fs: rel err {err:.2e}", flush=True)
                tol = 0.0 if kernel == 1 or nb in (3,) else 1e-13
                assert err <= tol, err
                assert rel_err(got2, got) <= tol


def test_noalloc_refusals():
    torch, _lib, scatt = _mods()
    print("case noalloc refusals", flush=True)
    nk, nb = 64, 2
    k1, k2, e, P1, P2, rho = _case(nk, nb, 5)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        out = np.empty_like(rho)
        import ctypes
        rc = _lib.load().dmd_scatt_eval_noalloc(op._h, 0.0, ctypes.c_void_p(rho.ctypes.data),
                                                ctypes.c_void_p(out.ctypes.data))
        assert rc == _lib.DMD_ERR_ARG and "device memory" in _lib.last_error()
    # an operator whose rotated state alone exceeds the 16 MiB array
    nk, nb = 70000, 4
    k1 = np.arange(0, nk - 1, dtype=np.uint64)
    k2 = k1 + 1
    e = np.zeros((nk, nb))
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(np.zeros((len(k1), nb ** 4), np.complex128), np.zeros((len(k1), nb ** 4), np.complex128))
        r = torch.zeros((nk, nb * nb), dtype=torch.complex128, device="cuda")
        with pytest.raises(_lib.DMDError) as ex:
            op.eval_noalloc(0.0, r)
        assert ex.value.code == _lib.DMD_ERR_ARG and "scratch" in str(ex.value)
