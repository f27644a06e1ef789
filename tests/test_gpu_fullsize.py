"""The operator at the workload's own sizes.

nb = 2, nk = 4096 (964,065 stored pairs, the reference-size case): P generated
on the device against the host generator, and evaluation and RK4 against the
oracle on all pairs.  nb = 2, nk = 65536 (58.2 M pairs, 29.8 GB of P) and
nb = 4, nk = 32768 (18.1 M pairs, power-law partners, 148 GB of P), too large
for the oracle in a test, through properties that hold at any size:
- the map is exactly quadratic in rho: F(r + 2x) + F(r - 2x) - 2F(r) equals
  4 (F(r + x) + F(r - x) - 2F(r));
- ddm is Hermitian, and sum_k Tr ddm_k = 0 (every stored pair feeds both of
  its rows, so particles are conserved);
- rho = 0 and rho = 1 give ddm = 0;
- picture covariance: the interaction-picture ddm at t equals the
  Schroedinger-picture ddm of the rotated state, rotated back."""
import numpy as np
import pytest

from helpers import FS, rel_err

pytestmark = pytest.mark.gpu


def _mods():
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    return scatt, oracle


def _herm(nk, nb, rng, scale, diag):
    A = rng.normal(size=(nk, nb, nb)) + 1j * rng.normal(size=(nk, nb, nb))
    return (scale * 0.5 * (A + A.conj().transpose(0, 2, 1)) + diag * np.eye(nb)).reshape(nk, nb * nb)


def test_reference_size_case_against_the_oracle():
    scatt, oracle = _mods()
    print("case full size nb=2 nk=4096", flush=True)
    p = scatt.synth_params(1)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    P1, P2 = scatt.synth_P(p, e, k1, k2)
    rng = np.random.default_rng(1)
    nt = max(1, min(16, oracle.max_threads()))
    prefac = 2 * np.pi / p.nk
    with scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk) as op:
        op.generate_P(p)
        D1, D2 = op.download_P()
        assert np.abs(D1 - P1).max() <= 1e-14 * np.abs(P1).max() and np.abs(D2 - P2).max() <= 1e-14 * np.abs(P2).max()
        for t_fs in (0.0, 137.0):
            rho = _herm(p.nk, p.nb, rng, 0.05, 0.3)
            got = op.eval(t_fs * FS, rho)
            ref = oracle.evolve(k1, k2, P1, P2, e, rho, t=t_fs * FS, prefac=prefac, nthreads=nt)
            assert rel_err(got, ref) < 1e-12, (t_fs, rel_err(got, ref))
        rho0 = _herm(p.nk, p.nb, rng, 0.02, 0.3)
        got, _ = op.integrate_rk4(rho0, 0.0, FS, 5)
        ref, _ = oracle.rk4(k1, k2, P1, P2, e, rho0, FS, 5, prefac=prefac, nthreads=nt)
        d = ref - rho0
        assert np.abs(got - ref).max() < 1e-11 * np.abs(d).max()


@pytest.mark.parametrize("preset", [2, 4])
def test_workload_sizes_keep_the_operator_identities(preset):
    scatt, oracle = _mods()
    p = scatt.synth_params(preset)
    print(f"case full size preset {preset}: nb={p.nb} nk={p.nk}", flush=True)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    rng = np.random.default_rng(preset)
    nk, nb = p.nk, p.nb
    t = 211 * FS
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, nk_full=nk) as op:
        op.generate_P(p)
        r = _herm(nk, nb, rng, 0.05, 0.3)
        x = _herm(nk, nb, rng, 0.1, 0.05)
        F0 = op.eval(t, r)
        q1 = op.eval(t, r + x) + op.eval(t, r - x) - 2 * F0
        q2 = op.eval(t, r + 2 * x) + op.eval(t, r - 2 * x) - 2 * F0
        assert rel_err(q2, 4 * q1) < 1e-10, rel_err(q2, 4 * q1)
        d = F0.reshape(nk, nb, nb)
        assert np.abs(d - d.conj().transpose(0, 2, 1)).max() <= 1e-14 * np.abs(d).max()
        tr = np.trace(d, axis1=1, axis2=2)
        assert abs(tr.sum()) < 1e-11 * np.abs(tr).sum(), (abs(tr.sum()), np.abs(tr).sum())
        assert np.abs(op.eval(t, np.zeros_like(r))).max() == 0.0
        one = np.tile(np.eye(nb).reshape(1, -1), (nk, 1)).astype(np.complex128)
        assert np.abs(op.eval(t, one)).max() < 1e-13 * np.abs(F0).max()
        # picture covariance: rho_S = rho_I e^{-i(e_i - e_j)t}, ddm_I = ddm_S e^{+i(e_i - e_j)t}
        ph = np.exp(1j * (e[:, :, None] - e[:, None, :]) * t).reshape(nk, nb * nb)
        op.set_flags(expt=False)
        S = op.eval(t, r * ph.conj())
        op.set_flags(expt=True)
        assert rel_err(S * ph, F0) < 1e-12, rel_err(S * ph, F0)
