"""Every benchmark workload at its full size through the C-ABI.

- nb = 2, nk = 4096 (964,065 stored pairs): S_z(t) over 200 RK4 steps of
  1 fs and the fitted tau_s against the oracle's own RK4 of the same run
  (spin_relaxation over many tau_s of this workload).
- nb = 4, nk = 20000 (9.75 M pairs, 6 branches + the e-i term): the FP64 panel
  kernel keeps the operator identities (exactly quadratic in rho, Hermitian,
  trace-free, zero at rho = 0 and 1, picture covariant); complex64 and BF16
  evaluations equal the FP64 kernel on the very P they store (copied into an
  FP64 handle on the device) to FP32 accuracy of the drho-driven part; the
  e-i term is linear: P_eph + P_ei evaluates to the sum of the two.
- nb = 4, nk = 32768 (18.1 M pairs, power-law rows), complex64 and BF16
  stores: the batched matrix-core pass with B = 16 trajectories equals the
  unbatched evaluation of each trajectory on the same stored P.
- nb = 2, nk = 16384 (7.8 M pairs, 7 P components): the VJP against central
  differences of the FP64 forward along random Hermitian directions, and the
  checkpointed adjoint RK4 over 64 steps against central differences
  (h = 1e-6) of the forward run for two branch scales and the impurity
  density.
"""
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


def _pair(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag))


def test_spin_signal_and_tau_against_the_oracle_rk4():
    """nb = 2, nk = 4096: 200 RK4 steps of 1 fs from the first-order spin
    state; S_z at every step and the fitted tau_s against the oracle's RK4
    (per-element phases, the reference's loop order) to 1e-10."""
    scatt, oracle = _mods()
    from denmat_dynm_development_amd import taufit
    p = scatt.synth_params(1)
    print(f"case spin signal nb={p.nb} nk={p.nk}", flush=True)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    rho0, f = scatt.synth_initial_state(p, e, s3)
    n, dt = 200, 1.0 * FS
    nt = max(1, min(16, oracle.max_threads()))
    P1, P2 = scatt.synth_P(p, e, k1, k2, nt)
    with scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk) as op:
        op.generate_P(p)
        op.set_observable(s3[:, 2], f, 0, p.nb, 1.0 / p.nk)
        rho_g, obs_g = op.integrate_rk4(rho0, 0.0, dt, n, report_every=1)
    print(f"  oracle rk4 on {nt} threads", flush=True)
    rho_o, obs_o = oracle.rk4(k1, k2, P1, P2, e, rho0, dt, n, prefac=2 * np.pi / p.nk, report_every=1,
                              s_obs=s3[:, 2], f_dm=f, i0=0, i1=p.nb, obs_prefac=1.0 / p.nk, nthreads=nt)
    times = np.arange(n + 1) * dt
    tau_g, tau_o = taufit.fit_tau(times, obs_g), taufit.fit_tau(times, obs_o)
    dS = np.abs(obs_g - obs_o).max() / np.abs(obs_o).max()
    print(f"  S_z {obs_o[0]:.6e} -> {obs_o[-1]:.6e}; max |dS|/max|S| {dS:.2e}; tau_s {tau_o:.6e} ps "
          f"(run spans {n * 1e-3 / tau_o:.1f} tau_s), |dtau|/tau {abs(tau_g / tau_o - 1):.2e}", flush=True)
    assert obs_o[-1] < 0.5 * obs_o[0]  # the run covers the decay
    assert dS <= 1e-10
    assert abs(tau_g / tau_o - 1) <= 1e-10
    assert rel_err(rho_g - rho0, rho_o - rho0) <= 1e-10


def _identities(op, nk, nb, e, rng, t):
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
    ph = np.exp(1j * (e[:, :, None] - e[:, None, :]) * t).reshape(nk, nb * nb)
    op.set_flags(expt=False)
    S = op.eval(t, r * ph.conj())
    op.set_flags(expt=True)
    assert rel_err(S * ph, F0) < 1e-12, rel_err(S * ph, F0)


@pytest.mark.parametrize("prec", [1, 2])
def test_nb4_ei_workload_reduced_forms_on_their_stored_P(prec):
    """nb = 4, nk = 20000 with the e-i term: the reduced store against the
    FP64 kernel on the same stored values, and the FP64 identities."""
    scatt, oracle = _mods()
    p = scatt.synth_params(3)
    print(f"case nb=4 e-i workload prec={prec}: nk={p.nk}", flush=True)
    assert p.imp_fraction > 0
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    states, f = scatt.synth_batch_states(p, e, s3, 3)
    rng = np.random.default_rng(30 + prec)
    t = 37 * FS
    nk, nb = p.nk, p.nb
    rho_f = np.zeros((nk, nb * nb), np.complex128)
    rho_f[:, ::nb + 1] = f
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, nk_full=nk, precision=prec) as red, \
            scatt.ScatteringOperator(nk, nb, k1, k2, e, nk_full=nk) as ref:
        red.generate_P(p)
        red.set_split(f)
        ref.copy_P_from(red)
        base_r, base_64 = red.eval(t, rho_f), ref.eval(t, rho_f)
        # the equilibrium part is FP64 arithmetic on the stored P in both
        assert rel_err(base_r, base_64) < 1e-12, rel_err(base_r, base_64)
        for b in range(3):
            got, want = red.eval(t, states[b]), ref.eval(t, states[b])
            e_d = rel_err(got - base_r, want - base_64)
            print(f"  state {b}: drho part {e_d:.2e}, whole {rel_err(got, want):.2e}", flush=True)
            assert e_d < 2e-6 and rel_err(got, want) < 1e-7
        if prec == 1:
            _identities(ref, nk, nb, e, rng, 211 * FS)


def test_nb4_ei_term_is_linear_at_full_size():
    """P = P_eph + P_ei (the generator's branch masks: 0x3f the six phonon
    branches, bit 31 the e-i term): eval of the whole P equals the sum of the
    two evaluations, and the e-i part is a visible share of it."""
    scatt, oracle = _mods()
    p = scatt.synth_params(3)
    print(f"case nb=4 e-i linearity nk={p.nk}", flush=True)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    rng = np.random.default_rng(33)
    rho = _herm(p.nk, p.nb, rng, 0.05, 0.3)
    t = 91 * FS
    out = {}
    with scatt.ScatteringOperator(p.nk, p.nb, k1, k2, e, nk_full=p.nk) as op:
        for name, mask in (("all", 0), ("eph", 0x3F), ("ei", 1 << 31)):
            q = scatt.synth_params(3)
            q.branch_mask = mask
            op.generate_P(q)
            out[name] = op.eval(t, rho)
    share = np.linalg.norm(out["ei"]) / np.linalg.norm(out["all"])
    err = rel_err(out["eph"] + out["ei"], out["all"])
    err_ei = rel_err(out["all"] - out["eph"], out["ei"])
    print(f"  e-i share {share:.3e}, |eph + ei - all| / |all| {err:.2e}, "
          f"|(all - eph) - ei| / |ei| {err_ei:.2e}", flush=True)
    # the e-i part of this workload is ~1e-6 of the whole (n_i V_cell = 4.5e-5,
    # sigma_imp = sigma_g / 2); FP64 resolves it to ~1e-10 of itself
    assert 1e-8 < share < 1.0
    assert err < 1e-12 and err_ei < 1e-6


@pytest.mark.parametrize("prec", [1, 2])
def test_batched_workload_full_size_matches_unbatched(prec):
    """nb = 4, nk = 32768 power-law rows: B = 16 trajectories in one
    matrix-core pass against the unbatched evaluation of three of them on the
    same stored P.  Complex64 store (BF16 hi + lo records, 16-bit operands,
    FP32 sums): the drho-driven part to 5e-6; BF16 store (FP16 records, one
    11-bit word per entry): to 1e-4."""
    scatt, oracle = _mods()
    p = scatt.synth_params(4)
    print(f"case batched full size nk={p.nk} prec={prec}", flush=True)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    B = 16
    states, f = scatt.synth_batch_states(p, e, s3, B)
    nk, nb = p.nk, p.nb
    rho_f = np.zeros((nk, nb * nb), np.complex128)
    rho_f[:, ::nb + 1] = f
    t = 53 * FS
    tol_d, tol = (5e-6, 1e-7) if prec == 1 else (1e-4, 1e-6)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, nk_full=nk, precision=prec) as op:
        op.generate_P(p)
        op.set_split(f)
        got = op.eval_batched(t, states)
        base = op.eval(t, rho_f)
        for b in (0, 7, 15):
            one = op.eval(t, states[b])
            e_d = rel_err(got[b] - base, one - base)
            print(f"  trajectory {b}: drho part {e_d:.2e}, whole {rel_err(got[b], one):.2e}", flush=True)
            assert e_d < tol_d and rel_err(got[b], one) < tol


def test_adjoint_workload_full_size():
    """nb = 2, nk = 16384, 7.8 M pairs: VJP of the FP64 operator against
    central differences of its own forward (quadratic, so exact up to
    rounding) along 3 Hermitian directions; the adjoint RK4 over 64 steps
    against central differences (h = 1e-6) of the forward run for branch
    scales 0 and 3 and the impurity density."""
    scatt, oracle = _mods()
    from denmat_dynm_development_amd.adjoint import SpinLifetimeProblem
    p = scatt.synth_params(5)
    print(f"case adjoint full size nk={p.nk}", flush=True)
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    s3 = scatt.synth_spin(p)
    rho0, f = scatt.synth_initial_state(p, e, s3)
    nk, nb = p.nk, p.nb
    rng = np.random.default_rng(55)
    t = 40 * FS
    with scatt.ScatteringOperator(nk, nb, k1, k2, e, nk_full=nk) as op:
        op.generate_P(p)
        v = rng.normal(size=(nk, nb * nb)) + 1j * rng.normal(size=(nk, nb * nb))
        g = op.vjp(t, rho0, v)
        h = 1e-3
        for trial in range(3):
            d = _herm(nk, nb, rng, 1.0, 0.0) * 0.05
            fd = (_pair(v, op.eval(t, rho0 + h * d)) - _pair(v, op.eval(t, rho0 - h * d))) / (2 * h)
            an = _pair(g, d)
            print(f"  vjp direction {trial}: {an:.12e} fd {fd:.12e}", flush=True)
            assert abs(an - fd) <= 1e-8 * abs(fd), (trial, an, fd)
    masks = [1 << a for a in range(p.n_modes)] + [1 << 31]
    params = []
    for m in masks:
        q = scatt.synth_params(5)
        q.branch_mask = m
        params.append(q)
    n, rep = 64, 1
    w = rng.normal(size=n // rep + 1)
    pr = SpinLifetimeProblem.from_device_generator(
        params, nk, nb, k1, k2, e, rho0=rho0, s_obs=s3[:, 2], f=f, i0=0, i1=nb, obs_prefac=1.0 / nk,
        t0=0.0, dt=1.0 * FS, n_steps=n, report_every=rep, checkpoint_every=16, nk_full=nk)
    try:
        theta = np.ones(len(masks))
        L = lambda th: float(np.dot(w, pr.observe(th)))
        dth, _ = pr.grad_obs_functional(theta, w)
        hh = 1e-6
        for c in (0, 3, 6):
            tp, tm = theta.copy(), theta.copy()
            tp[c] += hh
            tm[c] -= hh
            fd = (L(tp) - L(tm)) / (2 * hh)
            print(f"  theta[{c}]: adjoint {dth[c]:.10e} fd {fd:.10e}", flush=True)
            assert abs(dth[c] - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(dth).max()), (c, dth[c], fd)
    finally:
        pr.close()
