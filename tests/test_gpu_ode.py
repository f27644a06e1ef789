"""The adaptive RKF45 driver and the Euler stepper of the reference's
dm_dynamics (evolve_gsl with rkf45, evolve_euler) on the device: the
Fehlberg tableau pinned by the closed form of one step on -i[e, rho] (the
fourth-order stability polynomial 1 + z + z^2/2 + z^3/6 + z^4/24 + z^5/104
and the embedded error (R5 - R4)(z), R5 = ... + z^5/120 + z^6/2080), Euler's
1 + z, the adaptive solution of a Lindblad system against a converged oracle
RK4, and the step control's response to the tolerance."""
import numpy as np
import pytest

from helpers import FS, random_P, random_energies, random_pairs, random_rho, rel_err

pytestmark = pytest.mark.gpu


def _mods():
    from denmat_dynm_development_amd import _lib, scatt
    from oracle import oracle
    return _lib, scatt, oracle


def _coherent_only(scatt, nk, nb, rng):
    e = random_energies(nk, nb, rng, emax=0.02)
    op = scatt.ScatteringOperator(nk, nb, np.zeros(0, np.uint64), np.zeros(0, np.uint64), e)
    op.upload_P(np.zeros((0, nb ** 4), np.complex128), np.zeros((0, nb ** 4), np.complex128))
    op.set_flags(expt=False, schrodinger=True)  # the right-hand side is -i[e, rho] alone
    return op, e


def test_rkf45_one_step_is_the_fehlberg_polynomial():
    _lib, scatt, oracle = _mods()
    print("case rkf45 closed form", flush=True)
    rng = np.random.default_rng(81)
    nk, nb = 6, 3
    op, e = _coherent_only(scatt, nk, nb, rng)
    with op:
        rho = random_rho(nk, nb, rng, scale=0.3, diag=0.4)
        h = 7.0 * FS
        y, t, hn, st = op.rkf45_apply(rho, 0.0, h, h, 0.0, h, 1.0, 0.0)  # loose tolerance: one accepted step
        assert st["steps"] == 1 and st["rejected"] == 0 and st["rhs_evals"] == 6 and t == h
        z = (-1j * (e[:, :, None] - e[:, None, :]) * h).reshape(nk, nb * nb)
        R4 = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24 + z ** 5 / 104
        R5 = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24 + z ** 5 / 120 + z ** 6 / 2080
        assert np.abs(y - R4 * rho).max() < 1e-14
        err = (R5 - R4) * rho
        ratio = max(np.abs(err.real).max(), np.abs(err.imag).max()) / 1.0
        assert st["last_err_ratio"] == pytest.approx(ratio, rel=1e-6)
        # a planted mistake: the fifth-order weights would give R5, which this is not
        assert np.abs(y - R5 * rho).max() > 1e3 * np.abs(y - R4 * rho).max()
        # Euler: 1 + z per step
        ye = op.integrate_euler(rho, 0.0, h, 1)
        assert np.abs(ye - (1 + z) * rho).max() < 1e-15


def test_rkf45_adaptive_follows_converged_oracle_and_tolerance():
    _lib, scatt, oracle = _mods()
    print("case rkf45 adaptive", flush=True)
    rng = np.random.default_rng(82)
    nk, nb = 24, 2
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    P1, P2 = 0.2 * P1, 0.2 * P2
    prefac = 2 * np.pi / nk
    f = np.sort(rng.uniform(0.1, 0.5, (nk, nb)), axis=1)[:, ::-1].copy()
    rho0 = np.zeros((nk, nb, nb), np.complex128)
    rho0[:, np.arange(nb), np.arange(nb)] = f
    rho0 = rho0.reshape(nk, -1) + random_rho(nk, nb, rng, scale=0.02, diag=0.0)
    T = 200 * FS
    ref, _ = oracle.rk4(k1, k2, P1, P2, e, rho0, T / 4000, 4000, prefac=prefac)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        errs, steps = [], []
        for eps in (1e-6, 1e-9, 1e-12):
            y, t, hn, st = op.rkf45_apply(rho0, 0.0, T, 1e-3 * FS, 0.0, 50 * FS, eps, 0.0)
            assert t == T and st["last_err_ratio"] <= 1.1
            errs.append(np.abs(y - ref).max())
            steps.append(st["steps"])
            print(f"eps {eps:g}: {st}, max |y - ref| {errs[-1]:.3e}")
        assert steps[0] < steps[1] < steps[2]
        assert errs[2] < errs[1] < errs[0] and errs[2] < 1e-10
        # the driver state carries over: continuing from T to 2T with the suggestion
        y2, t2, hn2, st2 = op.rkf45_apply(y, T, 2 * T, hn, 0.0, 50 * FS, 1e-12, 0.0)
        ref2, _ = oracle.rk4(k1, k2, P1, P2, e, ref, T / 4000, 4000, prefac=prefac, t0=T)
        assert t2 == 2 * T and np.abs(y2 - ref2).max() < 1e-10
        # Euler against a numpy Euler with the oracle's right-hand side
        dt, n = 2 * FS, 15
        ye = op.integrate_euler(rho0, 0.0, dt, n)
        yr = rho0.copy()
        for i in range(n):
            k = oracle.evolve(k1, k2, P1, P2, e, yr, t=i * dt, prefac=prefac).reshape(nk, nb, nb)
            yr = yr + 0.5 * dt * (k + k.conj().transpose(0, 2, 1)).reshape(nk, -1)
        assert rel_err(ye, yr) < 1e-12


def test_compute_rhs_is_the_whole_right_hand_side():
    """dmd_scatt_compute_rhs: the right-hand side the integrators step, at t,
    with the band window, the stored ddmdt_eq and the coherent term, against
    the oracle's evolve on the cut and its coherent term on the whole matrix."""
    _lib, scatt, oracle = _mods()
    print("case compute_rhs window", flush=True)
    rng = np.random.default_rng(83)
    nk, nb, nbd, b0 = 30, 2, 4, 1
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    e = random_energies(nk, nb, rng)
    e_dm = np.sort(rng.uniform(0, 4e-3, (nk, nbd)), axis=1)
    e_dm[:, b0:b0 + nb] = e
    P1, P2 = random_P(len(k1), nb, rng)
    prefac = 2 * np.pi / nk
    H = random_rho(nk, nbd, rng, scale=1e-4, diag=1e-4)
    eq = random_rho(nk, nb, rng, scale=1e-6, diag=1e-6)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        op.set_window(nbd, b0, e_dm)
        op.set_coherent(H)
        op.set_ddmdt_eq(eq)
        for t_fs in (0.0, 23.0):
            t = t_fs * FS
            rho = random_rho(nk, nbd, rng, scale=0.05, diag=0.3)
            got = op.compute_rhs(t, rho)
            cut = rho.reshape(nk, nbd, nbd)[:, b0:b0 + nb, b0:b0 + nb].reshape(nk, -1)
            inner = oracle.evolve(k1, k2, P1, P2, e, cut, t=t, prefac=prefac, ddmdt_eq=eq)
            ref = oracle.coherent(H, e_dm, rho, t).reshape(nk, nbd, nbd)
            ref[:, b0:b0 + nb, b0:b0 + nb] += inner.reshape(nk, nb, nb)
            assert rel_err(got, ref.reshape(nk, -1)) < 1e-12, (t_fs, rel_err(got, ref.reshape(nk, -1)))
            # one Euler step is rho + dt * (this right-hand side), Hermitian already
            dt = 0.5 * FS
            assert np.abs(op.integrate_euler(rho, t, dt, 1) - (rho + dt * got)).max() < 1e-15 + 1e-13 * dt * np.abs(got).max()


def test_rkf45_stops_when_the_step_underflows():
    """An error bound no step can meet (eps_abs = 1e-300) shrinks h by 0.2 per
    rejection until t + h == t: the driver returns an error there instead of
    looping (GSL's driver also gives up once no progress is possible)."""
    _lib, scatt, oracle = _mods()
    print("case rkf45 underflow", flush=True)
    rng = np.random.default_rng(84)
    nk, nb = 6, 2
    op, e = _coherent_only(scatt, nk, nb, rng)
    with op:
        rho = random_rho(nk, nb, rng, scale=0.3, diag=0.4)
        t0 = 100 * FS
        with pytest.raises(_lib.DMDError, match="underflowed"):
            op.rkf45_apply(rho, t0, t0 + FS, FS, 0.0, FS, 1e-300, 0.0)


def test_semiclassical_right_hand_side_and_steppers():
    """alg_semiclassical (DensityMatrix_Dynamics.h:249,277): the right-hand
    side is zero_offdiag(F(zero_offdiag(y))) with the coherent term gone; RK4
    and RKF45 keep the state's initial off-diagonals, Euler zeroes them as the
    reference's in-place compute() does; eval (evolve) is unaffected."""
    _lib, scatt, oracle = _mods()
    print("case semiclassical", flush=True)
    rng = np.random.default_rng(85)
    nk, nb = 20, 2
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    P1, P2 = 0.2 * P1, 0.2 * P2
    prefac = 2 * np.pi / nk
    H = random_rho(nk, nb, rng, scale=1e-4, diag=1e-4)
    rho0 = random_rho(nk, nb, rng, scale=0.02, diag=0.3)

    def zoff(y):
        z = y.reshape(nk, nb, nb).copy()
        z[:, ~np.eye(nb, dtype=bool)] = 0
        return z.reshape(nk, -1)

    def rhs(t, y):
        return zoff(oracle.evolve(k1, k2, P1, P2, e, zoff(y), t=t, prefac=prefac))

    dt, n = 2 * FS, 6
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        op.set_coherent(H)
        plain = op.eval(3 * FS, rho0)
        op.set_semiclassical(True)
        assert rel_err(op.eval(3 * FS, rho0), plain) < 1e-14  # evolve itself is unchanged
        got = op.compute_rhs(3 * FS, rho0)
        assert rel_err(got, rhs(3 * FS, rho0)) < 1e-12
        y4, _ = op.integrate_rk4(rho0, 0.0, dt, n)
        y = rho0.copy()
        for i in range(n):
            t = i * dt
            s1 = rhs(t, y)
            s2 = rhs(t + dt / 2, y + dt / 2 * s1)
            s3 = rhs(t + dt / 2, y + dt / 2 * s2)
            s4 = rhs(t + dt, y + dt * s3)
            y = y + dt / 6 * (s1 + 2 * s2 + 2 * s3 + s4)
        assert np.abs(y4 - y).max() < 1e-12 * np.abs(y).max()
        assert np.array_equal(zoff(y4) - y4, zoff(rho0) - rho0)  # off-diagonals untouched
        ye = op.integrate_euler(rho0, 0.0, dt, 1)
        ref = zoff(rho0) + dt * rhs(0.0, rho0)
        assert np.abs(ye - ref).max() < 1e-14 * np.abs(ref).max()
        with pytest.raises(_lib.DMDError):
            op.integrate_rk4_batched(np.stack([rho0, rho0]), 0.0, dt, 1)


def test_phenomenological_relaxation_term():
    """alg_phenom_relax (phenom_relax, scattering_model/phenomenon_relax.h:42-51):
    the integrators' right-hand side gains -(rho - diag f_eq)/tau on the
    density-matrix bands [b_start, b_end), here overlapping the e-ph window
    partly; compute_rhs and RK4 against the oracle plus that term in numpy."""
    _lib, scatt, oracle = _mods()
    print("case phenom relax", flush=True)
    rng = np.random.default_rng(86)
    nk, nb, nbd, b0 = 24, 2, 4, 1
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    e = random_energies(nk, nb, rng)
    e_dm = np.sort(rng.uniform(0, 4e-3, (nk, nbd)), axis=1)
    e_dm[:, b0:b0 + nb] = e
    P1, P2 = random_P(len(k1), nb, rng)
    prefac = 2 * np.pi / nk
    feq = rng.uniform(0.1, 0.9, (nk, nbd))
    tau, bs, be = 300 * FS, 0, 3
    rho0 = random_rho(nk, nbd, rng, scale=0.03, diag=0.3)

    def rhs(t, y):
        yy = y.reshape(nk, nbd, nbd)
        inner = oracle.evolve(k1, k2, P1, P2, e, yy[:, b0:b0 + nb, b0:b0 + nb].reshape(nk, -1), t=t, prefac=prefac)
        out = np.zeros_like(yy)
        out[:, b0:b0 + nb, b0:b0 + nb] = inner.reshape(nk, nb, nb)
        d = yy[:, bs:be, bs:be].copy()
        d[:, np.arange(be - bs), np.arange(be - bs)] -= feq[:, bs:be]
        out[:, bs:be, bs:be] -= d / tau
        return out.reshape(nk, -1)

    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        op.set_window(nbd, b0, e_dm)
        op.set_phenom_relax(tau, bs, be, feq)
        assert rel_err(op.compute_rhs(9 * FS, rho0), rhs(9 * FS, rho0)) < 1e-12
        dt, n = 2 * FS, 5
        got, _ = op.integrate_rk4(rho0, 0.0, dt, n)
        y = rho0.copy()
        for i in range(n):
            t = i * dt
            s1 = rhs(t, y)
            s2 = rhs(t + dt / 2, y + dt / 2 * s1)
            s3 = rhs(t + dt / 2, y + dt / 2 * s2)
            s4 = rhs(t + dt, y + dt * s3)
            y = y + dt / 6 * (s1 + 2 * s2 + 2 * s3 + s4)
            yy = y.reshape(nk, nbd, nbd)
            y = (0.5 * (yy + yy.conj().transpose(0, 2, 1))).reshape(nk, -1)
        assert np.abs(got - y).max() < 1e-11 * np.abs(y - rho0).max()
        with pytest.raises(_lib.DMDError):
            op.set_phenom_relax(tau, 2, 5, feq)  # beyond nb_dm
        op.set_phenom_relax(0.0, 0, 0, None)  # off
        inner_only = op.compute_rhs(9 * FS, rho0)
        assert rel_err(inner_only, rhs(9 * FS, rho0) + 0) > 1e-6


def test_rhs_hook_rebuilds_P_from_every_stage_state():
    """A right-hand-side hook (update_scatt_inside) that sets P from the stage
    state: P = (1 + 2 Re rho_0[0,0]) P0 before every stage.  RK4 (stages at
    t0 + (n + a) dt, uncaptured), Euler and compute_rhs against the oracle
    applying the same rule; rkf45 calls it once per stage evaluation; a hook
    that fails stops the integrator."""
    _lib, scatt, oracle = _mods()
    print("case rhs hook", flush=True)
    rng = np.random.default_rng(91)
    nk, nb = 40, 2
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    rho0 = random_rho(nk, nb, rng)
    prefac = 2 * np.pi / nk
    dt, n = 0.05 * FS, 4
    scale = lambda y: 1.0 + 2.0 * y.reshape(nk, nb * nb)[0, 0].real
    g = lambda t, y: oracle.evolve(k1, k2, scale(y) * P1, scale(y) * P2, e, y, t=t, prefac=prefac)
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        calls = []

        def hook(t, rho):
            calls.append(t)
            op.upload_P(scale(rho) * P1, scale(rho) * P2)

        op.set_rhs_hook(hook)
        got, _ = op.integrate_rk4(rho0, 0.5 * FS, dt, n)
        assert len(calls) == 4 * n
        assert np.allclose(calls, [0.5 * FS + (i + a) * dt for i in range(n) for a in (0, 0.5, 0.5, 1)],
                           rtol=0, atol=1e-9)
        y = rho0.copy()
        for i in range(n):
            t = 0.5 * FS + i * dt
            a1 = g(t, y)
            a2 = g(t + dt / 2, y + dt / 2 * a1)
            a3 = g(t + dt / 2, y + dt / 2 * a2)
            a4 = g(t + dt, y + dt * a3)
            y = y + dt / 6 * (a1 + 2 * a2 + 2 * a3 + a4)
            yy = y.reshape(nk, nb, nb)
            y = (0.5 * (yy + yy.conj().transpose(0, 2, 1))).reshape(nk, -1)
        assert rel_err(got, y) < 1e-12
        calls.clear()
        got = op.integrate_euler(rho0, 0.0, dt, n)
        y = rho0.copy()
        for i in range(n):
            k = g(i * dt, y).reshape(nk, nb, nb)
            y = y + 0.5 * dt * (k + k.conj().transpose(0, 2, 1)).reshape(nk, -1)
        assert len(calls) == n and rel_err(got, y) < 1e-12
        calls.clear()
        assert rel_err(op.compute_rhs(7 * FS, rho0), g(7 * FS, rho0)) < 1e-12 and len(calls) == 1
        calls.clear()
        _, _, _, st = op.rkf45_apply(rho0, 0.0, 2 * dt, dt, 0.0, dt, 1e-6, 0.0)
        assert len(calls) == st["rhs_evals"] > 0

        def bad(t, rho):
            raise RuntimeError("stop")

        op.set_rhs_hook(bad)
        with pytest.raises(_lib.DMDError):
            op.integrate_rk4(rho0, 0.0, dt, 2)
        op.set_rhs_hook(None)
        op.upload_P(P1, P2)
        got, _ = op.integrate_rk4(rho0, 0.0, dt, 2)  # no hook: the captured-graph path again
        assert np.all(np.isfinite(got))
