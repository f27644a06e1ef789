"""The adjoint through the C-ABI: the VJP kernel of the scattering term
against central differences of the CPU oracle (exact for a quadratic
right-hand side up to rounding), the checkpointed reverse RK4 against central
differences of the forward run, and torch's gradcheck of the differentiable
spin lifetime at nk = 256."""
import numpy as np
import pytest

from helpers import FS, random_P, random_energies, random_pairs, random_rho, rel_err

pytestmark = pytest.mark.gpu


def _mods():
    from denmat_dynm_development_amd import scatt
    from oracle import oracle
    return scatt, oracle


def _pair(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag))


@pytest.mark.parametrize("nb", [1, 2])
@pytest.mark.parametrize("t_fs", [0.0, 90.0])
def test_vjp_matches_oracle_central_differences(nb, t_fs):
    scatt, oracle = _mods()
    print(f"case vjp nb={nb} t={t_fs}", flush=True)
    rng = np.random.default_rng(3 + nb)
    nk = 45
    k1, k2 = random_pairs(nk, 6, rng, diag=True)
    e = random_energies(nk, nb, rng)
    P1, P2 = random_P(len(k1), nb, rng)
    rho = random_rho(nk, nb, rng)
    v = rng.normal(size=(nk, nb * nb)) + 1j * rng.normal(size=(nk, nb * nb))
    t = t_fs * FS
    prefac = 2 * np.pi / nk
    with scatt.ScatteringOperator(nk, nb, k1, k2, e) as op:
        op.upload_P(P1, P2)
        g = op.vjp(t, rho, v)
    gm = g.reshape(nk, nb, nb)
    assert np.abs(gm - gm.conj().transpose(0, 2, 1)).max() == 0.0  # Hermitian by construction
    h = 1e-3
    for trial in range(3):
        d = random_rho(nk, nb, rng, scale=1.0, diag=0.0) + np.diag(rng.normal(size=nb)).reshape(-1)
        fp = oracle.evolve(k1, k2, P1, P2, e, rho + h * d, t=t, prefac=prefac)
        fm = oracle.evolve(k1, k2, P1, P2, e, rho - h * d, t=t, prefac=prefac)
        fd = (_pair(v, fp) - _pair(v, fm)) / (2 * h)
        an = _pair(g, d)
        assert abs(an - fd) <= 1e-9 * max(abs(fd), 1e-30) + 1e-12 * np.abs(g).max(), (trial, an, fd)


def test_adjoint_rk4_matches_forward_central_differences():
    scatt, oracle = _mods()
    print("case adjoint rk4", flush=True)
    rng = np.random.default_rng(17)
    nk, nb, ncomp = 24, 2, 3
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    e = random_energies(nk, nb, rng)
    comps = []
    for _ in range(ncomp):
        P1, P2 = random_P(len(k1), nb, rng)
        comps.append((0.15 * P1, 0.15 * P2))
    theta = np.array([0.7, 1.1, 0.9])
    f = np.sort(rng.uniform(0.1, 0.5, (nk, nb)), axis=1)[:, ::-1].copy()
    rho0 = np.zeros((nk, nb, nb), np.complex128)
    rho0[:, np.arange(nb), np.arange(nb)] = f
    rho0 = rho0.reshape(nk, nb * nb) + random_rho(nk, nb, rng, scale=0.02, diag=0.0)
    s = random_rho(nk, nb, rng, scale=0.5, diag=0.0)
    dt, n, rep, cpe = 2 * FS, 40, 5, 8
    w = rng.normal(size=n // rep + 1)  # L = sum_m w_m S_m
    from denmat_dynm_development_amd.adjoint import SpinLifetimeProblem
    pr = SpinLifetimeProblem(nk, nb, k1, k2, e, comps, rho0, s, f, 0, nb, 1.0, 0.0, dt, n, rep,
                             checkpoint_every=cpe)
    try:
        L = lambda th: float(np.dot(w, pr.observe(th)))
        dth, drho = pr.grad_obs_functional(theta, w)
        h = 1e-6
        for c in range(ncomp):
            tp, tm = theta.copy(), theta.copy()
            tp[c] += h
            tm[c] -= h
            fd = (L(tp) - L(tm)) / (2 * h)
            print(f"theta[{c}]: adjoint {dth[c]:.10e} fd {fd:.10e}")
            assert abs(dth[c] - fd) <= 1e-4 * abs(fd), (c, dth[c], fd)
        # initial state along a Hermitian direction
        d = random_rho(nk, nb, rng, scale=1.0, diag=0.0)
        pr.rho0 = rho0 + h * d
        Lp = L(theta)
        pr.rho0 = rho0 - h * d
        Lm = L(theta)
        pr.rho0 = rho0
        fd = (Lp - Lm) / (2 * h)
        print(f"rho0: adjoint {_pair(drho, d):.10e} fd {fd:.10e}")
        assert abs(_pair(drho, d) - fd) <= 1e-4 * abs(fd)
    finally:
        pr.close()


@pytest.mark.parametrize("nb", [1, 2])
def test_adjoint_segment_lengths_agree(nb):
    """One segment (every stage state kept), segments of 8 and a ragged last
    segment (7 into 40) give the same gradients: the recomputed stage states
    are the forward sweep's."""
    scatt, oracle = _mods()
    print(f"case adjoint segments nb={nb}", flush=True)
    rng = np.random.default_rng(29 + nb)
    nk, ncomp = 30, 2
    k1, k2 = random_pairs(nk, 5, rng, diag=True)
    e = random_energies(nk, nb, rng)
    comps = []
    for _ in range(ncomp):
        P1, P2 = random_P(len(k1), nb, rng)
        comps.append((0.15 * P1, 0.15 * P2))
    theta = np.array([0.8, 1.2])
    f = np.sort(rng.uniform(0.1, 0.5, (nk, nb)), axis=1)[:, ::-1].copy()
    rho0 = np.zeros((nk, nb, nb), np.complex128)
    rho0[:, np.arange(nb), np.arange(nb)] = f
    rho0 = rho0.reshape(nk, nb * nb) + random_rho(nk, nb, rng, scale=0.02, diag=0.0)
    s = random_rho(nk, nb, rng, scale=0.5, diag=0.0)
    dt, n, rep = 2 * FS, 40, 5
    w = rng.normal(size=n // rep + 1)
    from denmat_dynm_development_amd.adjoint import SpinLifetimeProblem
    got = []
    for cpe in (1000, 8, 7):
        pr = SpinLifetimeProblem(nk, nb, k1, k2, e, comps, rho0, s, f, 0, nb, 1.0, 0.0, dt, n, rep,
                                 checkpoint_every=cpe)
        try:
            got.append(pr.grad_obs_functional(theta, w))
        finally:
            pr.close()
    for dth, drho in got[1:]:
        assert np.abs(dth - got[0][0]).max() <= 1e-12 * np.abs(got[0][0]).max()
        assert np.abs(drho - got[0][1]).max() <= 1e-12 * np.abs(got[0][1]).max()


def test_tau_gradcheck_nk256():
    """torch.autograd.gradcheck of tau_s(theta) over the six branch scales and
    the impurity density of a synthetic nk = 256 problem (FP64 forward)."""
    import torch
    scatt, oracle = _mods()
    from denmat_dynm_development_amd.adjoint import SpinLifetimeProblem, TauFunction
    print("case tau gradcheck", flush=True)
    p = scatt.synth_params(5)
    p.nk, p.partners = 256, 24
    e = scatt.synth_energies(p)
    k1, k2 = scatt.synth_kpairs(p)
    masks = [1 << a for a in range(6)] + [1 << 31]
    params = []
    for m in masks:
        q = scatt.synth_params(5)
        q.nk, q.partners, q.branch_mask = 256, 24, m
        params.append(q)
    T = p.temperature
    mu = -0.1 / 27.21138505
    f = 1.0 / (np.exp((e - mu) / T) + 1.0)
    s3 = scatt.synth_spin(p)
    sz = s3[:, 2].reshape(p.nk, p.nb, p.nb)
    rng = np.random.default_rng(0)
    rho0 = np.zeros((p.nk, p.nb, p.nb), np.complex128)
    rho0[:, np.arange(p.nb), np.arange(p.nb)] = f
    rho0 = (rho0 + 1e-3 * f.mean() * (sz + sz.conj().transpose(0, 2, 1)) / 2).reshape(p.nk, -1)
    pr = SpinLifetimeProblem.from_components_generator(
        params, p.nk, p.nb, k1, k2, e, rho0=rho0, s_obs=s3[:, 2], f=f, i0=0, i1=p.nb,
        obs_prefac=1.0 / p.nk, t0=0.0, dt=1.0 * FS, n_steps=120, report_every=4, checkpoint_every=32)
    try:
        theta = torch.ones(7, dtype=torch.float64, requires_grad=True)
        tau = TauFunction.apply(theta, pr)
        print(f"tau_s = {tau.item():.6f} ps")
        assert np.isfinite(tau.item()) and tau.item() > 0
        assert torch.autograd.gradcheck(lambda th: TauFunction.apply(th, pr), (theta,), eps=1e-6, atol=1e-9,
                                        rtol=1e-4, nondet_tol=1e-9)
    finally:
        pr.close()
