"""Differentiable spin lifetime.

tau_s of an RK4 trajectory of the on-device scattering operator, as a
function of the weights theta_c of its P components (P = sum_c theta_c P_c:
per-branch coupling scales and the impurity density, each a linear factor on
its part of P, as scale_eph is on the whole, src_v4.5.7/ElectronPhonon.cpp:
225,232 and :87-97).  The gradient comes from the hand-written adjoint
(dmd_scatt_adjoint_rk4: checkpointed reverse RK4 with the VJP kernel of the
scattering term) and the implicit derivative of the exponential fit
(taufit.dtau_ds).  TauFunction wraps both as a torch.autograd.Function.
"""
import numpy as np
import torch

from . import _lib
from . import scatt
from . import taufit


class SpinLifetimeProblem:
    """One trajectory setup: pairs, energies, the P components, initial state,
    observable and integration grid.  All handles are FP64."""

    def __init__(self, nk, nb, k1, k2, e, components, rho0, s_obs, f, i0, i1, obs_prefac, t0, dt, n_steps,
                 report_every, checkpoint_every=64, nk_full=None, expt=True):
        self.nk, self.nb = int(nk), int(nb)
        self.rho0 = np.ascontiguousarray(rho0, np.complex128).reshape(self.nk, self.nb * self.nb)
        self.t0, self.dt, self.n_steps = float(t0), float(dt), int(n_steps)
        self.report_every, self.checkpoint_every = int(report_every), int(checkpoint_every)
        self.comps = []
        for P1, P2 in components:
            c = scatt.ScatteringOperator(nk, nb, k1, k2, e, nk_full=nk_full)
            c.upload_P(P1, P2)
            c.set_flags(expt=expt, kernel=_lib.DMD_KERNEL_AUTO)
            self.comps.append(c)
        self.op = scatt.ScatteringOperator(nk, nb, k1, k2, e, nk_full=nk_full)
        self.op.set_flags(expt=expt, kernel=_lib.DMD_KERNEL_AUTO)
        self.op.set_observable(s_obs, f, i0, i1, obs_prefac)

    @classmethod
    def from_device_generator(cls, params_by_component, nk, nb, k1, k2, e, **kw):
        """Components written by the on-device generator (no host copy of P)."""
        self = cls(nk, nb, k1, k2, e, [], **kw)
        for q in params_by_component:
            c = scatt.ScatteringOperator(nk, nb, k1, k2, e, nk_full=kw.get("nk_full"))
            c.generate_P(q)
            c.set_flags(expt=kw.get("expt", True), kernel=_lib.DMD_KERNEL_AUTO)
            self.comps.append(c)
        return self

    @classmethod
    def from_components_generator(cls, params_by_component, nk, nb, k1, k2, e, **kw):
        """Components from the synthetic generator, one dmd_synth_params each
        (e.g. branch_mask = 1 << a for branch a, 1 << 31 for the e-i term)."""
        comps = [scatt.synth_P(p, e, k1, k2) for p in params_by_component]
        return cls(nk, nb, k1, k2, e, comps, **kw)

    def close(self):
        for c in self.comps:
            c.close()
        self.op.close()

    def times_au(self):
        nobs = self.n_steps // self.report_every + 1
        return self.t0 + np.arange(nobs) * self.report_every * self.dt

    def observe(self, theta):
        """S at the reports for weights theta."""
        self.op.combine_P(self.comps, np.asarray(theta, np.float64))
        _, obs = self.op.integrate_rk4(self.rho0, self.t0, self.dt, self.n_steps, report_every=self.report_every)
        return obs

    def grad_obs_functional(self, theta, dL_dobs):
        """dL/dtheta (and dL/drho0) of a loss with gradient dL_dobs on the reports."""
        self.op.combine_P(self.comps, np.asarray(theta, np.float64))
        return self.op.adjoint_rk4(self.rho0, self.t0, self.dt, self.n_steps, self.report_every,
                                   self.checkpoint_every, dL_dobs, self.comps)

    def tau(self, theta):
        return taufit.fit_tau(self.times_au(), self.observe(theta))

    def tau_and_grad(self, theta):
        tau, dtds = taufit.dtau_ds(self.times_au(), self.observe(theta))
        dth, _ = self.grad_obs_functional(theta, dtds)
        return tau, dth


class TauFunction(torch.autograd.Function):
    """tau_s(theta) in ps; backward through the adjoint and the fit."""

    @staticmethod
    def forward(ctx, theta, problem):
        th = theta.detach().cpu().numpy().astype(np.float64)
        tau, dtds = taufit.dtau_ds(problem.times_au(), problem.observe(th))
        ctx.problem, ctx.th, ctx.dtds = problem, th, dtds
        return torch.tensor(tau, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, g):
        dth, _ = ctx.problem.grad_obs_functional(ctx.th, ctx.dtds)
        return g * torch.as_tensor(dth, dtype=g.dtype, device=g.device), None


def tau_loss(theta, problem, tau_target):
    """(tau_s(theta) - tau_target)^2, differentiable in theta."""
    return (TauFunction.apply(theta, problem) - tau_target) ** 2
