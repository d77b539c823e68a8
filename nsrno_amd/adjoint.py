"""Receiver traces with an implicit-adjoint gradient (torch.autograd.Function).

Drop-in for the pattern of nsrno/dense_wave.py:227-259 (ImplicitChannels):
forward solves the frequency-domain problem for every plane-wave angle and
returns receiver traces; backward solves the adjoint (conjugate-transpose)
problem once per angle with the trace gradient injected at the receivers and
forms the material gradient from the discrete adjoint-state identity

    dL/dm = -Re( sum_q lambda^H (dA/dm_q) u ),   A^H lambda = P^T dL/dd*,

with the sensitivities sum_r lam_r^H (dA/d eta_rel) u_r and
sum_r lam_r^H (dA/d kappa_c) u_r computed on the GPU
(helm_material_sensitivity).  Parameters are log c and log rho (per node),
the attenuation alpha [Np/m] is held fixed.  Non-convergence raises
RuntimeError, as the reference does.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib
from .helmholtz import HelmholtzPlan, Ports


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Problem:
    """Grid, frequency, acquisition and solver settings of one implicit solve."""

    def __init__(self, nz, nx, h_mm, frequency_mhz, pml_mm, weights, amplitudes, *, alpha_np_m=None, c0=1500.0,
                 rho0=1000.0, tolerance=1e-8, restart=30, max_iterations=20000, mg=None, x0_mm=0.0, z0_mm=0.0):
        self.plan = HelmholtzPlan(nz, nx, h_mm, frequency_mhz, pml_mm, x0_mm=x0_mm, z0_mm=z0_mm, c0_m_s=c0,
                                  nrhs=len(amplitudes), restart=restart)
        self.ports = Ports(weights)
        self.amp = torch.as_tensor(np.asarray(amplitudes), dtype=torch.complex128, device="cuda")
        self.alpha = None if alpha_np_m is None else torch.as_tensor(alpha_np_m, dtype=torch.float64,
                                                                     device="cuda").reshape(nz, nx)
        self.rho0, self.c0, self.tol, self.restart, self.max_it = rho0, c0, tolerance, restart, max_iterations
        self.mg = dict(levels=5, shift=0.5, pre=2, post=2, coarse_sweeps=10) if mg is None else mg
        self.h2 = (h_mm * 1e-3) ** 2
        self.omega = 2 * np.pi * frequency_mhz * 1e6

    def set_medium(self, c, rho):
        self.plan.set_medium_physical(c, rho, self.alpha, rho0=self.rho0)
        self.plan.setup_multigrid(**self.mg)

    def solve(self, rhs, adjoint=False):
        x, info = self.plan.solve(rhs, tolerance=self.tol, restart=self.restart, max_iterations=self.max_it,
                                  precond="mg", adjoint=adjoint)
        if any(i.reason != "converged" for i in info):
            raise RuntimeError(("Adjoint" if adjoint else "Forward") + " failed to converge: " + str(info))
        return x, info


class ImplicitReceivers(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_c, log_rho, problem):
        c = torch.exp(log_c.detach()).to(torch.float64).contiguous()
        rho = torch.exp(log_rho.detach()).to(torch.float64).contiguous()
        problem.set_medium(c, rho)
        b = problem.plan.inject(problem.ports, problem.amp)
        u, info = problem.solve(b)
        ctx.problem = problem
        ctx.info = info
        ctx.save_for_backward(c, rho, u)
        return problem.plan.sample(problem.ports, u)

    @staticmethod
    def backward(ctx, grad):
        c, rho, u = ctx.saved_tensors
        pr = ctx.problem
        pr.set_medium(c, rho)
        # adjoint right-hand side P^T g: helm_inject scales by -1/h^2
        rhs = pr.plan.inject(pr.ports, (-pr.h2) * grad.to(torch.complex128).contiguous())
        lam, _ = pr.solve(rhs, adjoint=True)
        n = pr.plan.n
        s_eta = torch.empty(n, dtype=torch.complex128, device="cuda")
        s_kap = torch.empty_like(s_eta)
        check(lib().helm_material_sensitivity(pr.plan._h, ctypes.c_void_p(u.data_ptr()),
                                              ctypes.c_void_p(lam.data_ptr()), u.shape[0],
                                              ctypes.c_void_p(s_eta.data_ptr()), ctypes.c_void_p(s_kap.data_ptr()),
                                              _s()), "helm_material_sensitivity")
        s_eta = s_eta.reshape(c.shape)
        s_kap = s_kap.reshape(c.shape)
        eta = pr.rho0 / rho
        att = torch.ones_like(c, dtype=torch.complex128)
        if pr.alpha is not None:
            att = torch.complex(torch.ones_like(c), pr.alpha * c / pr.omega)
        kappa = pr.rho0 * pr.c0 ** 2 / (rho * c * c) * att * att
        # d eta/d log rho = -eta ; d kappa/d log rho = -kappa ; d kappa/d log c = -2 kappa/att
        g_logrho = (s_eta * eta + s_kap * kappa).real
        g_logc = (2.0 * s_kap * kappa / att).real
        return g_logc, g_logrho, None


def implicit_receivers(log_c, log_rho, problem):
    """Receiver traces [R, P] (complex128) with implicit-adjoint gradients."""
    return ImplicitReceivers.apply(log_c, log_rho, problem)
