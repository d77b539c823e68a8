"""Frequency solve on the GPU: drop-in for nsrno/wave_solver.py.

``solve_frequency(model, medium, frequency_hz, angle_index, tolerance,
max_iterations, restart)`` keeps the reference's signature, guards
(ValueError for a medium that requires grad or a budget that is not a
positive multiple of restart), return value (prediction [192] complex64,
info dict with exactly the keys reason, initial_neural_iterations,
solver_iterations, relative_residual, seconds, trace) and semantics
(wave_solver.py:38-67).  The solve itself is libhelm.so's flexible GMRES with
FP64 true-residual refinement and the reference's own preconditioner,
ifft2(multiplier fft2(r)) without the gain (wave_solver.py:56-59; the
project's mixed-radix FFT kernels), so ``relative_residual`` is the true
FP64 ||f - L u|| / ||f||, and tolerances below the complex64 rounding floor
are reached.

``gmres`` keeps the reference's callable-based signature
(wave_solver.py:9-36) for any caller operator and preconditioner (this
package's are ``PlanOperator`` and ``PlanPreconditioner``);
``solve_frequency_batch`` solves several angles of one frequency as one
batched solve (shared coefficients, one launch sequence).
"""
from __future__ import annotations

import time

import torch

from .wave_operator import WaveResidualOperator


class PlanOperator:
    """u -> L u for one SpectralData (what the reference passes as ``operator``)."""

    def __init__(self, model: WaveResidualOperator, data):
        self.model, self.data = model, data

    def __call__(self, u):
        return self.model.apply_operator(u, self.data)


class PlanPreconditioner:
    """r -> k0^2 M^{-H} r: the plan's multigrid V-cycle as a preconditioner of
    the reference's operator L (the plan holds A = k0^2 conj(L), so
    L^{-1} = k0^2 (A^H)^{-1}); replaces the Fourier shifted-Laplacian
    ``precondition`` closure of wave_solver.py:58-61."""

    def __init__(self, model: WaveResidualOperator, data):
        self.model, self.data = model, data

    def __call__(self, r):
        v = r.to(self.model.device, torch.complex64).contiguous()
        z = self.data.plan.precondition(v, adjoint=True) * self.data.k0 ** 2
        return z.to(r.dtype)


def gmres(operator, precondition, rhs, x, callback, restart=32, max_cycles=80, tolerance=1e-4):
    """Right-preconditioned restarted GMRES (wave_solver.py:9-36) -> (x, reason).

    ``operator`` and ``precondition`` are any tensor -> tensor callables on
    tensors of ``rhs``'s shape and dtype (``precondition=None``: none), as in
    the reference; ``PlanOperator`` / ``PlanPreconditioner`` give this
    package's operator and multigrid.  The whole ``rhs`` is one Krylov vector
    (the reference's semantics); libhelm.so's reverse-communication kernels
    orthogonalise (DCGS2, FP64 sums), solve the least squares (FP64 Givens)
    and accumulate x in complex128.  ``callback(applications, x,
    relative_residual)`` fires at every restart with the current iterate and
    the true relative residual; reason is 'converged' or 'iteration_budget'.
    """
    from .krylov import fgmres
    pre = precondition if precondition is not None else (lambda r: r)
    return fgmres(operator, pre, rhs, x, callback, restart=restart, max_cycles=max_cycles, tolerance=tolerance)


def _check(medium, max_iterations, restart):
    if medium.requires_grad:
        raise ValueError("This reference has no implicit-adjoint backward; use the neural forward for input gradients.")
    if max_iterations < restart or max_iterations % restart:
        raise ValueError("max_iterations must be a positive multiple of restart")


@torch.no_grad()
def solve_frequency_batch(model: WaveResidualOperator, medium, frequency_hz, angle_indices, tolerance=1e-4,
                          max_iterations=2560, restart=32):
    """Solve several plane-wave angles of one frequency together.

    Returns (predictions [A, 192] complex64, fields [A, H, W] complex128,
    infos: one dict per angle with the keys of solve_frequency).
    """
    _check(medium, max_iterations, restart)
    started = time.perf_counter()
    model.eval()
    angles = list(angle_indices)
    data, _, _ = model.setup(medium, frequency_hz)
    inc = torch.stack([model.incident64(frequency_hz, a, data) for a in angles])
    rhs = model.scattered_rhs(inc, data)
    x = torch.zeros_like(rhs)
    data.plan.warm_start(rhs, x, model.iterations, precond="spectral", adjoint=True)
    traces = [[] for _ in angles]
    calls = [0]

    def monitor(applications, rel):
        calls[0] += 1
        for t, r in zip(traces, rel):
            t.append(dict(iterations=applications - calls[0], relative_residual=float(r)))

    # GMRES preconditioned by ifft2(multiplier fft2(r)) without the gain, as
    # wave_solver.py:56-59 (the learned boundary taper is left out there)
    x, info = data.plan.solve(rhs, x, tolerance=tolerance, restart=restart, max_iterations=max_iterations,
                              precond="spectral_plain", adjoint=True, callback=monitor)
    cols = model.rx_cols.to(model.device)
    preds = torch.stack([x[k, model.face, cols] * model.source_factor(frequency_hz, a).to(model.device)
                         for k, a in enumerate(angles)]).to(torch.complex64)
    seconds = time.perf_counter() - started
    infos = []
    for k in range(len(angles)):
        infos.append(dict(reason=info[k].reason, initial_neural_iterations=model.iterations,
                          solver_iterations=info[k].iterations, relative_residual=info[k].relative_residual,
                          seconds=seconds, trace=traces[k]))
    return preds, x, infos


@torch.no_grad()
def solve_frequency(model, medium, frequency_hz, angle_index, tolerance=1e-4, max_iterations=2560, restart=32):
    """Return complex receiver spectrum and measured convergence status (wave_solver.py:38-67)."""
    preds, _, infos = solve_frequency_batch(model, medium, frequency_hz, [angle_index], tolerance, max_iterations,
                                            restart)
    return preds[0], infos[0]
