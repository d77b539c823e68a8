"""Frequency solve on the GPU: drop-in for nsrno/wave_solver.py.

``solve_frequency(model, medium, frequency_hz, angle_index, tolerance,
max_iterations, restart)`` keeps the reference's signature, guards
(ValueError for a medium that requires grad or a budget that is not a
positive multiple of restart), return value (prediction [192] complex64,
info dict with exactly the keys reason, initial_neural_iterations,
solver_iterations, relative_residual, seconds, trace) and semantics
(wave_solver.py:38-67).  The solve itself is libhelm.so's flexible GMRES with
FP64 true-residual refinement and the multigrid preconditioner, so
``relative_residual`` is the true FP64 ||f - L u|| / ||f||, and tolerances
below the complex64 rounding floor are reached.

``gmres`` keeps the reference's callable-based signature
(wave_solver.py:9-36) for operators built by this package
(``PlanOperator``); ``solve_frequency_batch`` solves several angles of one
frequency as one batched solve (shared coefficients, one launch sequence).
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


def gmres(operator, precondition, rhs, x, callback, restart=32, max_cycles=80, tolerance=1e-4):
    """Right-preconditioned restarted GMRES (wave_solver.py:9-36) -> (x, reason).

    ``operator`` must be a PlanOperator; ``precondition`` is ignored in favour
    of the plan's multigrid (None or any callable is accepted).  ``callback``
    (applications, x, relative_residual) fires once per restart with the true
    FP64 relative residual; ``x`` passed to it is the current iterate.
    """
    if not isinstance(operator, PlanOperator):
        raise TypeError("gmres takes a PlanOperator built from WaveResidualOperator.setup")
    model, data = operator.model, operator.data
    k0sq = data.k0 ** 2
    b = (rhs.to(model.device, torch.complex128) * k0sq).contiguous()
    x0 = x.to(model.device, torch.complex128).contiguous()
    calls = [0]

    def cb(applications, rel):
        calls[0] += 1
        callback(applications - calls[0], None, rel[0])

    xs, info = data.plan.solve(b.reshape(1, *b.shape[-2:]), x0.reshape(1, *b.shape[-2:]), tolerance=tolerance,
                               restart=restart, max_iterations=restart * max_cycles, precond="mg", adjoint=True,
                               callback=cb)
    return xs.reshape(rhs.shape).to(rhs.dtype), info[0].reason


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
    data.plan.warm_start(rhs, x, model.iterations, precond="mg", adjoint=True)
    traces = [[] for _ in angles]
    calls = [0]

    def monitor(applications, rel):
        calls[0] += 1
        for t, r in zip(traces, rel):
            t.append(dict(iterations=applications - calls[0], relative_residual=float(r)))

    x, info = data.plan.solve(rhs, x, tolerance=tolerance, restart=restart, max_iterations=max_iterations,
                              precond="mg", adjoint=True, callback=monitor)
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
