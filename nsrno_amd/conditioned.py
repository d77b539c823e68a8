"""Conditioned split-PML frequency system on the GPU (the dense path).

Drop-in for nsrno/conditioned_wave.py:12-92 (FrequencyGeometry,
ConditionedFrequencySystem) and the operator functions of
nsrno/dense_wave.py:59-225 (source_rhs, channels, projected_operator,
adjoint_operator, solve_system).  The operator is applied by the HIP 8-tap
kernels of libhelm.so (op_kind HELM_OP_COND8) instead of eleven FFTs; solves
run multigrid-preconditioned FGMRES or BiCGSTAB with FP32 vectors and FP64
true-residual refinement.  There is no CPU path: a missing libhelm.so or GPU
raises.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib
from ._lib import check, lib
from .helmholtz import HelmholtzPlan, _ptr, _stream

C0 = 1540.0
RHO0 = 1000.0


class FrequencyGeometry:
    """FrequencyGeometry (conditioned_wave.py:12-50): the padded box, native
    mask, receivers and source row.  Split-PML factors live in the plan."""

    def __init__(self, spec, device="cuda"):
        self.spec = spec
        self.device = torch.device(device)
        self.h, self.w = spec["medium_shape"][-2:]
        self.dx = spec["dx_m"]
        self.H = math.ceil((self.h + 64) / 32) * 32
        self.W = math.ceil((self.w + 64) / 32) * 32
        if "fft_shape" in spec:
            self.H, self.W = spec["fft_shape"]
            if self.H - self.h < 16 or self.W - self.w < 16:
                raise ValueError("FFT padding must separate the8th-order finite-stencil support")
        self.pz = (self.H - self.h) // 2
        self.px = (self.W - self.w) // 2
        self.pad = (self.px, self.W - self.w - self.px, self.pz, self.H - self.h - self.pz)
        self.mask = F.pad(torch.ones(self.h, self.w, dtype=torch.float64, device=self.device), self.pad)
        self.rx_cols = (torch.round((torch.arange(192, dtype=torch.float64) * .0002 - .0191 - spec["x0_m"])
                                    / self.dx).long() + self.px).to(self.device)
        self.face = round(-spec["z0_m"] / self.dx) + self.pz
        self.dt = spec["native_dt_s"]
        self.pml = spec["pml"]

    def pad_field(self, p):
        return F.pad(p, self.pad)

    def native(self, p):
        return p[..., self.pz:self.pz + self.h, self.px:self.px + self.w]

    def coefficients(self, m):
        """Crop the pml rim, replicate it back, replicate into the box (:45-49)."""
        p = self.pml
        m = F.pad(m[..., p:-p, p:-p][None], (p, p, p, p), mode="replicate")
        return F.pad(m, self.pad, mode="replicate")[0]

    def channels(self, p):
        return p[self.face, self.rx_cols]


def _mg_levels(n, coarsest=40):
    levels = 1
    while n > coarsest and levels < 8:
        n = (n + 1) // 2
        levels += 1
    return levels


class ConditionedFrequencySystem:
    """ConditionedFrequencySystem (conditioned_wave.py:52-92) on a HIP plan.

    ``system(p)`` is the projected operator M A(M p) + (1-M) p of
    dense_wave.projected_operator; it equals the reference's ``__call__`` on
    every field that vanishes outside the medium, which is what all of the
    reference's call sites pass (pad_field outputs and masked fields,
    dense_wave.py:56,253).  nrhs / restart size the plan's Krylov workspace.
    """

    def __init__(self, geometry, medium, alpha, frequency, *, nrhs=16, restart=8, multigrid=None):
        g = geometry
        if tuple(medium.shape) != (2, g.h, g.w):
            raise ValueError("Expected native [2,z,x] medium")
        self.g = g
        self.frequency = float(frequency)
        dev = g.device
        self.D = complex(math.cos(2 * math.pi * self.frequency * g.dt), math.sin(2 * math.pi * self.frequency * g.dt))
        self.q = 1 / self.D
        self.k0 = 2 * math.sin(math.pi * self.frequency * g.dt) / (g.dt * C0)
        self.plan = HelmholtzPlan(g.H, g.W, g.dx * 1e3, self.frequency / 1e6, 0.0, c0_m_s=C0, nrhs=nrhs,
                                  restart=restart, device=dev, op_kind=_lib.HELM_OP_COND8)
        spec = g.spec
        self._med = _lib.HelmConditionedMedium(g.h, g.w, g.pz, g.px, int(g.pml), float(g.dt), RHO0,
                                               (ctypes.c_double * 2)(*spec["eta_per_unit_alpha"]),
                                               (ctypes.c_double * 2)(*spec["taus_s"]))
        self.update(medium, alpha)
        mg = dict(levels=_mg_levels(max(g.H, g.W)), shift=0.5, pre=2, post=2, coarse_sweeps=10)
        mg.update(multigrid or {})
        self.plan.setup_multigrid(**mg)

    def update(self, medium, alpha):
        """New c, rho, alpha maps on the same plan (coefficients and multigrid
        rebuilt on the device; no reallocation)."""
        g = self.g
        if tuple(medium.shape) != (2, g.h, g.w):
            raise ValueError("Expected native [2,z,x] medium")
        medium = torch.as_tensor(medium, device=g.device).to(torch.float64)
        alpha = torch.as_tensor(alpha, device=g.device).to(torch.float64)
        self._c = medium[0].contiguous()
        self._rho = medium[1].contiguous()
        self._alpha = alpha.contiguous()
        check(lib().helm_set_medium_conditioned(self.plan._h, _ptr(self._c), _ptr(self._rho), _ptr(self._alpha),
                                                ctypes.byref(self._med), _stream()), "helm_set_medium_conditioned")
        box = g.coefficients(medium)
        self.c, self.rho = box[0], box[1]
        self.alpha = g.coefficients(alpha[None])[0]
        self.a = RHO0 / self.rho

    def close(self):
        self.plan.close()

    def autotune(self, nrhs=None):
        """Time the streaming operator's launch shapes and the multigrid's
        workgroup target on this plan and keep the fastest (helm_autotune)."""
        return self.plan.autotune(nrhs)

    def __call__(self, p):
        return self.plan.apply(p.contiguous())

    def adjoint(self, p):
        return self.plan.apply(p.contiguous(), adjoint=True)

    def source(self, line):
        """conditioned_wave.py:82-88: one line [w] -> [H, W]."""
        return source_rhs(self, line[None])[0]

    def fixed_preconditioner(self, shift=.15):
        """conditioned_wave.py:90-92: r -> M ifft2(fft2(r) / (1 + lap/k0^2 - i shift)),
        lap = -(|Dx|^2 + |Dz|^2) of the 8-tap symbols, with the library's FFTs
        (helm_fft2, in the precision of r); the solves here use the multigrid
        preconditioner instead."""
        from .spectra import fft2, ifft2
        g = self.g
        coeff = (1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168)

        def symbol(n):
            q = 2 * math.pi * torch.fft.fftfreq(n, dtype=torch.float64, device=g.device)
            return 2j * sum(a * torch.sin((j + .5) * q) for j, a in enumerate(coeff)) * torch.exp(.5j * q) / g.dx

        lap = -(symbol(g.W).abs().square()[None, :] + symbol(g.H).abs().square()[:, None])
        multiplier = 1 / (1 + lap / self.k0 ** 2 - 1j * shift)
        return lambda r: g.mask * ifft2(multiplier.to(r.dtype) * fft2(r.to(g.device)))


def source_rhs(system, line):
    """dense_wave.source_rhs (:59-62): lines [R, w] -> c128 [R, H, W]."""
    g = system.g
    line = torch.as_tensor(line, device=g.device).to(torch.complex128).contiguous()
    if line.dim() != 2 or line.shape[1] != g.w:
        raise ValueError(f"source lines must be [R, {g.w}]")
    out = torch.empty((line.shape[0], g.H, g.W), dtype=torch.complex128, device=g.device)
    zi, xi = g.spec["source_c_lookup_zx"]
    check(lib().helm_conditioned_source(system.plan._h, _ptr(system._c), _ptr(system._rho), _ptr(line), int(g.face),
                                        int(zi), int(xi), _ptr(out), int(line.shape[0]), _stream()),
          "helm_conditioned_source")
    return out


def channels(g, p):
    return p[..., g.face, g.rx_cols]


def projected_operator(system, p):
    """dense_wave.projected_operator (:66-68)."""
    return system(p)


def adjoint_operator(system, p):
    """dense_wave.adjoint_operator (:70-82)."""
    return system.adjoint(p)


def scattered_rhs(system, reference_system, pref, line):
    """The right-hand side DenseData.assemble forms (dense_wave.py:56):
    source(system) - source(reference) - A pref + A_ref pref."""
    pref = pref.to(torch.complex128).contiguous()
    return (source_rhs(system, line) - source_rhs(reference_system, line) - system(pref)
            + reference_system(pref))


def _direction(preconditioner, system, adjoint):
    """The callable a caller's preconditioner stands for.  Objects with a
    ``setup(features)`` method (the reference's LearnedPreconditioner,
    conditioned_wave.py:94-117, used as dense_wave.py:199-201 does) get the
    features namespace (c, rho, alpha as float32 box maps, g, frequency, k0)
    and return the callable; objects with ``adjoint(features)`` give the
    adjoint solve's callable; any other callable is used as it is."""
    from types import SimpleNamespace
    if hasattr(preconditioner, "setup"):
        features = SimpleNamespace(c=system.c.float(), rho=system.rho.float(), alpha=system.alpha.float(),
                                   g=system.g, frequency=system.frequency, k0=system.k0)
        if adjoint:
            if not hasattr(preconditioner, "adjoint"):
                raise ValueError("this preconditioner has no adjoint; adjoint solves take preconditioner=None")
            return preconditioner.adjoint(features)
        return preconditioner.setup(features)
    if callable(preconditioner):
        return preconditioner
    raise TypeError("preconditioner must be callable or have a setup(features) method")


def solve_system(system, rhs, preconditioner=None, adjoint=False, tolerance=1e-5, max_cycles=300, warm_steps=8,
                 mixed=False, method="gmres", initial=None, restart=8):
    """dense_wave.solve_system (:167-225) on the GPU.

    rhs [R, H, W] (complex64 or complex128).  Without a preconditioner every
    right-hand side runs its own Krylov space to ``tolerance`` on its true
    FP64 residual (multigrid FGMRES(``restart``) or BiCGSTAB, FP32 vectors,
    FP64 refinement), so the batch residual the reference tests (all angles
    flattened into one vector, dense_wave.py:88-89) is below ``tolerance`` as
    well; a BiCGSTAB right-hand side that ends unconverged is finished by
    multigrid FGMRES (the reference's fallback to GMRES with the fixed
    preconditioner, :212-216).

    ``preconditioner`` (a callable on [R, H, W] fields, or an object with
    ``setup(features)`` such as the reference's LearnedPreconditioner) is used
    as the reference uses it:
    * forward solves with ``warm_steps`` > 0 and no ``initial``: a warm start
      of one GMRES cycle of ``warm_steps`` steps from zero along its
      directions, all angles one Krylov vector (batched_gmres_cycle,
      :84-116, :196-201);
    * method 'gmres_learned', or 'bicgstab' with a preconditioner: flexible
      GMRES(24) with it throughout (libhelm.so's reverse-communication
      kernels; the reference's BiCGSTAB with the learned preconditioner is
      replaced by this), and any angle left unconverged is finished by
      multigrid FGMRES from that iterate (fallback, :212-216);
    * method 'gmres': after the warm start the fixed (here: multigrid)
      preconditioner solves.
    ``mixed`` only changes the reported dtype: the solve always runs FP32
    vectors with FP64 refinement.  ``max_cycles`` counts cycles of
    ``restart`` steps (default 300).  Returns (value, info) with the
    reference's info keys.
    """
    if method not in ("gmres", "bicgstab", "gmres_learned"):
        raise ValueError(f"unknown method {method!r}")
    start = time.perf_counter()
    dtype = rhs.dtype
    b = rhs.to(torch.complex128).contiguous()
    R = b.shape[0] if b.dim() == 3 else 1
    b = b.reshape(R, system.g.H, system.g.W)
    x0 = None if initial is None else initial.to(torch.complex128).reshape(R, system.g.H, system.g.W).contiguous()
    operator = system.adjoint if adjoint else system
    steps_before, fell = 0, False
    if preconditioner is not None:
        from .krylov import fgmres, gmres_cycle
        direction = _direction(preconditioner, system, adjoint)
        if warm_steps and not adjoint and initial is None:
            x0 = gmres_cycle(operator, direction, b, iterations=int(warm_steps)).contiguous()
            steps_before = int(warm_steps)
        if method in ("gmres_learned", "bicgstab"):
            cycles = max(1, (restart * max_cycles) // 24)
            seen = []
            xs, reasons = fgmres(operator, direction, b, x0 if x0 is not None else torch.zeros_like(b), restart=24,
                                 max_cycles=cycles, tolerance=tolerance, batched=True,
                                 callback=lambda k, xx, r: seen.append(k))
            x0 = xs.contiguous()
            steps_before += seen[-1] if seen else 0
            if all(r == "converged" for r in reasons):
                residual = operator(x0) - b
                return _info_result(x0, residual, b, rhs, dtype, "converged", steps_before, len(seen), start, mixed,
                                    method, tolerance)
            fell = True  # finished by the multigrid solve below
    x, infos = system.plan.solve(b, x0, tolerance=tolerance, restart=restart, max_iterations=restart * max_cycles,
                                 precond="mg", method="bicgstab" if method == "bicgstab" and preconditioner is None
                                 else "gmres", adjoint=adjoint, fallback=method == "bicgstab")
    residual = operator(x) - b
    relative = float(residual.norm() / b.norm().clamp_min(1e-30))
    reason = "converged" if relative < tolerance else next(
        (i.reason for i in infos if i.reason != "converged"), "iteration_budget")
    info = _info_result(x, residual, b, rhs, dtype, reason, steps_before + max(i.iterations for i in infos),
                        max(i.cycles for i in infos), start, mixed, method, tolerance)[1]
    if fell or any(i.fallback for i in infos):
        info["fallback"] = True
    return x.reshape(rhs.shape).to(dtype), info


def _info_result(x, residual, b, rhs, dtype, reason, steps, refinements, start, mixed, method, tolerance):
    relative = float(residual.norm() / b.norm().clamp_min(1e-30))
    if relative >= tolerance and reason == "converged":
        reason = "iteration_budget"
    individual = residual.flatten(1).norm(dim=1) / b.flatten(1).norm(dim=1).clamp_min(1e-30)
    info = {"reason": reason, "steps": int(steps), "residual": relative,
            "max_angle_residual": float(individual.max()), "seconds": time.perf_counter() - start,
            "dtype": "mixed FP32 inverse / FP64 residual" if mixed else str(dtype), "method": method,
            "refinements": int(refinements)}
    return x.reshape(rhs.shape).to(dtype), info


def receiver_response(frequency, dt=2.5e-9):
    """conditioned_data.receiver_response (conditioned_data.py:10-24): the
    squared magnitude of a 4th-order Butterworth band-pass (4-7.5 MHz, zero
    phase) at `frequency` times the real, delay-compensated response of the
    201-tap Kaiser(5) low-pass FIR (cutoff 1/10) of the 10x resampler."""
    from scipy.signal import butter, firwin, freqz, sosfreqz
    band = butter(4, [4e6, 7.5e6], fs=1 / dt, btype="bandpass", output="sos")
    _, hb = sosfreqz(band, worN=[frequency], fs=1 / dt)
    taps = firwin(201, 0.1, window=("kaiser", 5.0))
    w = 2 * np.pi * frequency * dt
    _, hf = freqz(taps, worN=[w])
    return float(abs(hb[0]) ** 2 * (hf[0] * np.exp(1j * w * 100)).real)


class DenseCase:
    """One frequency and angle set of the dense path: what DenseData.assemble
    supplies (dense_wave.py:46-57) -- geometry, frequency, source lines
    [R, w], reference fields pref [R, H, W] (zero when None) and the uniform
    reference system (c 1540, rho 1000, alpha .002).  The system's plan is
    created once and updated with each new medium."""

    def __init__(self, geometry, frequency, lines, pref=None, *, restart=8, multigrid=None):
        g = geometry
        self.g = g
        self.frequency = float(frequency)
        self.lines = torch.as_tensor(lines, device=g.device).to(torch.complex128).contiguous()
        R = self.lines.shape[0]
        self.pref = (torch.zeros((R, g.H, g.W), dtype=torch.complex128, device=g.device) if pref is None
                     else torch.as_tensor(pref, device=g.device).to(torch.complex128).contiguous())
        uniform = torch.empty((2, g.h, g.w), dtype=torch.float64, device=g.device)
        uniform[0] = C0
        uniform[1] = RHO0
        self.reference_system = ConditionedFrequencySystem(g, uniform, torch.full((g.h, g.w), .002, dtype=torch.float64,
                                                                                  device=g.device),
                                                           self.frequency, nrhs=R, restart=restart,
                                                           multigrid=multigrid)
        self.restart, self.multigrid, self.system = restart, multigrid, None

    def assemble(self, medium, alpha):
        if self.system is None:
            self.system = ConditionedFrequencySystem(self.g, medium, alpha, self.frequency, nrhs=self.lines.shape[0],
                                                     restart=self.restart, multigrid=self.multigrid)
        else:
            self.system.update(medium, alpha)
        rhs = scattered_rhs(self.system, self.reference_system, self.pref, self.lines)
        return self.system, rhs

    def close(self):
        for s in (self.system, self.reference_system):
            if s is not None:
                s.close()


def material_gradient(system, v, lam, line=None):
    """d/d(c, rho) of -Re sum_r <lam_r, M (A v_r - s_r)> on the native maps
    (helm_conditioned_gradient); v = u + pref, lam: c128 [R, H, W]."""
    g = system.g
    v = v.to(torch.complex128).contiguous()
    lam = lam.to(torch.complex128).contiguous()
    R = v.shape[0]
    if lam.shape != v.shape:
        raise ValueError("v and lam must have the same shape")
    ln = None if line is None else torch.as_tensor(line, device=g.device).to(torch.complex128).contiguous()
    grad = torch.empty((2, g.h, g.w), dtype=torch.float64, device=g.device)
    zi, xi = g.spec["source_c_lookup_zx"]
    check(lib().helm_conditioned_gradient(system.plan._h, _ptr(system._c), _ptr(system._rho), _ptr(system._alpha),
                                          _ptr(v), _ptr(lam), None if ln is None else _ptr(ln), int(R), int(g.face),
                                          int(zi), int(xi), _ptr(grad[0]), _ptr(grad[1]), _stream()),
          "helm_conditioned_gradient")
    return grad


class ImplicitChannels(torch.autograd.Function):
    """dense_wave.ImplicitChannels (:227-256): receiver channels of the
    converged forward solve times the receiver response; backward solves the
    adjoint system with the channel gradient injected at the receivers and
    forms the material gradient of -Re <lam, M (A (u + pref) - s)> on the GPU
    (alpha fixed).  With a preconditioner the solves take method 'bicgstab'
    and fall back to a solve without it when they do not converge, as the
    reference does (:234-235, :247-249).  Non-convergence raises
    RuntimeError."""

    @staticmethod
    def forward(ctx, medium, alpha, case, preconditioner, tolerance):
        system, rhs = case.assemble(medium.detach().double(), alpha.detach().double())
        method = "bicgstab" if preconditioner is not None else "gmres"
        u, info = solve_system(system, rhs, preconditioner, tolerance=tolerance, mixed=True, method=method)
        if info["reason"] != "converged" and preconditioner is not None:
            u, info = solve_system(system, rhs, None, tolerance=tolerance, mixed=True, method="gmres")
        if info["reason"] != "converged":
            raise RuntimeError("Forward failed to converge: " + str(info))
        ctx.case, ctx.tolerance, ctx.info, ctx.preconditioner = case, tolerance, info, preconditioner
        ctx.save_for_backward(medium, alpha, u)
        return channels(case.g, u) * receiver_response(case.frequency)

    @staticmethod
    def backward(ctx, gradient):
        medium, alpha, u = ctx.saved_tensors
        case = ctx.case
        g = case.g
        system, _ = case.assemble(medium.detach().double(), alpha.detach().double())
        rhs = torch.zeros_like(u)
        rhs[..., g.face, g.rx_cols] = gradient.to(torch.complex128) * receiver_response(case.frequency)
        pre = ctx.preconditioner if ctx.preconditioner is not None and hasattr(ctx.preconditioner, "adjoint") else None
        lam, info = solve_system(system, rhs, pre, adjoint=True, tolerance=ctx.tolerance, mixed=True,
                                 method="bicgstab" if pre is not None else "gmres")
        if info["reason"] != "converged" and pre is not None:
            lam, info = solve_system(system, rhs, None, adjoint=True, tolerance=ctx.tolerance, mixed=True)
        if info["reason"] != "converged":
            raise RuntimeError("Adjoint failed to converge: " + str(info))
        grad = material_gradient(system, u + case.pref, lam, case.lines)
        return grad.to(medium.dtype), None, None, None, None


def implicit_channels(medium, alpha, case, preconditioner=None, tolerance=1e-5):
    """dense_wave.implicit_channels (:258-259); ``case`` (DenseCase) stands
    for the reference's (data, bin_index, angles)."""
    return ImplicitChannels.apply(medium, alpha, case, preconditioner, tolerance)
