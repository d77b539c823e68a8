"""Spectral residual operator on the GPU: drop-in for nsrno/wave_operator.py.

``WaveResidualOperator`` keeps the reference's constructor, geometry
attributes (H, W, pz, px, pad, dx, x, z, sponge, rx_cols, face, angles, trefs)
and methods (setup, apply_operator, incident, source_factor, forward) with
the same arguments and meaning; the arithmetic runs in libhelm.so:

* setup            helm_set_medium_spectral  (wave_operator.py:46-70)
* apply_operator   helm_apply on the 8-tap periodic stencil whose Fourier
                   symbol is the reference's D (wave_operator.py:105-109)
* incident         helm_spectral_incident    (wave_operator.py:111-127)
* the scattered-field right-hand side        helm_spectral_rhs (:142-143)
* _iteration / the warm start of forward     helm_warm_start   (:129-135)

* the warm start's direction      helm_spectral_set_preconditioner + the
                                   library's FFTs: gain * ifft2(multiplier *
                                   fft2(r)) exactly as the reference forms it
                                   (:71-102, 129-135), with a trained model's
                                   context gain, spectral modulation and shift

Differences, by design:
* the Krylov solve after the warm start is preconditioned by the library's
  complex-shifted-Laplacian multigrid (a 5-point discretisation of the same
  coefficients) instead of the Fourier multiplier; the converged solution is
  the same;
* ``setup`` returns an opaque ``SpectralData`` (plan handle and k0) in place of
  the (ax, az, kappa_complex, k0) tensor tuple; ``coefficients`` gives the
  tensors when they are wanted.

Every trained parameter is honoured (``load_wave_operator``): ``alpha_log``
(attenuation), ``response`` (receiver scale), ``log_shift``, ``spectral`` and
the ``context`` net.
"""
from __future__ import annotations

import math
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .helmholtz import HelmholtzPlan, _ptr, _stream

C0 = 1540.0
RHO0 = 1000.0
APERTURE_M = 0.0191
SPONGE_CELLS = 48
SPONGE_AMPLITUDE = 2.0
RELAX_DT = 2.5e-9


def _mg_levels(nz, nx, coarsest=40):
    levels, n = 1, max(nz, nx)
    while n > coarsest and levels < 7:
        n = (n + 1) // 2
        levels += 1
    return max(levels, 2)


class SpectralData:
    """One medium at one frequency on the GPU (the reference's ``data`` tuple)."""

    def __init__(self, plan, k0):
        self.plan = plan
        self.k0 = k0


class WaveResidualOperator:
    """GPU mirror of nsrno/wave_operator.py:14-167 (see the module docstring)."""

    def __init__(self, spec, iterations=8, preconditioner="shifted", *, restart=32, nrhs=16, device="cuda"):
        if preconditioner not in ("shifted", "shifted_tapered", "free_space", "free_space_tapered"):
            raise ValueError(preconditioner)
        self.spec = spec
        self.iterations = int(iterations)
        self.preconditioner = preconditioner
        self.device = torch.device(device)
        h, w = spec["medium_shape"][-2:]
        self.h, self.w = int(h), int(w)
        self.H = math.ceil((self.h + 64) / 32) * 32
        self.W = math.ceil((self.w + 64) / 32) * 32
        self.pz = (self.H - self.h) // 2
        self.px = (self.W - self.w) // 2
        self.pad = (self.px, self.W - self.w - self.px, self.pz, self.H - self.h - self.pz)
        self.dx = float(spec["dx_m"])
        self.x = torch.arange(self.W, dtype=torch.float64) * self.dx + spec["x0_m"] - self.px * self.dx
        self.z = torch.arange(self.H, dtype=torch.float64) * self.dx + spec["z0_m"] - self.pz * self.dx
        iz = torch.minimum(torch.arange(self.H), torch.arange(self.H).flip(0))[:, None]
        ix = torch.minimum(torch.arange(self.W), torch.arange(self.W).flip(0))[None, :]
        self.sponge = SPONGE_AMPLITUDE * ((SPONGE_CELLS - torch.minimum(iz, ix)).clamp_min(0) / SPONGE_CELLS) ** 2
        self.rx_cols = torch.round((torch.arange(192, dtype=torch.float64) * 0.0002 - APERTURE_M - spec["x0_m"])
                                   / self.dx).long() + self.px
        self.face = int(round(-spec["z0_m"] / self.dx)) + self.pz
        self.angles = torch.tensor(spec["angles_deg"], dtype=torch.float64) * math.pi / 180
        self.trefs = torch.tensor(spec["source_tref_s"], dtype=torch.float64)
        # trained scalars (defaults = the reference's initial values)
        self.alpha_log = math.log(0.5)
        self.log_shift = math.log(0.15)
        self.response = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
        # learned preconditioner parts (wave_operator.py:37-41): the context net's
        # weights (None: its zero-initialised last layer, gain 1) and the
        # spectral modulation coefficients
        self.context = None
        self.spectral = np.zeros((2, 4))
        self.training = False
        self.restart_capacity = int(restart)
        self.nrhs_capacity = int(nrhs)
        self._plans = OrderedDict()

    # ------------------------------------------------------------ module-like API
    def eval(self):
        self.training = False
        return self

    def to(self, device):
        self.device = torch.device(device)
        self._plans.clear()
        return self

    def load_state_dict(self, state):
        """Take the scalars that change the solution from a reference state dict."""
        def val(key):
            v = state[key]
            return v.detach().cpu().double().numpy() if torch.is_tensor(v) else np.asarray(v, dtype=np.float64)
        if "alpha_log" in state:
            self.alpha_log = float(val("alpha_log"))
        if "log_shift" in state:
            self.log_shift = float(val("log_shift"))
        if "response" in state:
            self.response = val("response").reshape(2, 3)
        if "spectral" in state:
            self.spectral = val("spectral").reshape(2, 4)
        keys = ("context.0.weight", "context.0.bias", "context.2.weight", "context.2.bias")
        if all(k in state for k in keys):
            self.context = [val(k) for k in keys]
        return self

    @property
    def alpha_scale(self):
        return min(max(math.exp(self.alpha_log), 0.01), 3.0)

    def wavenumber(self, frequency):
        omega = 2 * math.pi * frequency
        if self.spec.get("attenuation") == "relaxation":
            return 2 * math.sin(omega * RELAX_DT / 2) / RELAX_DT / C0
        return omega / C0

    # ------------------------------------------------------------ plans
    def _plan(self, frequency):
        key = float(frequency)
        if key in self._plans:
            self._plans.move_to_end(key)
            return self._plans[key]
        plan = HelmholtzPlan(self.H, self.W, self.dx * 1e3, key / 1e6, 0.0, c0_m_s=C0, nrhs=self.nrhs_capacity,
                             restart=self.restart_capacity, device=self.device, op_kind=_lib.HELM_OP_STAG8)
        plan.medium_key = None
        self._plans[key] = plan
        while len(self._plans) > 4:
            self._plans.popitem(last=False)[1].close()
        return plan

    def _medium_struct(self):
        m = _lib.HelmSpectralMedium()
        m.h, m.w, m.pz, m.px = self.h, self.w, self.pz, self.px
        m.x0_m, m.z0_m = float(self.spec["x0_m"]), float(self.spec["z0_m"])
        m.rho0 = RHO0
        m.alpha_scale = self.alpha_scale
        m.relaxation = int(self.spec.get("attenuation") == "relaxation")
        m.sponge_cells, m.sponge_amplitude = SPONGE_CELLS, SPONGE_AMPLITUDE
        if m.relaxation:
            epa, taus = self.spec["eta_per_unit_alpha"], self.spec["taus_s"]
            m.eta_per_unit_alpha[0], m.eta_per_unit_alpha[1] = float(epa[0]), float(epa[1])
            m.taus_s[0], m.taus_s[1] = float(taus[0]), float(taus[1])
        m.relax_dt_s = RELAX_DT
        return m

    def _set_medium(self, plan, c, rho):
        c = c.to(self.device, torch.float32).contiguous()
        rho = rho.to(self.device, torch.float32).contiguous()
        m = self._medium_struct()
        check(lib().helm_set_medium_spectral(plan._h, _ptr(c), _ptr(rho), m, _stream()), "helm_set_medium_spectral")
        plan._medium = (c, rho)
        if plan.mg_levels == 0:
            plan.setup_multigrid(levels=_mg_levels(self.H, self.W), shift=0.5, pre=2, post=2, coarse_sweeps=10,
                                 omega=0.0)

    def _symbol(self, n):
        q = 2 * np.pi * np.fft.fftfreq(n)
        coeff = (1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168)
        return 2j * sum(a * np.sin((i + .5) * q) for i, a in enumerate(coeff)) * np.exp(.5j * q) / self.dx

    def preconditioner_maps(self, c, rho, frequency, learned=True):
        """(gain [H, W] or None, multiplier [H, W]) of the reference's spectral
        preconditioner (wave_operator.py:71-102, 'shifted' and
        'shifted_tapered'): multiplier exp(.5 tanh(mod0) + .5 i tanh(mod1)) /
        (1 - q^2 - i shift), the modulation from the spectral coefficients;
        gain from the context net on the 4x pooled (log c/C0, -log a,
        f/6 MHz - 1) maps (None when its last layer is zero: gain 1), times the
        sponge taper for 'shifted_tapered'.  Host-side set-up (float64, the
        context net in float32 as the reference's)."""
        if self.preconditioner not in ("shifted", "shifted_tapered"):
            raise NotImplementedError("the free-space Green preconditioner needs a 2H x 2W box; use 'shifted'")
        k0 = self.wavenumber(frequency)
        Dx = self._symbol(self.W)[None, :]
        Dz = self._symbol(self.H)[:, None]
        q2 = (np.abs(Dx) ** 2 + np.abs(Dz) ** 2) / k0 ** 2
        shift = min(max(math.exp(self.log_shift), .03), 1.0) if learned else .15
        mult = 1 / (1 - q2 - 1j * shift)
        if learned and np.any(self.spectral):
            angular = np.abs(Dx) ** 2 / (np.abs(Dx) ** 2 + np.abs(Dz) ** 2 + 1) * np.ones_like(q2)
            basis = np.stack((np.ones_like(q2), q2 / (1 + q2), angular, angular * q2 / (1 + q2)))
            mod = np.einsum("ck,khw->chw", self.spectral, basis)
            mult = np.exp(.5 * np.tanh(mod[0]) + .5j * np.tanh(mod[1])) * mult
        gain = None
        if learned and self.context is not None:
            import torch.nn.functional as F
            cp = torch.as_tensor(c, dtype=torch.float32).cpu()
            a = RHO0 / torch.as_tensor(rho, dtype=torch.float32).cpu()
            cp = F.pad(cp[None, None], self.pad, mode="replicate")[0, 0]
            a = F.pad(a[None, None], self.pad, mode="replicate")[0, 0]
            feats = torch.stack(((cp / C0).log(), -a.log(), torch.ones_like(a) * (frequency / 6e6 - 1)))[None]
            w0, b0, w2, b2 = (torch.as_tensor(v, dtype=torch.float32) for v in self.context)
            g = F.conv2d(torch.tanh(F.conv2d(F.avg_pool2d(feats, 4), w0, b0, padding=1)), w2, b2, padding=1)
            g = F.interpolate(g, size=(self.H, self.W), mode="bilinear", align_corners=False)[0].double().numpy()
            gain = (1 + .5 * np.tanh(g[0])) * np.exp(.4j * np.tanh(g[1]))
        if self.preconditioner == "shifted_tapered":
            taper = np.exp(-k0 * 48 * self.dx / 3 * (self.sponge.numpy() / 2) ** 1.5)
            gain = taper if gain is None else gain * taper
        return gain, mult

    def setup(self, medium, frequency, learned=True):
        """(data, gain, multiplier) as wave_operator.py:46-103; data is a
        SpectralData whose plan also holds the spectral preconditioner
        (gain * ifft2(multiplier * fft2(r)), helm_spectral_set_preconditioner)
        that the warm start applies."""
        if medium.shape[0] != 1:
            raise ValueError("One native medium per frequency solve")
        if tuple(medium.shape[-2:]) != (self.h, self.w):
            raise ValueError(f"medium must be [1, 2, {self.h}, {self.w}]")
        c, rho = medium.detach().float().unbind(1)
        plan = self._plan(frequency)
        self._set_medium(plan, c[0], rho[0])
        gain, mult = self.preconditioner_maps(c[0].cpu().numpy(), rho[0].cpu().numpy(), frequency, learned)
        plan.set_spectral_preconditioner(mult, gain)
        return SpectralData(plan, self.wavenumber(frequency)), (1.0 if gain is None else gain), mult

    def coefficients(self, data):
        """(kappa_S, L applied to ones) read back through the FP64 operator: kappa_S = L 1."""
        one = torch.ones((self.H, self.W), dtype=torch.complex128, device=self.device)
        return data.plan.apply(one, adjoint=True) / data.k0 ** 2

    def apply_operator(self, u, data):
        """L u for complex64 (FP32 stencil) or complex128 (FP64) fields [..., H, W]."""
        u = u.to(self.device).contiguous()
        return data.plan.apply(u, adjoint=True) / data.k0 ** 2

    def incident(self, frequency, angle, data=None):
        """Incident plane wave [H, W] complex64 for spec angle index ``angle``."""
        return self.incident64(frequency, angle, data).to(torch.complex64)

    def incident64(self, frequency, angle, data=None):
        plan = data.plan if data is not None else self._plan(frequency)
        if getattr(plan, "_medium", None) is None:
            water = torch.full((self.h, self.w), C0, dtype=torch.float32)
            self._set_medium(plan, water, torch.full((self.h, self.w), RHO0, dtype=torch.float32))
        out = torch.empty((self.H, self.W), dtype=torch.complex128, device=self.device)
        theta = float(self.angles[angle])
        check(lib().helm_spectral_incident(plan._h, math.sin(theta), APERTURE_M, _ptr(out), _stream()),
              "helm_spectral_incident")
        return out

    def scattered_rhs(self, incident, data):
        """-k0^2 (L - L0) u_inc (complex128), the right-hand side helm_solve(adjoint) takes."""
        inc = incident.to(self.device, torch.complex128).contiguous()
        nrhs = inc.numel() // (self.H * self.W)
        rhs = torch.empty_like(inc)
        check(lib().helm_spectral_rhs(data.plan._h, _ptr(inc), _ptr(rhs), nrhs, _stream()), "helm_spectral_rhs")
        return rhs

    def source_factor(self, frequency, angle):
        """Gaussian-sine source coefficient x exp(-i w t_ref) x response (wave_operator.py:156-167)."""
        f0 = self.spec.get("source_f0_hz", 7.5e6)
        sigma = 1 / f0
        duration = 2401 / 40e6
        omega = 2 * math.pi * frequency
        src = -0.5j * 1e5 * sigma * math.sqrt(2 * math.pi) / duration * (
            math.exp(-0.5 * (sigma * (omega - 2 * math.pi * f0)) ** 2)
            - math.exp(-0.5 * (sigma * (omega + 2 * math.pi * f0)) ** 2))
        src = src * complex(np.exp(-1j * omega * float(self.trefs[angle])))
        f = frequency / 6e6 - 1
        coef = self.response @ np.array([1.0, f, f * f])
        return torch.tensor(src * complex(coef[0], coef[1]), dtype=torch.complex64)

    def forward(self, medium, frequency, angle, learned=True, return_trace=False):
        """Warm-start-only forward (wave_operator.py:137-154): prediction, residual ratio, trace."""
        data, _, _ = self.setup(medium, frequency, learned)
        rhs = self.scattered_rhs(self.incident64(frequency, angle, data), data)
        x = torch.zeros_like(rhs)
        den = rhs.abs().square().mean().clamp_min(1e-20)
        trace = []
        for _ in range(self.iterations if return_trace else 0):
            data.plan.warm_start(rhs, x, 1, precond="spectral", adjoint=True)
            r = rhs - data.plan.apply(x, adjoint=True)
            trace.append(float((r.abs().square().mean() / den).sqrt()))
        if not return_trace:
            data.plan.warm_start(rhs, x, self.iterations, precond="spectral", adjoint=True)
        r = rhs - data.plan.apply(x, adjoint=True)
        loss = r.abs().square().mean() / den
        pred = x[self.face, self.rx_cols.to(self.device)] * self.source_factor(frequency, angle).to(self.device)
        return pred.to(torch.complex64), loss, trace

    __call__ = forward

    @torch.no_grad()
    def predict_band_rf(self, medium, angle, bin_indices):
        """Inference-only RF projected onto the given positive-frequency bins
        (wave_operator.py:169-180): the prediction of ``forward`` at each bin
        frequency index * 40 MHz / 2401, then the inverse real transform to
        2401 samples (helm_band_rf) -> float32 [192, 2401]."""
        from .spectra import band_rf
        bins = [int(b) for b in bin_indices]
        for index in bins:
            if not 0 < index < 1201:
                raise ValueError("Positive RF bin out of range")
        preds = torch.stack([self(medium, index * 40e6 / 2401, angle)[0] for index in bins], dim=1)
        return band_rf(preds, bins, 2401)


def load_wave_operator(path, device="cuda", iterations=None):
    """Build the mirror from a reference checkpoint (weights_only load, nsrno/wave_operator.py:182-186)."""
    saved = torch.load(path, map_location="cpu", weights_only=True)
    model = WaveResidualOperator(saved["spec"], saved["iterations"] if iterations is None else iterations,
                                 saved.get("preconditioner", "shifted"), device=device)
    model.load_state_dict(saved["model"])
    return model.eval()
