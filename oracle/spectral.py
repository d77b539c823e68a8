"""CPU oracle for the spectral-operator (S) frequency solve.

TEST INFRASTRUCTURE ONLY: imported by the tests, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, never by ``nsrno_amd``.  Plain
NumPy/SciPy in float64/complex128.

Restates nsrno/wave_operator.py and the driver nsrno/wave_solver.py:38-67
with a default-initialised model (zero last context layer: learned gain 1,
spectral modulation 0; ``alpha_scale`` = clamp(exp(alpha_log), .01, 3) and the
response polynomial are explicit arguments):

* ``Box``             wave_operator.py:17-36   (padded periodic box, sponge,
                                                coordinates, receiver row/columns)
* ``symbol``          wave_operator.py:22-28   (8th-order staggered derivative)
* ``setup``           wave_operator.py:46-70   (a, b, tissue mask, kappa_c,
                                                harmonic face coefficients, k0)
* ``apply_fft``       wave_operator.py:105-109 (Fourier form of the operator)
* ``apply_stencil``   the same operator as a periodic 8-tap stencil (own form)
* ``matrix``          the same operator assembled as a sparse matrix
* ``incident``        wave_operator.py:111-127 (Hann aperture, discrete-dispersion
                                                Newton for k_z)
* ``scattered_rhs``   wave_operator.py:142-143, wave_solver.py:53
* ``source_factor``   wave_operator.py:156-167
* ``multiplier``      wave_operator.py:93-102 ('shifted', untrained: the
                      Fourier shifted-Laplacian inverse 1/(1 - q^2 - i shift))
* ``solve_gmres``     wave_solver.py:38-67 in complex128: learned-free warm
                      start (wave_operator.py:129-135) and restarted GMRES
                      (oracle.helmholtz.gmres = wave_solver.py:9-36) with the
                      spectral preconditioner
* ``solve_direct``    the same system by splu: the exact solution.
* ``context_gain``    wave_operator.py:71-75 (learned context conv net gain)
* ``modulated_multiplier`` wave_operator.py:93-102 with the learned spectral
                      modulation and shift
* ``forward``         wave_operator.py:129-154 (warm-start-only prediction,
                      loss and trace)

S convention: torch/NumPy forward FFT kernel, outgoing waves exp(-i k r),
sponge and attenuation as -i terms (wave_operator.py:3,67).

How each function is pinned (tests/test_oracle_spectral.py):
* setup, incident, apply, source_factor, solve: golden vectors written by
  tests/golden/make_golden_s.py from the reference's own WaveResidualOperator
  and solve_frequency (no closed form pins the tissue mask, the sponge, the
  Newton iteration or the source spectrum);
* apply_fft == apply_stencil == matrix @ u (symbol vs stencil, brute-force
  columns); real coefficients give a real symmetric matrix;
* incident: the Newton root satisfies the discrete dispersion relation;
* solve_direct: residual < 1e-10 of the assembled system;
* context_gain, modulated_multiplier, forward: golden gain map, multiplier
  and forward predictions/losses/traces of the reference's own model, at its
  default initialisation and with perturbed context net, modulation and shift;
* solve_gmres: equal to solve_direct at convergence; the warm-start steps
  never increase the residual (scripts/check_wave.py:18); the reference's
  own complex64 prediction (golden) within its stagnation error.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy import sparse
from scipy.sparse.linalg import splu

FD = (1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168)
C0 = 1540.0
RHO0 = 1000.0
SPONGE_CELLS = 48
APERTURE_M = 0.0191
RELAX_DT = 2.5e-9


@dataclass
class Box:
    """Padded periodic box around an h x w medium (wave_operator.py:17-36)."""
    h: int
    w: int
    dx: float
    x0: float
    z0: float

    def __post_init__(self):
        self.H = math.ceil((self.h + 64) / 32) * 32
        self.W = math.ceil((self.w + 64) / 32) * 32
        self.pz = (self.H - self.h) // 2
        self.px = (self.W - self.w) // 2
        self.x = np.arange(self.W) * self.dx + self.x0 - self.px * self.dx
        self.z = np.arange(self.H) * self.dx + self.z0 - self.pz * self.dx
        iz = np.minimum(np.arange(self.H), np.arange(self.H)[::-1])[:, None]
        ix = np.minimum(np.arange(self.W), np.arange(self.W)[::-1])[None, :]
        self.sponge = 2.0 * (np.clip(SPONGE_CELLS - np.minimum(iz, ix), 0, None) / SPONGE_CELLS) ** 2
        self.rx_cols = np.round((np.arange(192) * 0.0002 - APERTURE_M - self.x0) / self.dx).astype(np.int64) + self.px
        self.face = int(round(-self.z0 / self.dx)) + self.pz


def symbol(n, dx):
    """Fourier symbol of the forward staggered derivative on n periodic points."""
    q = 2 * np.pi * np.fft.fftfreq(n)
    return 2j * sum(a * np.sin((i + 0.5) * q) for i, a in enumerate(FD)) * np.exp(0.5j * q) / dx


def wavenumber(frequency, attenuation=None):
    """k0 (wave_operator.py:55): omega/C0, or the 2.5 ns leapfrog-consistent value."""
    omega = 2 * np.pi * frequency
    if attenuation == "relaxation":
        return 2 * math.sin(omega * RELAX_DT / 2) / RELAX_DT / C0
    return omega / C0


def _pad(a, box):
    zi = np.clip(np.arange(box.H) - box.pz, 0, box.h - 1)
    xi = np.clip(np.arange(box.W) - box.px, 0, box.w - 1)
    return a[np.ix_(zi, xi)]


def setup(box, c, rho, frequency, alpha_scale=0.5, attenuation=None, eta_per_unit_alpha=None, taus=None):
    """(ax, az, kappa_c, k0) of wave_operator.py:46-70 (float64)."""
    c = np.asarray(c, dtype=np.float64)
    rho = np.asarray(rho, dtype=np.float64)
    a = _pad(RHO0 / rho, box)
    b = a * (C0 / _pad(c, box)) ** 2
    cp = _pad(c, box)
    omega = 2 * np.pi * frequency
    k0 = wavenumber(frequency, attenuation)
    alpha = alpha_scale * 100 * math.log(10) / 20 * (frequency / 1e6) ** 1.1
    # z >= 0 up to rounding: the row at z = 0 counts as tissue however its coordinate rounds
    tissue = (box.z[:, None] >= -1e-6 * box.dx) * (1 - np.exp(-(np.log(cp / C0) / 0.03) ** 2 - (np.log(a) / 0.05) ** 2))
    if attenuation == "relaxation":
        eta = [alpha_scale * tissue * cp / C0 * v for v in eta_per_unit_alpha]
        amplitude = (1 + 0.5 * (eta[0] + eta[1])) ** 2
        step = np.exp(1j * omega * RELAX_DT)
        hsum = sum(e / (amplitude * (1 + t * (step - 1) / RELAX_DT)) for e, t in zip(eta, taus))
        modulus = amplitude * (1 - hsum) / (1 - (1 - 1 / step) * hsum)
        kc = b / modulus - 1j * box.sponge
    else:
        kc = b * (1 - 1j * alpha * tissue * cp / omega) ** 2 - 1j * box.sponge
    ax = 2 * a * np.roll(a, -1, 1) / (a + np.roll(a, -1, 1))
    az = 2 * a * np.roll(a, -1, 0) / (a + np.roll(a, -1, 0))
    return ax, az, kc, k0


def apply_fft(u, ax, az, kc, k0, dx):
    """L u in the Fourier form of wave_operator.py:105-109."""
    H, W = u.shape
    Dx = symbol(W, dx)[None, :]
    Dz = symbol(H, dx)[:, None]
    s = np.fft.fft2(u)
    gx = np.fft.ifft2(Dx * s)
    gz = np.fft.ifft2(Dz * s)
    div = np.fft.ifft2(-np.conj(Dx) * np.fft.fft2(ax * gx) - np.conj(Dz) * np.fft.fft2(az * gz))
    return div / k0 ** 2 + kc * u


def _dplus(u, axis, dx):
    """(D+ u)_j = sum_m a_m (u_{j+m+1} - u_{j-m}) / dx, periodic (face j sits at j + 1/2)."""
    return sum(a * (np.roll(u, -(m + 1), axis) - np.roll(u, m, axis)) for m, a in enumerate(FD)) / dx


def _dminus(g, axis, dx):
    """(D- g)_j = sum_m a_m (g_{j+m} - g_{j-m-1}) / dx, periodic: D- = -(D+)^T."""
    return sum(a * (np.roll(g, -m, axis) - np.roll(g, m + 1, axis)) for m, a in enumerate(FD)) / dx


def apply_stencil(u, ax, az, kc, k0, dx):
    """The same operator as a real-space periodic 8-tap staggered stencil."""
    div = _dminus(ax * _dplus(u, -1, dx), -1, dx) + _dminus(az * _dplus(u, -2, dx), -2, dx)
    return div / k0 ** 2 + kc * u


def _circ_plus(n, dx):
    rows, cols, vals = [], [], []
    for j in range(n):
        for m, a in enumerate(FD):
            rows += [j, j]
            cols += [(j + m + 1) % n, (j - m) % n]
            vals += [a / dx, -a / dx]
    return sparse.csr_matrix((vals, (rows, cols)), shape=(n, n))


def matrix(ax, az, kc, k0, dx):
    """Sparse complex128 matrix of L on the flattened [z][x] box."""
    H, W = ax.shape
    dpx = sparse.kron(sparse.identity(H), _circ_plus(W, dx))
    dpz = sparse.kron(_circ_plus(H, dx), sparse.identity(W))
    lx = -dpx.T @ sparse.diags(ax.ravel()) @ dpx
    lz = -dpz.T @ sparse.diags(az.ravel()) @ dpz
    return ((lx + lz) / k0 ** 2 + sparse.diags(kc.ravel())).astype(np.complex128).tocsc()


def incident(box, frequency, theta_rad, attenuation=None):
    """Plane-wave incident field on the box (wave_operator.py:111-127)."""
    k0 = wavenumber(frequency, attenuation)
    x = box.x
    aperture = np.where(np.abs(x) <= APERTURE_M, 0.5 * (1 + np.cos(np.pi * x / APERTURE_M)), 0.0)
    boundary = aperture * np.exp(-1j * k0 * x * math.sin(theta_rad))
    sx = np.abs(symbol(box.W, box.dx)) ** 2
    desired = np.sqrt(np.clip(k0 ** 2 - sx, 0, None)) * box.dx
    q = np.clip(desired, 0, np.pi)
    for _ in range(6):
        value = 2 * sum(a * np.sin((j + 0.5) * q) for j, a in enumerate(FD))
        deriv = 2 * sum(a * (j + 0.5) * np.cos((j + 0.5) * q) for j, a in enumerate(FD))
        q = np.clip(q - (value - desired) / np.maximum(deriv, 0.05), 0, np.pi)
    kz = q / box.dx - 1j * np.sqrt(np.clip(sx - k0 ** 2, 0, None))
    prop = np.exp(-1j * np.abs(box.z)[:, None] * kz[None, :])
    return np.fft.ifft(np.fft.fft(boundary)[None, :] * prop, axis=-1), q, desired


def scattered_rhs(box, inc, ax, az, kc, k0):
    """-(L - L0) u_inc with L0 = (ax = az = 1, kappa = 1 - i sponge) (wave_operator.py:142-143)."""
    return -apply_stencil(inc, ax - 1, az - 1, kc - (1 - 1j * box.sponge), k0, box.dx)


def source_factor(frequency, tref, f0=7.5e6, response=((1.0, 0.0, 0.0), (0.0, 0.0, 0.0))):
    """Gaussian-sine source coefficient times the response polynomial (wave_operator.py:156-167)."""
    sigma = 1 / f0
    duration = 2401 / 40e6
    omega = 2 * np.pi * frequency
    src = -0.5j * 1e5 * sigma * math.sqrt(2 * math.pi) / duration * (
        math.exp(-0.5 * (sigma * (omega - 2 * math.pi * f0)) ** 2)
        - math.exp(-0.5 * (sigma * (omega + 2 * math.pi * f0)) ** 2))
    src = src * np.exp(-1j * omega * tref)
    f = frequency / 6e6 - 1
    coef = np.asarray(response, dtype=np.float64) @ np.array([1.0, f, f * f])
    return src * complex(coef[0], coef[1])


def solve_direct(box, c, rho, frequency, theta_rad, tref, alpha_scale=0.5, attenuation=None,
                 eta_per_unit_alpha=None, taus=None, f0=7.5e6):
    """(prediction [192], field [H, W], relative residual) of the exact frequency solve."""
    ax, az, kc, k0 = setup(box, c, rho, frequency, alpha_scale, attenuation, eta_per_unit_alpha, taus)
    inc, _, _ = incident(box, frequency, theta_rad, attenuation)
    rhs = scattered_rhs(box, inc, ax, az, kc, k0)
    A = matrix(ax, az, kc, k0, box.dx)
    u = splu(A).solve(rhs.ravel()).reshape(rhs.shape)
    rel = np.linalg.norm(A @ u.ravel() - rhs.ravel()) / np.linalg.norm(rhs)
    pred = u[box.face, box.rx_cols] * source_factor(frequency, tref, f0)
    return pred, u, rel


def multiplier(box, k0, shift=0.15):
    """Fourier multiplier of the untrained 'shifted' preconditioner (wave_operator.py:93-102)."""
    Dx = symbol(box.W, box.dx)[None, :]
    Dz = symbol(box.H, box.dx)[:, None]
    q2 = (np.abs(Dx) ** 2 + np.abs(Dz) ** 2) / k0 ** 2
    return 1 / (1 - q2 - 1j * shift)


def _conv3x3(x, w, b):
    """Conv2d(k=3, padding=1) of x [Cin, H, W] with w [Cout, Cin, 3, 3], b [Cout]."""
    cin, H, W = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1)))
    out = np.empty((w.shape[0], H, W))
    for o in range(w.shape[0]):
        acc = np.full((H, W), float(b[o]))
        for i in range(cin):
            for dz in range(3):
                for dx in range(3):
                    acc += w[o, i, dz, dx] * xp[i, dz:dz + H, dx:dx + W]
        out[o] = acc
    return out


def _bilinear(x, H, W):
    """F.interpolate(x, size=(H, W), mode='bilinear', align_corners=False) of x [C, h, w]."""
    c, h, w = x.shape

    def taps(n_out, n_in):
        s = (np.arange(n_out) + 0.5) * n_in / n_out - 0.5
        s = np.clip(s, 0, None)
        i0 = np.minimum(np.floor(s).astype(int), n_in - 1)
        i1 = np.minimum(i0 + 1, n_in - 1)
        return i0, i1, s - i0

    z0, z1, fz = taps(H, h)
    x0, x1, fx = taps(W, w)
    top = x[:, z0][:, :, x0] * (1 - fx) + x[:, z0][:, :, x1] * fx
    bot = x[:, z1][:, :, x0] * (1 - fx) + x[:, z1][:, :, x1] * fx
    return top * (1 - fz)[None, :, None] + bot * fz[None, :, None]


def context_gain(box, c, rho, frequency, params):
    """The learned complex gain of wave_operator.py:71-75: a 3 -> 12 -> 2 conv
    net (tanh) on the 4x average-pooled (log c/C0, -log a, f/6 MHz - 1) maps,
    bilinearly upsampled to the box, gain = (1 + .5 tanh g0) exp(.4 i tanh g1).
    params: context_0_weight/bias, context_2_weight/bias arrays."""
    c = np.asarray(c, dtype=np.float64)
    a = _pad(RHO0 / np.asarray(rho, dtype=np.float64), box)
    cp = _pad(c, box)
    feats = np.stack((np.log(cp / C0), -np.log(a), np.full(a.shape, frequency / 6e6 - 1)))
    pooled = feats.reshape(3, box.H // 4, 4, box.W // 4, 4).mean(axis=(2, 4))
    hid = np.tanh(_conv3x3(pooled, params["context_0_weight"], params["context_0_bias"]))
    g = _bilinear(_conv3x3(hid, params["context_2_weight"], params["context_2_bias"]), box.H, box.W)
    return (1 + 0.5 * np.tanh(g[0])) * np.exp(0.4j * np.tanh(g[1]))


def modulated_multiplier(box, k0, shift=0.15, spectral=None):
    """The 'shifted' multiplier with the learned spectral modulation
    exp(.5 tanh(mod0) + .5 i tanh(mod1)), mod = spectral @ (1, q^2/(1+q^2),
    angular, angular q^2/(1+q^2)) (wave_operator.py:93-102)."""
    Dx = symbol(box.W, box.dx)[None, :]
    Dz = symbol(box.H, box.dx)[:, None]
    q2 = (np.abs(Dx) ** 2 + np.abs(Dz) ** 2) / k0 ** 2
    green = 1 / (1 - q2 - 1j * shift)
    if spectral is None:
        return green
    angular = np.abs(Dx) ** 2 / (np.abs(Dx) ** 2 + np.abs(Dz) ** 2 + 1)
    basis = np.stack((np.ones_like(q2), q2 / (1 + q2), angular * np.ones_like(q2), angular * q2 / (1 + q2)))
    mod = np.einsum("ck,khw->chw", np.asarray(spectral, dtype=np.float64), basis)
    return np.exp(0.5 * np.tanh(mod[0]) + 0.5j * np.tanh(mod[1])) * green


def forward(box, c, rho, frequency, theta_rad, tref, iterations=8, gain=1.0, mult=None, alpha_scale=0.5,
            f0=7.5e6):
    """The warm-start-only forward of wave_operator.py:137-154 in complex128:
    `iterations` steps u += alpha d, d = gain ifft2(mult fft2(r)),
    alpha = <L d, r> / ||L d||^2 (:129-135).  Returns (prediction, loss =
    mean|r|^2 / mean|rhs|^2, trace of sqrt(loss) after every step)."""
    ax, az, kc, k0 = setup(box, c, rho, frequency, alpha_scale)
    inc, _, _ = incident(box, frequency, theta_rad)
    rhs = scattered_rhs(box, inc, ax, az, kc, k0)
    mult = multiplier(box, k0) if mult is None else mult

    def op(u):
        return apply_stencil(u, ax, az, kc, k0, box.dx)

    den = max(np.mean(np.abs(rhs) ** 2), 1e-20)
    u = np.zeros_like(rhs)
    trace = []
    for _ in range(iterations):
        r = rhs - op(u)
        d = gain * np.fft.ifft2(mult * np.fft.fft2(r))
        image = op(d)
        u = u + np.vdot(image, r) / max(np.vdot(image, image).real, 1e-24) * d
        trace.append(float(np.sqrt(np.mean(np.abs(rhs - op(u)) ** 2) / den)))
    loss = float(np.mean(np.abs(rhs - op(u)) ** 2) / den)
    pred = u[box.face, box.rx_cols] * source_factor(frequency, tref, f0)
    return pred, loss, trace


def solve_gmres(box, c, rho, frequency, theta_rad, tref, iterations=8, tolerance=1e-10, restart=32,
                max_iterations=4096, alpha_scale=0.5, f0=7.5e6):
    """(prediction, field, info) of solve_frequency in complex128 arithmetic."""
    from .helmholtz import gmres
    ax, az, kc, k0 = setup(box, c, rho, frequency, alpha_scale)
    inc, _, _ = incident(box, frequency, theta_rad)
    rhs = scattered_rhs(box, inc, ax, az, kc, k0)
    mult = multiplier(box, k0)

    def op(u):
        return apply_stencil(u, ax, az, kc, k0, box.dx)

    def pre(r):
        return np.fft.ifft2(mult * np.fft.fft2(r))

    u = np.zeros_like(rhs)
    warm = []
    for _ in range(iterations):  # wave_operator.py:129-135 with gain 1
        r = rhs - op(u)
        d = pre(r)
        image = op(d)
        u = u + np.vdot(image, r) / max(np.vdot(image, image).real, 1e-24) * d
        warm.append(np.linalg.norm(rhs - op(u)) / np.linalg.norm(rhs))
    trace = []
    u, reason = gmres(op, pre, rhs, u, lambda a, x, r: trace.append(dict(iterations=a, relative_residual=r)),
                      restart=restart, max_cycles=max_iterations // restart, tolerance=tolerance)
    pred = u[box.face, box.rx_cols] * source_factor(frequency, tref, f0)
    return pred, u, dict(reason=reason, trace=trace, warm=warm)
