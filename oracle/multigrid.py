"""CPU restatement of the device multigrid V-cycle (complex-shifted Laplacian).

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker of the HIP
preconditioner).  The reference has no multigrid: its preconditioner is the
FFT shifted-Laplacian multiplier of nsrno/wave_operator.py:93-96.  This
restatement is pinned by (a) equality of one V-cycle with the HIP one on the
same input, (b) Richardson iteration with it converging to the direct
(splu) solution of the shifted operator, (c) for constant
coefficients without PML, the per-node weight's predicted worst |g| equals
the spectral radius of the Jacobi iteration matrix computed by brute force.
"""
from __future__ import annotations

import numpy as np

from . import helmholtz as H


def node_weights(coef, coarsest, cap=1.1):
    """Per-node complex Jacobi weight from the local Fourier symbol.

    A mode with cos(theta_x) = t1, cos(theta_z) = t2 is multiplied by
    g = 1 - w z, z = 1 + t1 a + t2 b, a = (fw + fe)/dg, b = (fn + fs)/dg.
    Candidates w = r e^{i phi}, phi = -pi/2 + pi k/32, r = 2 m/32; smoothing
    levels minimise max |g| over the corners of the oscillatory region
    (t1 <= 0 or t2 <= 0) among weights with max |g| <= cap over the full
    square, else (and on the coarsest level) minimise max |g| over the square.
    """
    dg, fw, fe, fn, fs = coef
    a = (fw + fe) / dg
    b = (fn + fs) / dg
    t1 = np.array([-1, -1, 1, 1, -1, -1, 0, 0, -1, 1, 1], float)
    t2 = np.array([-1, 1, -1, 1, -1, 1, -1, 1, 0, -1, 0], float)
    z = 1 + t1[:, None, None] * a[None] + t2[:, None, None] * b[None]          # [11, nz, nx]
    phi = -np.pi / 2 + np.pi * np.arange(33) / 32
    r = 2 * np.arange(1, 33) / 32
    cand = (r[None, :] * np.exp(1j * phi[:, None])).ravel()                    # [C]
    g2 = np.abs(1 - cand[:, None, None, None] * z[None]) ** 2                   # [C, 11, nz, nx]
    full = g2[:, :4].max(axis=1)
    hi = g2[:, 4:].max(axis=1)
    score = full if coarsest else np.where(full <= cap * cap, hi, 1e3 + full)
    return cand[np.argmin(score, axis=0)]


class Level:
    def __init__(self, fine_grid, h_mm, nz, nx, frequency_mhz, c0, pml_strength):
        self.nz, self.nx, self.h_mm = nz, nx, h_mm
        g = fine_grid
        x = g.x_mm[0] + h_mm * np.arange(nx)
        z = g.z_mm[0] + h_mm * np.arange(nz)
        xe = g.x_mm[0] + h_mm * (np.arange(nx + 1) - 0.5)
        ze = g.z_mm[0] + h_mm * (np.arange(nz + 1) - 0.5)
        omega = 2 * np.pi * frequency_mhz * 1e6
        args = dict(pml_mm=g.pml_mm, omega=omega, c0=c0, pml_strength=pml_strength)
        self.sx = H.stretch(x, g.x_mm[0], g.x_mm[-1], **args)
        self.sz = H.stretch(z, g.z_mm[0], g.z_mm[-1], **args)
        self.sxe = H.stretch(xe, g.x_mm[0], g.x_mm[-1], **args)
        self.sze = H.stretch(ze, g.z_mm[0], g.z_mm[-1], **args)
        self.k0 = omega / c0
        self.h_m = h_mm * 1e-3

    def coef(self, eta, kap, shift):
        """Stencil coefficients (dg, fw, fe, fn, fs) of the shifted operator."""
        h2 = self.h_m ** 2
        ep = np.pad(eta, 1, constant_values=1.0)
        fw = self.sz[:, None] / self.sxe[None, :-1] * 0.5 * (eta + ep[1:-1, :-2]) / h2
        fe = self.sz[:, None] / self.sxe[None, 1:] * 0.5 * (eta + ep[1:-1, 2:]) / h2
        fn = self.sx[None, :] / self.sze[:-1, None] * 0.5 * (eta + ep[:-2, 1:-1]) / h2
        fs = self.sx[None, :] / self.sze[1:, None] * 0.5 * (eta + ep[2:, 1:-1]) / h2
        mass = self.k0 ** 2 * self.sz[:, None] * self.sx[None, :] * kap * (1 + 1j * shift)
        return mass - (fw + fe + fn + fs), fw, fe, fn, fs

    def apply(self, c, u):
        dg, fw, fe, fn, fs = c
        up = np.pad(u, [(0, 0), (1, 1), (1, 1)])
        return (dg * u + fw * up[:, 1:-1, :-2] + fe * up[:, 1:-1, 2:] + fn * up[:, :-2, 1:-1]
                + fs * up[:, 2:, 1:-1])


def coarsen_medium(eta, kap):
    nzf, nxf = eta.shape
    nzc, nxc = (nzf + 1) // 2, (nxf + 1) // 2
    w = np.array([0.25, 0.5, 0.25])
    ec = np.zeros((nzc, nxc))
    kc = np.zeros((nzc, nxc), complex)
    for a in (-1, 0, 1):
        i = np.clip(2 * np.arange(nzc) + a, 0, nzf - 1)
        for b in (-1, 0, 1):
            j = np.clip(2 * np.arange(nxc) + b, 0, nxf - 1)
            ec += w[a + 1] * w[b + 1] * eta[np.ix_(i, j)]
            kc += w[a + 1] * w[b + 1] * kap[np.ix_(i, j)]
    return ec, kc


def restrict(r, nzc, nxc):
    R, nzf, nxf = r.shape
    w = np.array([0.25, 0.5, 0.25])
    rp = np.pad(r, [(0, 0), (1, 2), (1, 2)])
    out = np.zeros((R, nzc, nxc), complex)
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            out += w[a + 1] * w[b + 1] * rp[:, 1 + a:1 + a + 2 * nzc:2, 1 + b:1 + b + 2 * nxc:2]
    return out


def prolong(uc, nzf, nxf):
    R, nzc, nxc = uc.shape
    up = np.pad(uc, [(0, 0), (0, 1), (0, 1)])
    out = np.zeros((R, 2 * nzc, 2 * nxc), complex)
    out[:, 0::2, 0::2] = up[:, :-1, :-1]
    out[:, 1::2, 0::2] = 0.5 * (up[:, :-1, :-1] + up[:, 1:, :-1])
    out[:, 0::2, 1::2] = 0.5 * (up[:, :-1, :-1] + up[:, :-1, 1:])
    out[:, 1::2, 1::2] = 0.25 * (up[:, :-1, :-1] + up[:, 1:, :-1] + up[:, :-1, 1:] + up[:, 1:, 1:])
    return out[:, :nzf, :nxf]


class Multigrid:
    """V-cycle for A(kappa_c (1 + i shift)) with the device's components."""

    def __init__(self, grid, frequency_mhz, eta, kap, *, levels=5, shift=0.5, pre=2, post=2, coarse=30,
                 omega=0.0, c0=1540.0, pml_strength=12.0):
        self.shift, self.pre, self.post, self.coarse = shift, pre, post, coarse
        self.levels, self.coefs, self.weights = [], [], []
        nz, nx = grid.shape
        e, k = np.asarray(eta, float).reshape(nz, nx), np.asarray(kap, complex).reshape(nz, nx)
        h = grid.h_mm
        for l in range(levels):
            if l > 0:
                nzc, nxc = (nz + 1) // 2, (nx + 1) // 2
                if nzc < 3 or nxc < 3:
                    break
                e, k = coarsen_medium(e, k)
                nz, nx, h = nzc, nxc, 2 * h
            lev = Level(grid, h, nz, nx, frequency_mhz, c0, pml_strength)
            self.levels.append(lev)
            self.coefs.append(lev.coef(e, k, shift))
        for l, c in enumerate(self.coefs):
            if omega > 0:
                self.weights.append(np.full(c[0].shape, complex(omega)))
            else:
                self.weights.append(node_weights(c, l + 1 == len(self.levels)))

    def jacobi(self, l, u, f):
        c = self.coefs[l]
        w = self.weights[l] / c[0]
        if u is None:
            return w * f
        return u + w * (f - self.levels[l].apply(c, u))

    def vcycle(self, f, l=0):
        coarsest = l + 1 == len(self.levels)
        u = None
        for _ in range(self.coarse if coarsest else self.pre):
            u = self.jacobi(l, u, f)
        if u is None:
            u = np.zeros_like(f)
        if coarsest:
            return u
        r = f - self.levels[l].apply(self.coefs[l], u)
        nxt = self.levels[l + 1]
        uc = self.vcycle(restrict(r, nxt.nz, nxt.nx), l + 1)
        u = u + prolong(uc, self.levels[l].nz, self.levels[l].nx)
        for _ in range(self.post):
            u = self.jacobi(l, u, f)
        return u
