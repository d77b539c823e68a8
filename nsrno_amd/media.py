"""Seeded synthetic media and acquisitions for the benchmark workloads.

All media are [z, x] float64 maps on the full grid (PML included), seeded
with ``numpy.random.default_rng(seed)``:

* ``inclusion_medium``  - water background with Gaussian inclusions that
  perturb c and rho (the 128^2 + 20-cell-PML direct-solve case at 1 MHz);
  the perturbation vanishes in the PML so the theranostic contrast checks
  (wave.py:158-172) accept it.
* ``smooth_random_medium`` - Gaussian-filtered white noise rescaled to
  c in [1400, 1600] m/s, rho in [900, 1100] kg/m^3 and a power-law
  attenuation alpha0 in [0.3, 1.0] dB/cm/MHz^1.1, generated on the interior
  and replicated into the PML (the S-path padding rule,
  nsrno/wave_operator.py:51-53).
* ``line_array`` / ``plane_wave_amplitudes`` - a linear array of
  finite-line elements one cell inside the top PML edge, transmitting plane
  waves by linear phase delays with a Hann apodisation (the S path's
  19.1 mm Hann aperture and exp(-i k0 x sin(theta)) boundary phase,
  nsrno/wave_operator.py:111-116, in the exp(-i omega t) convention).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.ndimage import gaussian_filter


@dataclass
class Workload:
    """A grid + medium + acquisition: everything one batched solve needs."""
    nz: int
    nx: int
    h_mm: float
    pml_cells: int
    frequency_mhz: float
    c0: float
    rho0: float
    c: np.ndarray
    rho: np.ndarray
    alpha_np_m: np.ndarray | None
    x0_mm: float = 0.0
    z0_mm: float = 0.0

    @property
    def pml_mm(self):
        return self.pml_cells * self.h_mm

    @property
    def x_mm(self):
        return self.x0_mm + self.h_mm * np.arange(self.nx)

    @property
    def z_mm(self):
        return self.z0_mm + self.h_mm * np.arange(self.nz)

    def eta_rel(self):
        return self.rho0 / self.rho

    def kappa_c(self):
        omega = 2 * np.pi * self.frequency_mhz * 1e6
        k = self.rho0 * self.c0 ** 2 / (self.rho * self.c ** 2)
        if self.alpha_np_m is None:
            return k.astype(np.complex128)
        return k * (1 + 1j * self.alpha_np_m * self.c / omega) ** 2


def alpha_np_per_m(alpha0_db_cm_mhz, frequency_mhz, power=1.1):
    """dB/cm/MHz^y at f MHz -> Np/m (unit conversion of wave_operator.py:56)."""
    return np.asarray(alpha0_db_cm_mhz) * 100 * np.log(10) / 20 * frequency_mhz ** power


def spacing_mm(c_min, frequency_mhz, ppw=10):
    """dx = lambda_min / ppw."""
    return c_min / (frequency_mhz * 1e6) / ppw * 1e3


def inclusion_medium(seed=0, interior=128, pml=20, frequency_mhz=1.0, n_inclusions=5, c0=1500.0, rho0=1000.0,
                     sigma_cells=(3, 10), dc=60.0, drho=80.0):
    rng = np.random.default_rng(seed)
    n = interior + 2 * pml
    h = spacing_mm(c0 - dc, frequency_mhz)
    zz, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    c = np.full((n, n), c0)
    rho = np.full((n, n), rho0)
    for _ in range(n_inclusions):
        s = rng.uniform(*sigma_cells)
        cz, cx = rng.uniform(pml + 2 * s, n - pml - 1 - 2 * s, size=2)
        g = np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * s * s))
        c += rng.uniform(-dc, dc) * g
        rho += rng.uniform(-drho, drho) * g
    inside = np.zeros((n, n), bool)
    inside[pml:n - pml, pml:n - pml] = True
    c[~inside] = c0
    rho[~inside] = rho0
    return Workload(n, n, h, pml, frequency_mhz, c0, rho0, c, rho, None)


def _smooth_field(rng, shape, sigma, lo, hi):
    f = gaussian_filter(rng.standard_normal(shape), sigma, mode="wrap")
    f = (f - f.min()) / max(f.max() - f.min(), 1e-30)
    return lo + (hi - lo) * f


def smooth_random_medium(seed=0, interior=512, pml=32, frequency_mhz=3.0, sigma_cells=8.0, c_range=(1400, 1600),
                         rho_range=(900, 1100), alpha0_range=(0.3, 1.0), c0=1500.0, rho0=1000.0, h_mm=None,
                         attenuation=True):
    rng = np.random.default_rng(seed)
    h = spacing_mm(c_range[0], frequency_mhz) if h_mm is None else h_mm
    shape = (interior, interior)
    c = np.pad(_smooth_field(rng, shape, sigma_cells, *c_range), pml, mode="edge")
    rho = np.pad(_smooth_field(rng, shape, sigma_cells, *rho_range), pml, mode="edge")
    alpha = None
    if attenuation:
        a0 = np.pad(_smooth_field(rng, shape, sigma_cells, *alpha0_range), pml, mode="edge")
        alpha = alpha_np_per_m(a0, frequency_mhz)
    n = interior + 2 * pml
    return Workload(n, n, h, pml, frequency_mhz, c0, rho0, c, rho, alpha)


def line_array(work: Workload, n_elements, margin_cells=None, row_offset_cells=1.0):
    """Element centres [E,2] (x,z mm), normals and widths of a linear array
    spanning the interior width at one cell below the top PML edge."""
    pml = work.pml_cells
    margin = pml + 2 if margin_cells is None else margin_cells
    x_lo = work.x0_mm + margin * work.h_mm
    x_hi = work.x0_mm + (work.nx - 1 - margin) * work.h_mm
    pitch = (x_hi - x_lo) / n_elements
    xs = x_lo + pitch * (np.arange(n_elements) + 0.5)
    z = work.z0_mm + (pml + row_offset_cells) * work.h_mm
    positions = np.stack((xs, np.full(n_elements, z)), axis=1)
    normals = np.tile([0.0, 1.0], (n_elements, 1))
    widths = np.full(n_elements, pitch * 0.999)
    return positions, normals, widths


def plane_wave_amplitudes(positions_mm, angles_deg, frequency_mhz, c0, apodize=True):
    """Transmit amplitudes [n_angles, E]: exp(i k0 x sin(theta)) (exp(-i omega t)
    convention: a plane wave leaving the array at angle theta from +z)."""
    x = positions_mm[:, 0] * 1e-3
    x = x - x.mean()
    k0 = 2 * np.pi * frequency_mhz * 1e6 / c0
    th = np.deg2rad(np.asarray(angles_deg, dtype=np.float64))
    amp = np.exp(1j * k0 * np.sin(th)[:, None] * x[None, :])
    if apodize:
        e = len(x)
        amp = amp * (0.5 - 0.5 * np.cos(2 * np.pi * (np.arange(e) + 0.5) / e))[None, :]
    return amp
