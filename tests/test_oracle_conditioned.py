"""Pins of the conditioned split-PML oracle (oracle/conditioned.py).

Golden vectors: tests/golden/conditioned_small.npz, written by
tests/golden/make_golden_c.py from nsrno/conditioned_wave.py:12-92 and
nsrno/dense_wave.py:59-82 in float64.  Beyond the goldens: the real-space
stencil equals the Fourier form, the analytic adjoint satisfies
<y, A x> = <A^H y, x> (the reference's own check, scripts/check_fast_fft.py:32-36,
bound 1e-10), the operator restricted to the native domain does not depend on
the FFT padding (scripts/check_fast_fft.py:18-24), and in a uniform lossless
interior it reduces to the closed form 1 - |D(q)|^2 / k0^2 on a plane wave.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import conditioned as oc

GOLD = Path(__file__).parent / "golden" / "conditioned_small.npz"


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


@pytest.fixture(scope="module")
def system(gold):
    spec = json.loads(str(gold["spec_json"]))
    g = oc.Geometry(spec)
    medium = np.stack((gold["c"], gold["rho"]))
    return oc.System(g, medium, gold["alpha"], float(gold["f"]))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_geometry_matches_golden(gold, system):
    g = system.g
    assert (g.H, g.W, g.pz, g.px, g.face) == (int(gold["H"]), int(gold["W"]), int(gold["pz"]), int(gold["px"]),
                                               int(gold["face"]))
    np.testing.assert_array_equal(g.rx_cols, gold["rx_cols"])
    np.testing.assert_array_equal(g.mask, gold["mask"])
    for name in ("arx", "arz", "avx", "avz"):
        np.testing.assert_allclose(getattr(g, name), gold[name], rtol=0, atol=1e-15, err_msg=name)


def test_coefficients_match_golden(gold, system):
    s = system
    np.testing.assert_array_equal(s.c, gold["sc"])
    np.testing.assert_array_equal(s.rho, gold["srho"])
    np.testing.assert_array_equal(s.alpha, gold["salpha"])
    assert abs(s.D - complex(gold["D"])) < 1e-15 and abs(s.k0 / float(gold["k0"]) - 1) < 1e-15
    for name in ("ax", "az", "mass", "svx", "svz", "gvx", "gvz", "mx", "mz"):
        assert rel(getattr(s, name), gold[name]) < 1e-14, name


def test_operator_matches_golden(gold, system):
    p = gold["p"]
    assert rel(system.apply_fft(p), gold["call"]) < 1e-13
    assert rel(system.apply(p), gold["call"]) < 1e-13
    assert rel(system.projected(p), gold["proj"]) < 1e-13
    assert rel(system.adjoint(p), gold["adj"]) < 1e-13
    assert rel(system.adjoint_fft(p), gold["adj"]) < 1e-13


def test_source_and_fixed_preconditioner_match_golden(gold, system):
    assert rel(oc.source_rhs(system, gold["line"]), gold["src"]) < 1e-14
    m = system.fixed_multiplier()
    fixed = system.g.mask * np.fft.ifft2(m * np.fft.fft2(gold["r"]))
    assert rel(fixed, gold["fixed"]) < 1e-13


def test_adjoint_identity(system):
    rng = np.random.default_rng(3)
    g = system.g
    x = (rng.normal(size=(g.H, g.W)) + 1j * rng.normal(size=(g.H, g.W))) * g.mask
    y = (rng.normal(size=(g.H, g.W)) + 1j * rng.normal(size=(g.H, g.W))) * g.mask
    left = np.vdot(y, system.projected(x))
    right = np.vdot(system.adjoint(y), x)
    assert abs(left - right) / max(abs(left), abs(right)) < 1e-10
    # a planted mistake (conjugate of the mass term dropped) breaks the identity
    bad = system.adjoint(y) + (system.mass - np.conj(system.mass)) * g.mask * y
    assert abs(np.vdot(bad, x) - left) / abs(left) > 1e-6


def test_native_operator_independent_of_padding(gold, system):
    spec = json.loads(str(gold["spec_json"]))
    big = oc.Geometry({**spec, "fft_shape": [160, 192]})
    s2 = oc.System(big, np.stack((gold["c"], gold["rho"])), gold["alpha"], float(gold["f"]))
    g1, g2 = system.g, big
    rng = np.random.default_rng(5)
    native = rng.normal(size=(g1.h, g1.w)) + 1j * rng.normal(size=(g1.h, g1.w))

    def pad(g, a):
        out = np.zeros((g.H, g.W), complex)
        out[g.pz:g.pz + g.h, g.px:g.px + g.w] = a
        return out

    def crop(g, a):
        return a[g.pz:g.pz + g.h, g.px:g.px + g.w]

    a1 = crop(g1, system.projected(pad(g1, native)))
    a2 = crop(g2, s2.projected(pad(g2, native)))
    assert rel(a2, a1) < 1e-13


def test_uniform_interior_closed_form():
    """Uniform lossless water, no PML inside the domain: A e^{iqx} = (1 - |D(q)|^2/k0^2) e^{iqx}."""
    spec = {"medium_shape": [2, 64, 64], "dx_m": 50e-6, "x0_m": -.0016, "z0_m": -.0016, "native_dt_s": 5e-9,
            "pml": 1, "eta_per_unit_alpha": [.015, .025], "taus_s": [12e-9, 35e-9], "fft_shape": [64, 64]}
    g = oc.Geometry(spec)
    # pml = 1: every node of the (unpadded) box is interior-damped only at its edges; drop the edges by
    # evaluating the bulk formula on the full periodic box with the PML factors switched off
    s = oc.System(g, np.stack((np.full((64, 64), oc.C0), np.full((64, 64), oc.RHO0))), np.zeros((64, 64)), 5e6)
    s.svx = np.ones_like(s.svx)
    s.svz = np.ones_like(s.svz)
    s.gvx = np.ones_like(s.gvx)
    s.gvz = np.ones_like(s.gvz)
    s.mx = np.zeros_like(s.mx)
    s.mz = np.zeros_like(s.mz)
    g.arx = np.ones_like(g.arx)
    g.arz = np.ones_like(g.arz)
    kx, kz = 5, 3
    zz, xx = np.meshgrid(np.arange(64), np.arange(64), indexing="ij")
    wave = np.exp(2j * np.pi * (kx * xx + kz * zz) / 64)
    dx = oc.symbol(64, g.dx)[kx]
    dz = oc.symbol(64, g.dx)[kz]
    expected = (1 - (abs(dx) ** 2 + abs(dz) ** 2) / s.k0 ** 2) * wave
    assert rel(s.apply(wave), expected) < 1e-12
    assert s.mass[5, 5] == pytest.approx(1.0, abs=1e-15)


def test_probed_matrix_and_direct_solve(system):
    g = system.g
    A = oc.matrix(system)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(g.H, g.W)) + 1j * rng.normal(size=(g.H, g.W))
    assert rel((A @ x.ravel()).reshape(g.H, g.W), system.projected(x)) < 1e-13
    # the adjoint stencil is the conjugate transpose of the probed matrix
    assert rel((A.conj().T @ x.ravel()).reshape(g.H, g.W), system.adjoint(x)) < 1e-13
    b = rng.normal(size=(2, g.H, g.W)) + 1j * rng.normal(size=(2, g.H, g.W))
    u = oc.solve_direct(system, b)
    for r in range(2):
        assert rel(system.projected(u[r]), b[r]) < 1e-10
    lam = oc.solve_direct(system, b, adjoint=True)
    assert rel(system.adjoint(lam[0]), b[0]) < 1e-10
