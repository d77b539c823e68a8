"""Pins of the spectral-operator (S) oracle: golden vectors of the reference's
own WaveResidualOperator / solve_frequency (tests/golden/make_golden_s.py),
the symbol == stencil == matrix identities, symmetry, the discrete
dispersion relation of the incident field, and solver agreement."""
from pathlib import Path

import numpy as np
import pytest

from oracle import spectral as S

G = np.load(Path(__file__).parent / "golden" / "spectral_small.npz")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def golden_box():
    return S.Box(int(G["h"]), int(G["w"]), float(G["dx"]), float(G["x0"]), float(G["z0"]))


def test_box_geometry_matches_reference():
    box = golden_box()
    assert (box.H, box.W, box.pz, box.px, box.face) == (int(G["H"]), int(G["W"]), int(G["pz"]), int(G["px"]),
                                                        int(G["face"]))
    assert np.array_equal(box.rx_cols, G["rx_cols"])


def test_setup_matches_golden():
    box = golden_box()
    ax, az, kc, k0 = S.setup(box, G["c"], G["rho"], float(G["f"]))
    # the reference computes in float32
    assert rel(ax, G["ax"]) < 1e-6 and rel(az, G["az"]) < 1e-6
    assert rel(kc, G["kc"]) < 1e-6
    assert abs(k0 - float(G["k0"])) < 1e-9 * k0


def test_setup_attenuation_matches_golden():
    """Power-law and discrete-relaxation kappa_c on a medium where the tissue mask is active."""
    box = golden_box()
    f = float(G["f"])
    _, _, kc, _ = S.setup(box, G["tissue_c"], G["tissue_rho"], f)
    assert rel(kc, G["power_kc_tissue"]) < 1e-6
    assert np.abs((kc + 1j * box.sponge).imag).max() > 1e-4  # attenuation present
    _, _, kr, k0 = S.setup(box, G["tissue_c"], G["tissue_rho"], f, attenuation="relaxation",
                           eta_per_unit_alpha=G["relax_epa"], taus=G["relax_taus"])
    assert rel(kr, G["relax_kc"]) < 1e-6
    assert abs(k0 - float(G["relax_k0"])) < 1e-9 * k0


def test_apply_matches_golden_and_forms_agree():
    box = golden_box()
    ax, az, kc, k0 = S.setup(box, G["c"], G["rho"], float(G["f"]))
    u = G["u"].astype(np.complex128)
    lu_fft = S.apply_fft(u, ax, az, kc, k0, box.dx)
    lu_st = S.apply_stencil(u, ax, az, kc, k0, box.dx)
    assert rel(lu_fft, G["lu"]) < 2e-6  # complex64 FFTs in the reference
    assert rel(lu_st, lu_fft) < 1e-13   # symbol == 8-tap stencil


def test_stencil_matrix_columns_and_symmetry():
    rng = np.random.default_rng(3)
    H, W, dx, k0 = 24, 32, 1e-4, 7000.0
    ax = 1 + 0.2 * rng.random((H, W))
    az = 1 + 0.2 * rng.random((H, W))
    kc = 1 + 0.1 * rng.random((H, W)) - 0.3j * rng.random((H, W))
    A = S.matrix(ax, az, kc, k0, dx)
    for col in (0, 5, W - 1, 7 * W + 3, H * W - 1):  # brute-force columns incl. wrap-around
        e = np.zeros(H * W, complex)
        e[col] = 1
        assert rel(A @ e, S.apply_stencil(e.reshape(H, W), ax, az, kc, k0, dx).ravel()) < 1e-14
        assert rel(A @ e, S.apply_fft(e.reshape(H, W), ax, az, kc, k0, dx).ravel()) < 1e-12
    # complex symmetric; real coefficients give a real symmetric matrix
    assert abs(A - A.T).max() < 1e-9 * abs(A).max()
    Ar = S.matrix(ax, az, kc.real, k0, dx)
    assert abs(Ar.imag).max() == 0
    # the divergence is minus the adjoint of the gradient: sum of each row of the
    # Laplacian part is zero (constants are in its null space)
    lap = S.matrix(ax, az, np.zeros_like(kc), k0, dx)
    assert abs(lap @ np.ones(H * W)).max() < 1e-9 * abs(lap).max()


def test_incident_matches_golden_and_dispersion():
    box = golden_box()
    f = float(G["f"])
    theta = float(G["angles"][1]) * np.pi / 180
    inc, q, desired = S.incident(box, f, theta)
    # the reference evaluates coordinates in float32: phase noise ~1e-5
    assert rel(inc, G["inc"]) < 5e-5
    prop = desired > 0
    assert prop.any()
    value = 2 * sum(a * np.sin((j + 0.5) * q) for j, a in enumerate(S.FD))
    assert np.abs(value - desired)[prop & (desired < 2.5)].max() < 1e-10
    # at z = 0 the field is the aperture trace itself
    row = int(np.argmin(np.abs(box.z)))
    assert abs(box.z[row]) < 1e-12
    x = box.x
    ap = np.where(np.abs(x) <= S.APERTURE_M, 0.5 * (1 + np.cos(np.pi * x / S.APERTURE_M)), 0)
    assert rel(inc[row], ap.astype(complex)) < 1e-12


def test_source_factor_matches_golden():
    sf = S.source_factor(float(G["f"]), float(G["trefs"][2]))
    assert abs(sf - complex(G["sf"])) < 2e-6 * abs(sf)


def small_case():
    # 0.6 mm spacing so the 192 receivers (0.2 mm pitch) fit a 66-column medium
    h, w, dx = 16, 66, 6e-4
    box = S.Box(h, w, dx, -0.0191 - 2 * dx, -3 * dx)
    zz, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    blob = np.exp(-((zz - 8) ** 2 + (xx - 33) ** 2) / 20.0)
    return box, 1540 + 30 * blob, 1000 - 40 * blob


def test_direct_and_gmres_agree_small():
    box, c, rho = small_case()
    f = 0.6e6
    pd, ud, rr = S.solve_direct(box, c, rho, f, 0.05, 0.0)
    assert rr < 1e-10
    pg, ug, info = S.solve_gmres(box, c, rho, f, 0.05, 0.0, iterations=4, tolerance=1e-11)
    assert info["reason"] == "converged"
    assert rel(ug, ud) < 1e-8 and rel(pg, pd) < 1e-8
    # the minimal-residual warm start never increases the residual
    assert all(b <= a * (1 + 1e-12) for a, b in zip(info["warm"], info["warm"][1:]))


def test_reference_prediction_within_its_stagnation():
    """The reference's solve (complex64) stopped at rel. residual 8e-6 with
    reason iteration_budget; its prediction is within a few percent of the
    converged complex128 solution of the same system."""
    box = golden_box()
    theta = float(G["angles"][1]) * np.pi / 180
    pred, u, info = S.solve_gmres(box, G["c"], G["rho"], float(G["f"]), theta, float(G["trefs"][1]), iterations=4,
                                  tolerance=1e-9)
    assert info["reason"] == "converged"
    assert str(G["pred_reason"]) == "iteration_budget" and float(G["pred_residual"]) < 1e-5
    assert rel(pred, G["pred"]) < 0.03


def _learned_params():
    return {k[len("learned_"):]: G[k].astype(np.float64) for k in G.files
            if k.startswith("learned_context") or k in ("learned_spectral", "learned_log_shift")}


def test_forward_matches_reference_default_model():
    """wave_operator.forward of the default model (gain 1, no modulation): the
    warm-start-only prediction, loss and trace (golden, complex64 reference)."""
    box = S.Box(int(G["h"]), int(G["w"]), float(G["dx"]), float(G["x0"]), float(G["z0"]))
    theta = float(G["angles"][1]) * np.pi / 180
    pred, loss, trace = S.forward(box, np.float32(G["c"]), np.float32(G["rho"]), float(G["f"]), theta,
                                  float(G["trefs"][1]), iterations=4)
    # the reference runs in complex64: its right-hand side and first direction
    # carry ~1e-5 relative rounding, its iterate ~4e-4; the receiver row holds
    # values ~1e-4 of the field's maximum, where that is ~2 % of the prediction
    ax, az, kc, k0 = S.setup(box, np.float32(G["c"]), np.float32(G["rho"]), float(G["f"]))
    inc, _, _ = S.incident(box, float(G["f"]), theta)
    rhs = S.scattered_rhs(box, inc, ax, az, kc, k0)
    assert rel(rhs, G["fwd_rhs"]) < 5e-5
    assert rel(np.fft.ifft2(S.multiplier(box, k0) * np.fft.fft2(rhs)), G["fwd_dir0"]) < 5e-5
    assert rel(pred, G["fwd_pred"]) < 5e-2
    assert abs(loss - float(G["fwd_loss"])) < 2e-3 * float(G["fwd_loss"])
    assert np.allclose(trace, G["fwd_trace"], rtol=1e-3)


def test_learned_gain_multiplier_and_forward_match_reference():
    box = S.Box(int(G["h"]), int(G["w"]), float(G["dx"]), float(G["x0"]), float(G["z0"]))
    prm = _learned_params()
    f = float(G["f"])
    gain = S.context_gain(box, np.float32(G["c"]), np.float32(G["rho"]), f, prm)
    assert rel(gain, G["learned_gain"]) < 1e-5
    k0 = S.wavenumber(f)
    shift = float(np.clip(np.exp(prm["log_shift"]), 0.03, 1.0))
    mult = S.modulated_multiplier(box, k0, shift, prm["spectral"])
    assert rel(mult, G["learned_mult"]) < 1e-5
    theta = float(G["angles"][1]) * np.pi / 180
    pred, loss, trace = S.forward(box, np.float32(G["c"]), np.float32(G["rho"]), f, theta, float(G["trefs"][1]),
                                  iterations=4, gain=gain, mult=mult)
    assert rel(pred, G["learned_pred"]) < 5e-2  # complex64 reference, see the default-model test
    assert np.allclose(trace, G["learned_trace"], rtol=1e-3)
