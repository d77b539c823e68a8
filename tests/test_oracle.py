"""Pins the CPU oracle against the reference's golden vectors and closed forms.

The golden file was written by tests/golden/make_golden.py from the
reference's own theranostic Solver.  The physics checks mirror
tests/test_theranostic.py:72-170 of the reference.
"""
from pathlib import Path

import numpy as np
import pytest
from scipy import sparse
from scipy.special import hankel1

from oracle import helmholtz as H

GOLDEN = np.load(Path(__file__).parent / "golden" / "theranostic_small.npz")


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def golden_grid():
    axis = GOLDEN["axis"]
    return H.Grid(axis, axis, float(GOLDEN["pml_mm"]))


def golden_weights(grid):
    return H.source_weights(grid, GOLDEN["positions"], GOLDEN["normals"], GOLDEN["widths"])


def test_coefficients_match_golden():
    grid = golden_grid()
    k0, mass, edge = H.coefficients(grid, float(GOLDEN["frequency_mhz"]))
    assert abs(k0 - GOLDEN["k0"]) / GOLDEN["k0"] < 1e-15
    np.testing.assert_allclose(mass, GOLDEN["cell_mass"], rtol=1e-15)
    np.testing.assert_allclose(edge, GOLDEN["edge_stretch"], rtol=1e-15)


def test_matrix_matches_golden():
    grid = golden_grid()
    a = H.assemble(grid, float(GOLDEN["frequency_mhz"]), GOLDEN["m_kappa"], GOLDEN["m_eta"])
    g = sparse.csc_matrix((GOLDEN["matrix_val"], (GOLDEN["matrix_row"], GOLDEN["matrix_col"])),
                          shape=a.shape)
    assert abs(a - g).max() <= 1e-12 * abs(g).max()


def test_matrix_without_pml_matches_golden():
    ax = np.linspace(0.0, 4.0, 9)
    az = np.linspace(0.0, 3.0, 7)
    grid = H.Grid(ax, az, 0.0)
    a = H.assemble(grid, 0.4, GOLDEN["nopml_mk"], GOLDEN["nopml_me"], c0=1500.0)
    g = sparse.csc_matrix((GOLDEN["nopml_val"], (GOLDEN["nopml_row"], GOLDEN["nopml_col"])), shape=a.shape)
    assert abs(a - g).max() <= 1e-12 * abs(g).max()


def test_matrix_free_apply_equals_matrix():
    """apply() is the oracle's matrix-free form; it must equal assemble() @ u."""
    grid = golden_grid()
    f = float(GOLDEN["frequency_mhz"])
    a = H.assemble(grid, f, GOLDEN["m_kappa"], GOLDEN["m_eta"])
    rng = np.random.default_rng(0)
    u = rng.normal(size=(3,) + grid.shape) + 1j * rng.normal(size=(3,) + grid.shape)
    out = H.apply(grid, f, u, 1 + GOLDEN["m_eta"], 1 + GOLDEN["m_kappa"])
    for r in range(3):
        assert rel(out[r].ravel(), a @ u[r].ravel()) < 1e-14
    # the same on the pml-free non-square grid (contrast on the ghost edges)
    grid0 = H.Grid(np.linspace(0.0, 4.0, 9), np.linspace(0.0, 3.0, 7), 0.0)
    a0 = H.assemble(grid0, 0.4, GOLDEN["nopml_mk"], GOLDEN["nopml_me"], c0=1500.0)
    u0 = rng.normal(size=grid0.shape) + 1j * rng.normal(size=grid0.shape)
    out0 = H.apply(grid0, 0.4, u0, 1 + GOLDEN["nopml_me"], 1 + GOLDEN["nopml_mk"], c0=1500.0)
    assert rel(out0.ravel(), a0 @ u0.ravel()) < 1e-14


def test_source_weights_match_golden():
    grid = golden_grid()
    w = golden_weights(grid)
    g = sparse.csc_matrix((GOLDEN["weights_val"], (GOLDEN["weights_row"], GOLDEN["weights_col"])), shape=w.shape)
    assert abs(w - g).max() < 1e-15
    np.testing.assert_allclose(np.asarray(w.sum(axis=0)).ravel(), 1, atol=1e-14)
    assert w.data.min() >= 0


def test_solve_matches_golden_fields():
    grid = golden_grid()
    w = golden_weights(grid)
    fields, diag = H.solve(grid, float(GOLDEN["frequency_mhz"]), w, GOLDEN["m_kappa"], GOLDEN["m_eta"])
    assert diag["relative_linear_residual"] < 1e-12
    assert rel(fields, GOLDEN["fields"]) < 1e-10
    assert rel(H.sample(fields, w), GOLDEN["data"]) < 1e-10


def test_reciprocal_not_hermitian():
    grid = golden_grid()
    w = golden_weights(grid)
    fields, _ = H.solve(grid, float(GOLDEN["frequency_mhz"]), w, GOLDEN["m_kappa"], GOLDEN["m_eta"])
    d = H.sample(fields, w)
    assert rel(d, d.T) < 1e-12
    assert np.linalg.norm(d - d.conj().T) > 1e-3
    a = H.assemble(grid, float(GOLDEN["frequency_mhz"]), GOLDEN["m_kappa"], GOLDEN["m_eta"])
    assert abs(a - a.T).max() == 0


def test_outgoing_green_function_and_grid_refinement():
    """Closed form (i/4) H0(k r), second-order refinement (reference test :141-158)."""
    f = 0.15
    positions = np.array([[-3.2, 0.0], [3.2, 0.0], [0.0, 4.8]])
    dist = np.linalg.norm(positions[:, None] - positions[None, :], axis=2) * 1e-3
    off = ~np.eye(3, dtype=bool)
    k0 = 2 * np.pi * f * 1e6 / 1540.0
    exact = 0.25j * hankel1(0, k0 * dist[off])
    errors = []
    for intervals in (40, 80):
        axis = np.linspace(-16.0, 16.0, intervals + 1)
        grid = H.Grid(axis, axis, 6.0)
        w = H.source_weights(grid, positions, np.tile([0.0, 1.0], (3, 1)), np.zeros(3))
        fields, _ = H.solve(grid, f, w)
        measured = H.sample(fields, w)
        if intervals == 40:
            assert rel(measured, GOLDEN["point_data"]) < 1e-10
        errors.append(rel(measured[off], exact))
    assert errors[1] < 0.02
    assert errors[1] < 0.4 * errors[0]


def test_contrast_validation():
    grid = golden_grid()
    bad = np.zeros(grid.shape)
    bad[20, 20] = -1.0
    with pytest.raises(ValueError):
        H.assemble(grid, 0.15, m_kappa=bad)
    bad[:] = 0
    bad[0, 0] = 0.01
    with pytest.raises(ValueError):
        H.assemble(grid, 0.15, m_eta=bad)


def test_attenuation_plane_wave_decay():
    """Closed form for the viscoacoustic extension: in a homogeneous medium
    the discrete 1-D operator u'' + k0^2 kappa_c u has the outgoing solution
    exp(i k z) with 2(cos(k h) - 1)/h^2 + k0^2 kappa_c = 0, and Im k tends
    to alpha as h -> 0.  The oracle's 2-D apply on a field uniform in x
    reduces to the same 1-D operator on the middle column."""
    f = 1.0
    h = 0.02
    nz = 1201
    z = h * np.arange(nz)
    omega = 2 * np.pi * f * 1e6
    alpha = 300.0  # Np/m
    hm = h * 1e-3
    k0 = omega / 1540.0
    kf = H.attenuation_factor(1540.0, alpha, omega)
    main = np.full(nz, -2.0 / hm ** 2 + k0 ** 2 * kf)
    off = np.ones(nz - 1) / hm ** 2
    a = sparse.diags([off, main, off], [-1, 0, 1], format="csc")
    rhs = np.zeros(nz, complex)
    rhs[100] = -1.0 / hm
    u = sparse.linalg.spsolve(a, rhs)
    k_discrete = np.arccos(1 - k0 ** 2 * kf * hm ** 2 / 2) / hm
    rate = -np.log(abs(u[500]) / abs(u[300])) / (200 * hm)
    assert abs(rate - k_discrete.imag) / alpha < 1e-3
    assert abs(k_discrete.imag - alpha) / alpha < 0.01
    phase = np.angle(u[301] / u[300]) / hm
    assert abs(phase - k_discrete.real) / k_discrete.real < 1e-3
    grid1 = H.Grid(np.array([0.0, h, 2 * h]), z, 0.0)
    uu = np.tile(u[:, None], (1, 3)).astype(complex)
    out = H.apply(grid1, f, uu, None, np.full(grid1.shape, kf))
    np.testing.assert_allclose(out[1:-1, 1], (a @ u)[1:-1], rtol=1e-9, atol=1e-9 * np.abs(a @ u).max())


def test_gmres_matches_direct_solve():
    grid = golden_grid()
    f = float(GOLDEN["frequency_mhz"])
    a = H.assemble(grid, f, GOLDEN["m_kappa"], GOLDEN["m_eta"])
    w = golden_weights(grid)
    rhs = -(w[:, 0].toarray().ravel() / grid.h_m ** 2).astype(complex)
    exact = sparse.linalg.spsolve(a, rhs)
    trace = []
    x, reason = H.gmres(lambda v: a @ v, lambda v: v, rhs, np.zeros_like(rhs),
                        lambda n, u, r: trace.append(r), restart=200, max_cycles=40, tolerance=1e-10)
    assert reason == "converged"
    assert rel(x, exact) < 1e-7
    assert trace[-1] < 1e-10


def test_gmres_reasons_and_exact_termination():
    rng = np.random.default_rng(1)
    n = 12
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 6 * np.eye(n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    # full Krylov space within one cycle -> exact (happy breakdown)
    x, reason = H.gmres(lambda v: a @ v, lambda v: v, b, np.zeros(n, complex), restart=n, max_cycles=2,
                        tolerance=1e-10)
    assert reason == "converged" and rel(a @ x, b) < 1e-10
    # one short cycle cannot converge -> iteration_budget
    _, reason = H.gmres(lambda v: a @ v, lambda v: v, b, np.zeros(n, complex), restart=2, max_cycles=1,
                        tolerance=1e-12)
    assert reason == "iteration_budget"
    # exact start -> converged with zero applications
    calls = []
    _, reason = H.gmres(lambda v: a @ v, lambda v: v, b, np.linalg.solve(a, b), lambda c, u, r: calls.append(c),
                        restart=4, max_cycles=3, tolerance=1e-8)
    assert reason == "converged" and calls == [0]


def test_node_weight_symbol_matches_brute_force_spectrum():
    """For constant coefficients and no PML the local symbol is exact: the
    worst |g| the weight rule predicts over all modes equals the spectral
    radius of the Jacobi iteration matrix I - w D^-1 M (periodic-free Dirichlet
    grid: eigenvalues sit inside the predicted range)."""
    from oracle import multigrid as M
    n = 15
    axis = np.arange(n) * 0.1
    grid = H.Grid(axis, axis, 0.0)
    f = 1.2
    lev = M.Level(grid, 0.1, n, n, f, 1540.0, 12.0)
    coef = lev.coef(np.ones((n, n)), np.ones((n, n), complex), 0.5)
    w = M.node_weights(coef, coarsest=True)
    wc = w[n // 2, n // 2]
    a = (coef[1] + coef[2])[n // 2, n // 2] / coef[0][n // 2, n // 2]
    b = (coef[3] + coef[4])[n // 2, n // 2] / coef[0][n // 2, n // 2]
    shifted = H.assemble(grid, f, kappa_factor=np.full(grid.size, 1 + 0.5j)).toarray()
    d = np.diag(shifted)
    it = np.eye(grid.size) - wc * shifted / d[:, None]
    rho = np.abs(np.linalg.eigvals(it)).max()
    t = np.cos(np.pi * np.arange(1, n + 1) / (n + 1))
    pred = np.abs(1 - wc * (1 + t[:, None] * a + t[None, :] * b)).max()
    assert abs(rho - pred) < 1e-9
    assert rho < 1.0
