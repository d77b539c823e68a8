"""GPU: the measured configurations at their stated sizes.

* Config 1 (the reference's CPU-direct case): 128^2 interior + 20-cell
  stretched PML (168^2 nodes), 1 MHz, dx = lambda_min/10, five seeded
  Gaussian inclusions, a 64-element line; the fields of every element port
  and of the 0-degree plane wave (the sum of the port columns), and the
  receiver traces, against the oracle's complex128 sparse LU.
* Config 3 (batched uneven solves): the lowest, a middle and the highest of
  the 64 frequencies in 4.0-8.9 MHz x 11 angles at their full grid sizes
  (interior 384^2 to 854^2 + 32 PML), 4 host threads on 4 streams with
  captured restart cycles; every solve converges to 1e-6 and the lowest
  frequency's fields match the oracle's sparse LU; failed captures are
  reported.
* Config 4 (adjoint gradient): misfit 0.5 ||P u - d||^2 on 256^2 + 32 PML at
  3 MHz, 11 angles, data synthesised from a seeded inclusion medium, the
  gradient at the uniform starting model w.r.t. log c and log rho against
  central finite differences (eps = 1e-4, and their Richardson extrapolation
  with eps/2, which removes the eps^2 term) along 20 random directions, all
  in complex128 solves.
"""
import numpy as np
import pytest
import torch
from scipy.ndimage import gaussian_filter

from oracle import helmholtz as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def test_config1_inclusion_medium_fields_and_traces():
    from nsrno_amd import media
    from nsrno_amd import theranostic as T
    w = media.inclusion_medium(seed=0, interior=128, pml=20, frequency_mhz=1.0)
    assert (w.nz, w.nx) == (168, 168)
    pos, nrm, wid = media.line_array(w, 64)
    grid = T.Grid(w.x_mm, w.z_mm, w.pml_mm)
    solver = T.Solver(grid, c0_m_s=w.c0)
    weights = solver.source_weights(pos, nrm, wid)
    mk = w.rho0 * w.c0 ** 2 / (w.rho * w.c ** 2) - 1
    me = w.rho0 / w.rho - 1
    cols = np.hstack((weights.toarray(), weights.toarray().sum(axis=1, keepdims=True)))  # 64 ports + 0 deg
    from scipy import sparse
    src = sparse.csc_matrix(cols)
    fields, diag = solver.solve(w.frequency_mhz, src, mk, me, return_diagnostics=True)
    traces = solver.sample(fields, weights)
    og = O.Grid(w.x_mm, w.z_mm, w.pml_mm)
    ref, odiag = O.solve(og, w.frequency_mhz, src, mk, me, c0=w.c0)
    assert fields.shape == ref.shape == (168 * 168, 65)
    err = np.linalg.norm(fields - ref, axis=0) / np.linalg.norm(ref, axis=0)
    print(f"config 1: field error max {err.max():.2e}, residual {diag['relative_linear_residual']:.1e}")
    assert err.max() < 1e-5
    ref_traces = O.sample(ref, weights)
    assert traces.shape == ref_traces.shape == (65, 64)
    assert rel(traces, ref_traces) < 1e-5


def test_config3_sweep_extremes_converge_and_match():
    from nsrno_amd.batch import run_sweep
    freqs = np.linspace(4.0, 8.9, 64)[[0, 31, 63]]
    angles = np.linspace(-15.0, 15.0, 11)
    cases, wall = run_sweep(freqs, angles, tol=1e-6, restart=3, workers=4, graph=True, keep_fields=(0,))
    print(f"config 3 subset: {[(c.frequency_mhz, c.interior, max(c.iterations), c.graph_fallbacks) for c in cases]}, "
          f"{wall:.2f} s")
    assert [c.interior for c in cases] == [384, 615, 854]
    for c in cases:
        assert c.reasons == ("converged",) and c.relative_residual < 1e-6
    low = cases[0]
    work, weights, amp = low.setup["work"], low.setup["weights"], low.setup["amplitudes"]
    og = O.Grid(work.x_mm, work.z_mm, work.pml_mm)
    rhs = -(weights @ amp.T) / og.h_m ** 2
    ref, _ = O.solve(og, work.frequency_mhz, None, work.rho0 * work.c0 ** 2 / (work.rho * work.c ** 2) - 1,
                     work.rho0 / work.rho - 1, rhs=rhs, c0=work.c0, check_pml=False)
    got = low.fields.reshape(len(angles), -1).T
    err = np.linalg.norm(got - ref, axis=0) / np.linalg.norm(ref, axis=0)
    print(f"config 3 lowest frequency field error vs sparse LU: max {err.max():.2e}")
    assert err.max() < 1e-5
    assert rel(low.traces, (weights.T @ ref).T) < 1e-5


def test_config4_adjoint_gradient_twenty_directions():
    from nsrno_amd import media
    from nsrno_amd.adjoint import Problem, implicit_receivers
    truth = media.inclusion_medium(seed=3, interior=256, pml=32, frequency_mhz=3.0)
    assert (truth.nz, truth.nx) == (320, 320)
    pos, nrm, wid = media.line_array(truth, 192)
    W = O.source_weights(O.Grid(truth.x_mm, truth.z_mm, truth.pml_mm), pos, nrm, wid)
    amp = media.plane_wave_amplitudes(pos, np.linspace(-15.0, 15.0, 11), 3.0, truth.c0)
    pr = Problem(truth.nz, truth.nx, truth.h_mm, 3.0, truth.pml_mm, W, amp, c0=truth.c0, tolerance=1e-12,
                 restart=20)
    with torch.no_grad():
        d_obs = implicit_receivers(torch.tensor(np.log(truth.c), device="cuda"),
                                   torch.tensor(np.log(truth.rho), device="cuda"), pr)
    lc0 = torch.full((truth.nz, truth.nx), float(np.log(truth.c0)), dtype=torch.float64, device="cuda")
    lr0 = torch.full((truth.nz, truth.nx), float(np.log(truth.rho0)), dtype=torch.float64, device="cuda")
    lc = lc0.clone().requires_grad_(True)
    lr = lr0.clone().requires_grad_(True)
    loss = 0.5 * (implicit_receivers(lc, lr, pr) - d_obs).abs().square().sum()
    loss.backward()
    gc, gr = lc.grad.cpu().numpy(), lr.grad.cpu().numpy()

    def L(dc, dr):
        with torch.no_grad():
            return float(0.5 * (implicit_receivers(lc0 + dc, lr0 + dr, pr) - d_obs).abs().square().sum())

    rng = np.random.default_rng(20)
    eps = 1e-4
    inner = np.zeros((truth.nz, truth.nx))
    inner[truth.pml_cells:-truth.pml_cells, truth.pml_cells:-truth.pml_cells] = 1.0  # contrast vanishes in the PML
    worst, worst_r = 0.0, 0.0
    for k in range(20):
        vc = gaussian_filter(rng.normal(size=inner.shape), 6) * inner
        vr = gaussian_filter(rng.normal(size=inner.shape), 6) * inner
        vc, vr = vc / np.abs(vc).max(), vr / np.abs(vr).max()
        tc, tr = torch.tensor(vc, device="cuda"), torch.tensor(vr, device="cuda")

        def central(e):
            return (L(e * tc, e * tr) - L(-e * tc, -e * tr)) / (2 * e)

        fd, fd2 = central(eps), central(eps / 2)
        # the central difference carries an eps^2 term (the data's phase moves by
        # ~k0 * 13 mm per unit log c): Richardson's (4 D(eps/2) - D(eps)) / 3 removes it
        rich = (4 * fd2 - fd) / 3
        an = float(np.sum(gc * vc) + np.sum(gr * vr))
        worst = max(worst, abs(fd - an) / abs(an))
        worst_r = max(worst_r, abs(rich - an) / abs(an))
        assert abs(fd - an) <= 1.5e-2 * abs(an), (k, fd, an)  # eps^2 truncation: 0.2-0.7 % at eps = 1e-4
        assert abs(rich - an) <= 2e-5 * abs(an), (k, rich, an)
    print(f"config 4: 20 directions, worst relative mismatch: central eps=1e-4 {worst:.2e}, "
          f"Richardson {worst_r:.2e}")
