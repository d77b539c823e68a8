"""GPU: implicit-adjoint material gradient against central finite differences
(the check of nsrno/dense_wave.py's ImplicitChannels, scripts/profile_dense_solve.py:19-27,
here in complex128 with tight solves) and against the CPU oracle's direct solve."""
import numpy as np
import pytest
import torch

from oracle import helmholtz as O

pytestmark = pytest.mark.gpu


def make_problem(tol=1e-11):
    from nsrno_amd import media
    from nsrno_amd.adjoint import Problem
    w = media.smooth_random_medium(seed=4, interior=64, pml=16, frequency_mhz=1.0, sigma_cells=4.0)
    pos, nrm, wid = media.line_array(w, 24)
    W = O.source_weights(O.Grid(w.x_mm, w.z_mm, w.pml_mm), pos, nrm, wid)
    amp = media.plane_wave_amplitudes(pos, [-10.0, 0.0, 12.0], 1.0, w.c0)
    pr = Problem(w.nz, w.nx, w.h_mm, 1.0, w.pml_mm, W, amp, alpha_np_m=w.alpha_np_m, c0=w.c0, tolerance=tol)
    return w, pr, W, amp


def test_gradient_matches_finite_differences():
    from nsrno_amd.adjoint import implicit_receivers
    w, pr, W, amp = make_problem()
    rng = np.random.default_rng(0)
    lc0 = torch.tensor(np.log(w.c), device="cuda")
    lr0 = torch.tensor(np.log(w.rho), device="cuda")
    with torch.no_grad():
        d_obs = implicit_receivers(lc0 + 0.01, lr0 - 0.02, pr)  # observed data from a shifted medium
    lc = lc0.clone().requires_grad_(True)
    lr = lr0.clone().requires_grad_(True)
    d = implicit_receivers(lc, lr, pr)
    loss = 0.5 * (d - d_obs).abs().square().sum()
    loss.backward()
    gc, gr = lc.grad.cpu().numpy(), lr.grad.cpu().numpy()

    def L(dc, dr):
        with torch.no_grad():
            dd = implicit_receivers(lc0 + dc, lr0 + dr, pr)
            return float(0.5 * (dd - d_obs).abs().square().sum())

    eps = 1e-4
    for k in range(4):
        from scipy.ndimage import gaussian_filter
        vc = gaussian_filter(rng.normal(size=w.c.shape), 3)
        vr = gaussian_filter(rng.normal(size=w.c.shape), 3) * (k % 2)
        vc, vr = vc / np.abs(vc).max(), vr / max(np.abs(vr).max(), 1e-30)
        tc, tr = torch.tensor(vc, device="cuda"), torch.tensor(vr, device="cuda")
        fd = (L(eps * tc, eps * tr) - L(-eps * tc, -eps * tr)) / (2 * eps)
        an = float(np.sum(gc * vc) + np.sum(gr * vr))
        assert abs(fd - an) <= 2e-5 * abs(fd) + 1e-12 * float(loss), (k, fd, an)


def test_forward_traces_match_direct_solve():
    from nsrno_amd.adjoint import implicit_receivers
    w, pr, W, amp = make_problem(tol=1e-11)
    with torch.no_grad():
        d = implicit_receivers(torch.tensor(np.log(w.c), device="cuda"), torch.tensor(np.log(w.rho), device="cuda"),
                               pr).cpu().numpy()
    grid = O.Grid(w.x_mm, w.z_mm, w.pml_mm)
    omega = 2 * np.pi * 1e6
    rhs = -(W @ amp.T) / grid.h_m ** 2
    x, _ = O.solve(grid, 1.0, None, w.rho0 * w.c0 ** 2 / (w.rho * w.c ** 2) - 1, w.rho0 / w.rho - 1,
                   kappa_factor=O.attenuation_factor(w.c, w.alpha_np_m, omega), rhs=rhs, c0=w.c0, check_pml=False)
    ref = (W.T @ x).T
    assert np.linalg.norm(d - ref) / np.linalg.norm(ref) < 1e-9
