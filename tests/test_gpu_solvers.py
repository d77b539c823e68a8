"""GPU: multigrid preconditioner and BiCGSTAB against the CPU oracle."""
from pathlib import Path

import numpy as np
import pytest
import torch
from scipy.sparse.linalg import splu

from oracle import helmholtz as O

pytestmark = pytest.mark.gpu
GOLDEN = np.load(Path(__file__).parent / "golden" / "theranostic_small.npz")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def setup(nrhs=4, levels=3):
    from nsrno_amd.helmholtz import HelmholtzPlan, Ports
    axis = GOLDEN["axis"]
    grid = O.Grid(axis, axis, 6.0)
    f = float(GOLDEN["frequency_mhz"])
    plan = HelmholtzPlan(grid.shape[0], grid.shape[1], grid.h_mm, f, 6.0, x0_mm=axis[0], z0_mm=axis[0],
                         nrhs=nrhs, restart=30)
    plan.set_medium(1 + GOLDEN["m_eta"], (1 + GOLDEN["m_kappa"]).astype(complex))
    plan.setup_multigrid(levels=levels, shift=0.5, pre=2, post=2, coarse_sweeps=30)
    w = O.source_weights(grid, GOLDEN["positions"], GOLDEN["normals"], GOLDEN["widths"])
    rhs = plan.inject(Ports(w), np.eye(4)[:nrhs])
    return grid, f, plan, w, rhs


def test_vcycle_iteration_solves_shifted_operator():
    """Richardson iteration with the V-cycle converges to the direct solution
    of the shifted operator A(kappa_c (1 + 0.5 i)) (oracle splu)."""
    grid, f, plan, w, rhs = setup(nrhs=1)
    shifted = O.assemble(grid, f, GOLDEN["m_kappa"], GOLDEN["m_eta"], kappa_factor=np.full(grid.size, 1 + 0.5j))
    b = rhs.cpu().numpy().ravel()
    exact = splu(shifted.tocsc()).solve(b)
    x = np.zeros_like(b)
    scale = np.linalg.norm(b)
    for _ in range(60):
        r = (b - shifted @ x) / scale
        z = plan.precondition(torch.from_numpy(r.astype(np.complex64).reshape(1, *grid.shape)).cuda())
        x = x + scale * z.cpu().numpy().ravel().astype(complex)
    assert rel(x, exact) < 1e-4
    # linear and deterministic
    v = torch.randn(2, *grid.shape, dtype=torch.complex64, device="cuda")
    z1 = plan.precondition(v)
    z2 = plan.precondition(v)
    assert torch.equal(z1, z2)
    z3 = plan.precondition((2 * v).contiguous())
    assert rel(z3.cpu().numpy(), 2 * z1.cpu().numpy()) < 1e-5


@pytest.mark.parametrize("method,precond", [("gmres", "mg"), ("bicgstab", "mg")])
def test_preconditioned_and_bicgstab_solves_match_direct(method, precond):
    grid, f, plan, w, rhs = setup()
    x, infos = plan.solve(rhs, tolerance=1e-10, precond=precond, method=method, max_iterations=20000)
    assert all(i.reason == "converged" for i in infos), infos
    fields = x.cpu().numpy().reshape(4, -1).T
    assert rel(fields, GOLDEN["fields"]) < 1e-7
    a = O.assemble(grid, f, GOLDEN["m_kappa"], GOLDEN["m_eta"])
    b = rhs.cpu().numpy().reshape(4, -1).T
    assert rel(a @ fields, b) < 1e-10


def test_multigrid_reduces_iterations():
    grid, f, plan, w, rhs = setup()
    _, plain = plan.solve(rhs, tolerance=1e-8, precond="none", max_iterations=20000)
    _, mg = plan.solve(rhs, tolerance=1e-8, precond="mg", max_iterations=20000)
    assert all(i.reason == "converged" for i in plain + mg)
    assert max(i.iterations for i in mg) < 0.5 * max(i.iterations for i in plain)


def test_adjoint_mg_solve():
    grid, f, plan, w, rhs = setup(nrhs=2)
    x, infos = plan.solve(rhs, tolerance=1e-10, precond="mg", adjoint=True, max_iterations=20000)
    a = O.assemble(grid, f, GOLDEN["m_kappa"], GOLDEN["m_eta"])
    xs = x.cpu().numpy().reshape(2, -1).T
    b = rhs.cpu().numpy().reshape(2, -1).T
    assert all(i.reason == "converged" for i in infos)
    assert rel(a.conj().T @ xs, b) < 1e-10


def test_bicgstab_recurrence_matches_oracle():
    """Six unpreconditioned BiCGSTAB steps: the device recurrence equals the
    oracle restatement of dense_wave.py:118-148 (FP32 vectors vs FP64).  A
    PML-free, strongly damped grid keeps the recurrence well conditioned so
    FP32 and FP64 iterates stay close (BiCGSTAB is not a minimising method:
    on ill-conditioned grids rounding alone separates the two precisions)."""
    from nsrno_amd.helmholtz import HelmholtzPlan
    axis = np.linspace(0.0, 4.0, 33)
    grid = O.Grid(axis, axis, 0.0)
    f = 0.4
    rng = np.random.default_rng(11)
    mk = 0.1 * rng.uniform(-1, 1, grid.shape)
    me = 0.1 * rng.uniform(-1, 1, grid.shape)
    plan = HelmholtzPlan(33, 33, axis[1] - axis[0], f, 0.0, c0_m_s=1500.0, nrhs=2, restart=8)
    plan.set_medium(1 + me, (1 + mk) * (1 + 2j))
    a = O.assemble(grid, f, mk, me, kappa_factor=np.full(grid.size, 1 + 2j), c0=1500.0, check_pml=False)
    b = rng.normal(size=(2,) + grid.shape) + 1j * rng.normal(size=(2,) + grid.shape)
    x, infos = plan.solve(torch.from_numpy(b).cuda(), tolerance=1e-14, method="bicgstab", precond="none",
                          max_iterations=6)
    for r in range(2):
        ref, info = O.bicgstab(lambda u: a @ u, lambda u: u, b[r].ravel(), np.zeros(grid.size, complex),
                               tolerance=1e-14, max_iterations=6)
        assert info["iterations"] == 6 and info["reason"] == "iteration_budget"
        assert rel(x[r].cpu().numpy().ravel(), ref) < 1e-4


@pytest.mark.parametrize("pre,post", [(2, 1), (2, 2), (3, 2)])
def test_vcycle_matches_oracle_restatement(pre, post):
    """One V-cycle (fixed real weight so no arg-min ties) equals the CPU
    restatement in oracle/multigrid.py; (2, 2) runs the fused level kernels
    (smoothing + residual + restriction, prolongation + two sweeps)."""
    from oracle import multigrid as M
    grid, f, plan, w, rhs = setup(nrhs=3, levels=3)
    plan.setup_multigrid(levels=3, shift=0.5, pre=pre, post=post, coarse_sweeps=7, omega=0.6)
    mg = M.Multigrid(grid, f, 1 + GOLDEN["m_eta"], (1 + GOLDEN["m_kappa"]).astype(complex), levels=3, shift=0.5,
                     pre=pre, post=post, coarse=7, omega=0.6)
    v = np.random.default_rng(3).normal(size=(3,) + grid.shape) + 1j * np.random.default_rng(4).normal(
        size=(3,) + grid.shape)
    z = plan.precondition(torch.from_numpy(v.astype(np.complex64)).cuda()).cpu().numpy()
    assert rel(z, mg.vcycle(v)) < 2e-5
    if (pre, post) != (2, 1):
        return
    # per-node weights: the oracle's V-cycle with its own weights converges
    mg2 = M.Multigrid(grid, f, 1 + GOLDEN["m_eta"], (1 + GOLDEN["m_kappa"]).astype(complex), levels=3, shift=0.5,
                      pre=2, post=2, coarse=30)
    shifted = O.assemble(grid, f, GOLDEN["m_kappa"], GOLDEN["m_eta"], kappa_factor=np.full(grid.size, 1 + 0.5j))
    b = rhs[0].cpu().numpy().ravel()
    exact = splu(shifted.tocsc()).solve(b)
    xx = np.zeros_like(b)
    for _ in range(30):
        xx = xx + mg2.vcycle((b - shifted @ xx).reshape(1, *grid.shape)).ravel()
    assert rel(xx, exact) < 1e-3


def test_graph_replayed_cycles_match_eager():
    """Restart cycles replayed as a captured HIP graph give bit-identical
    solutions to the eagerly launched cycles, replay after replay, with other
    work and synchronisations in between."""
    from nsrno_amd.helmholtz import HelmholtzPlan
    grid = O.grid_with_pml(120, 100, 12, 0.05)
    nz, nx = grid.shape
    rng = np.random.default_rng(3)
    eta = 1 + 0.1 * rng.random((nz, nx))
    kap = (1 + 0.2 * rng.random((nz, nx))).astype(complex)
    plan = HelmholtzPlan(nz, nx, grid.h_mm, 2.0, grid.pml_mm, x0_mm=grid.x_mm[0], z0_mm=grid.z_mm[0], nrhs=3,
                         restart=8)
    plan.set_medium(eta, kap)
    plan.setup_multigrid(levels=3, shift=0.5, pre=2, post=2, coarse_sweeps=10)
    b = torch.from_numpy(rng.normal(size=(3, nz, nx)) + 1j * rng.normal(size=(3, nz, nx))).cuda()
    x_e, i_e = plan.solve(b, tolerance=1e-9, restart=8, max_iterations=800, precond="mg")
    assert all(i.reason == "converged" for i in i_e) and i_e[0].cycles > 3
    for rep in range(3):
        noise = plan.apply(torch.randn(3, nz, nx, dtype=torch.complex64, device="cuda"))  # eager work between
        torch.cuda.synchronize()
        x_g, i_g = plan.solve(b, tolerance=1e-9, restart=8, max_iterations=800, precond="mg", graph=True)
        torch.cuda.synchronize()
        assert [i.iterations for i in i_g] == [i.iterations for i in i_e], rep
        assert torch.equal(x_g, x_e), rep
        del noise
