"""GPU parity of the stencil, ports and Krylov solve against the CPU oracle."""
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import helmholtz as O

pytestmark = pytest.mark.gpu
GOLDEN = np.load(Path(__file__).parent / "golden" / "theranostic_small.npz")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def golden_setup():
    from nsrno_amd.helmholtz import HelmholtzPlan
    axis = GOLDEN["axis"]
    grid = O.Grid(axis, axis, 6.0)
    f = float(GOLDEN["frequency_mhz"])
    plan = HelmholtzPlan(grid.shape[0], grid.shape[1], grid.h_mm, f, 6.0, x0_mm=axis[0], z0_mm=axis[0],
                         nrhs=4, restart=30)
    plan.set_medium(1 + GOLDEN["m_eta"], (1 + GOLDEN["m_kappa"]).astype(complex))
    return grid, f, plan


def random_fields(shape, r, seed, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=(r,) + shape) + 1j * rng.normal(size=(r,) + shape)).astype(dtype)


def test_apply_fp64_matches_oracle():
    grid, f, plan = golden_setup()
    u = random_fields(grid.shape, 3, 0)
    ref = O.apply(grid, f, u, 1 + GOLDEN["m_eta"], 1 + GOLDEN["m_kappa"])
    out = plan.apply(torch.from_numpy(u).cuda()).cpu().numpy()
    assert rel(out, ref) < 1e-13


def test_apply_fp32_and_adjoint():
    grid, f, plan = golden_setup()
    u = random_fields(grid.shape, 2, 1)
    ref = O.apply(grid, f, u, 1 + GOLDEN["m_eta"], 1 + GOLDEN["m_kappa"])
    out = plan.apply(torch.from_numpy(u.astype(np.complex64)).cuda()).cpu().numpy()
    assert rel(out, ref) < 2e-6
    # adjoint identity <y, A x> = <A^H y, x>
    y = random_fields(grid.shape, 2, 2)
    ax = plan.apply(torch.from_numpy(u).cuda()).cpu().numpy()
    ahy = plan.apply(torch.from_numpy(y).cuda(), adjoint=True).cpu().numpy()
    lhs = np.vdot(y, ax)
    rhs = np.vdot(ahy, u)
    assert abs(lhs - rhs) / abs(lhs) < 1e-13
    # A^H = conj(A) for the complex-symmetric operator
    assert rel(ahy, np.conj(O.apply(grid, f, np.conj(y), 1 + GOLDEN["m_eta"], 1 + GOLDEN["m_kappa"]))) < 1e-13


def test_attenuation_medium_physical():
    from nsrno_amd.helmholtz import HelmholtzPlan
    rng = np.random.default_rng(5)
    ax = np.arange(48) * 0.1
    grid = O.Grid(ax, ax[:40], 0.8)
    f = 1.3
    c = 1500 + 40 * rng.uniform(-1, 1, grid.shape)
    rho = 1000 + 60 * rng.uniform(-1, 1, grid.shape)
    alpha = 30 * rng.uniform(0, 1, grid.shape)
    plan = HelmholtzPlan(grid.shape[0], grid.shape[1], 0.1, f, 0.8, c0_m_s=1500.0)
    plan.set_medium_physical(c, rho, alpha, rho0=1000.0)
    omega = 2 * np.pi * f * 1e6
    kap = 1000 * 1500 ** 2 / (rho * c ** 2) * O.attenuation_factor(c, alpha, omega)
    u = random_fields(grid.shape, 1, 3)
    ref = O.apply(grid, f, u, 1000 / rho, kap, c0=1500.0)
    out = plan.apply(torch.from_numpy(u).cuda()).cpu().numpy()
    assert rel(out, ref) < 1e-13
    # and the matrix form of the oracle agrees on the same coefficients
    a = O.assemble(grid, f, 1000 * 1500 ** 2 / (rho * c ** 2) - 1, 1000 / rho - 1,
                   kappa_factor=O.attenuation_factor(c, alpha, omega), c0=1500.0, check_pml=False)
    assert rel(out[0].ravel(), a @ u[0].ravel()) < 1e-13


def test_inject_and_sample_match_oracle():
    from nsrno_amd.helmholtz import Ports
    grid, f, plan = golden_setup()
    w = O.source_weights(grid, GOLDEN["positions"], GOLDEN["normals"], GOLDEN["widths"])
    ports = Ports(w)
    rng = np.random.default_rng(4)
    amp = rng.normal(size=(3, 4)) + 1j * rng.normal(size=(3, 4))
    rhs = plan.inject(ports, amp).cpu().numpy().reshape(3, -1)
    ref = -(w @ amp.T).T / grid.h_m ** 2
    assert rel(rhs, ref) < 1e-15
    u = random_fields(grid.shape, 3, 6)
    tr = plan.sample(ports, torch.from_numpy(u).cuda()).cpu().numpy()
    assert rel(tr, (w.T @ u.reshape(3, -1).T).T) < 1e-14


def test_gmres_solve_matches_direct_fields():
    from nsrno_amd.helmholtz import Ports
    grid, f, plan = golden_setup()
    w = O.source_weights(grid, GOLDEN["positions"], GOLDEN["normals"], GOLDEN["widths"])
    rhs = plan.inject(Ports(w), np.eye(4))
    trace = []
    x, infos = plan.solve(rhs, tolerance=1e-11, restart=30, max_iterations=20000,
                          callback=lambda n, r: trace.append((n, max(r))))
    fields = x.cpu().numpy().reshape(4, -1).T
    assert all(i.reason == "converged" for i in infos), infos
    assert max(i.relative_residual for i in infos) < 1e-11
    # the reference's own direct solve (golden) and the oracle's
    assert rel(fields, GOLDEN["fields"]) < 1e-8
    a = O.assemble(grid, f, GOLDEN["m_kappa"], GOLDEN["m_eta"])
    true = np.linalg.norm(a @ fields - (-(w.toarray()) / grid.h_m ** 2)) / np.linalg.norm(w.toarray() / grid.h_m ** 2)
    assert true < 2e-11
    assert trace[0][0] == 1 and trace[-1][1] < 1e-11


def test_solve_reasons_budget_and_exact_start():
    from nsrno_amd.helmholtz import Ports
    grid, f, plan = golden_setup()
    w = O.source_weights(grid, GOLDEN["positions"], GOLDEN["normals"], GOLDEN["widths"])
    rhs = plan.inject(Ports(w), np.eye(4)[:2])
    _, infos = plan.solve(rhs, tolerance=1e-12, restart=5, max_iterations=5)
    assert all(i.reason == "iteration_budget" and i.iterations == 5 for i in infos)
    exact = torch.from_numpy(np.ascontiguousarray(GOLDEN["fields"][:, :2].T.reshape(2, *grid.shape))).cuda()
    _, infos = plan.solve(rhs, x0=exact, tolerance=1e-9)
    assert all(i.reason == "converged" and i.iterations == 0 for i in infos)
    zero = torch.zeros_like(rhs)
    _, infos = plan.solve(zero, tolerance=1e-9)
    assert all(i.reason == "converged" for i in infos)


def test_adjoint_solve():
    grid, f, plan = golden_setup()
    b = random_fields(grid.shape, 1, 9)
    x, infos = plan.solve(torch.from_numpy(b).cuda(), tolerance=1e-11, adjoint=True, max_iterations=20000)
    a = O.assemble(grid, f, GOLDEN["m_kappa"], GOLDEN["m_eta"])
    xs = x.cpu().numpy().ravel()
    assert infos[0].reason == "converged"
    assert rel(a.conj().T @ xs, b.ravel()) < 2e-11


def test_theranostic_dropin_matches_golden():
    from nsrno_amd import theranostic as T
    axis = GOLDEN["axis"]
    solver = T.Solver(T.Grid(axis, axis, 6.0))
    w = solver.source_weights(GOLDEN["positions"], GOLDEN["normals"], GOLDEN["widths"])
    fields, diag = solver.solve(float(GOLDEN["frequency_mhz"]), w, GOLDEN["m_kappa"], GOLDEN["m_eta"],
                                return_diagnostics=True, source_chunk=3)
    assert diag["relative_linear_residual"] < 1e-10
    assert rel(fields, GOLDEN["fields"]) < 1e-8
    assert rel(solver.sample(fields, w), GOLDEN["data"]) < 1e-8
    d = solver.sample(fields, w)
    assert rel(d, d.T) < 1e-8
    op = solver.matrix(float(GOLDEN["frequency_mhz"]), GOLDEN["m_kappa"], GOLDEN["m_eta"])
    a = O.assemble(O.Grid(axis, axis, 6.0), float(GOLDEN["frequency_mhz"]), GOLDEN["m_kappa"], GOLDEN["m_eta"])
    v = np.random.default_rng(0).normal(size=a.shape[0]) + 0j
    assert rel(op @ v, a @ v) < 1e-13


def test_born_data_and_jacobian_match_oracle():
    from scipy import sparse as sp
    from nsrno_amd import theranostic as T
    axis = GOLDEN["axis"]
    grid = O.Grid(axis, axis, 6.0)
    solver = T.Solver(T.Grid(axis, axis, 6.0))
    f = float(GOLDEN["frequency_mhz"])
    w = solver.source_weights(GOLDEN["positions"], GOLDEN["normals"], GOLDEN["widths"])
    xx, zz = np.meshgrid(axis, axis)

    def bump(x0, z0, a, b):
        return np.maximum(1 - ((xx - x0) / a) ** 2 - ((zz - z0) / b) ** 2, 0) ** 2

    basis = sp.csr_matrix(np.column_stack([bump(-2.0, 1.0, 2.2, 1.7).ravel(), bump(2.0, 4.0, 1.6, 2.1).ravel()]))
    u0 = solver.solve(f, w)
    jac = solver.born_jacobian(f, u0, np.array([3, 0]), np.array([2, 1, 0]), basis, pair_chunk=2)
    ref = O.born_jacobian(grid, f, u0, np.array([3, 0]), np.array([2, 1, 0]), basis)
    assert rel(jac, ref) < 1e-12
    mk, me = basis @ np.array([0.05, -0.03]), basis @ np.array([-0.02, 0.04])
    bd = solver.born_data(f, w, mk, me, source_chunk=3)
    assert rel(bd, O.born_data(grid, f, w, mk, me)) < 1e-8
    bd2 = solver.born_data(f, w, mk, me, background_fields=u0)
    assert rel(bd2, bd) < 1e-8


@pytest.mark.parametrize("shape", [None, (8, 4, 0), (3, 1, 0), (2, 2, 1), (4, 4, 1), (8, 1, 1), (16, 4, 1),
                                   (4, 2, 2), (16, 1, 2), (2, 2, 3), (4, 4, 3), (8, 1, 4), (2, 4, 4),
                                   (2, 8, 1), (4, 8, 3), (2, 2, 5), (8, 8, 5), (2, 4, 6), (4, 8, 6)])
@pytest.mark.parametrize("nrhs", [1, 5])
@pytest.mark.parametrize("width", [170, 171])
def test_apply_fp32_streaming_shapes(nrhs, shape, width):
    """The FP32 stencil's streaming, block and paired block forms (rows per
    run, RHS per group, workgroup order) on grids that are no multiple of the
    strip (64 or 128 columns) or of the row runs, with a RHS count that leaves
    a partial group; the odd row length sends the paired form to its block
    fallback, the even one runs it."""
    from nsrno_amd.helmholtz import HelmholtzPlan
    grid = O.grid_with_pml(width, 131, 12, 0.05)  # 194 or 195 columns
    nz, nx = grid.shape
    rng = np.random.default_rng(5)
    eta = 1 + 0.1 * rng.random((nz, nx))
    kap = (1 + 0.2 * rng.random((nz, nx))) * (1 + 0.05j * rng.random((nz, nx)))
    plan = HelmholtzPlan(nz, nx, grid.h_mm, 2.0, grid.pml_mm, x0_mm=grid.x_mm[0], z0_mm=grid.z_mm[0], nrhs=nrhs,
                         restart=4)
    plan.set_medium(eta, kap)
    if shape is not None:
        plan.set_stencil_shape(*shape)
    u = random_fields(grid.shape, nrhs, 7)
    ref = O.apply(grid, 2.0, u, eta, kap)
    out = plan.apply(torch.from_numpy(u.astype(np.complex64)).cuda()).cpu().numpy()
    assert rel(out, ref) < 2e-6
    outa = plan.apply(torch.from_numpy(u.astype(np.complex64)).cuda(), adjoint=True).cpu().numpy()
    assert rel(outa, np.conj(O.apply(grid, 2.0, np.conj(u), eta, kap))) < 2e-6
