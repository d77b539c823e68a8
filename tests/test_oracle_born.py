"""Born data / Jacobian of the oracle, pinned like tests/test_theranostic.py:90-139
of the reference (finite differences of the full solve, Born vs J, quadratic remainder)."""
import numpy as np
from scipy import sparse

from oracle import helmholtz as H


def fixture():
    axis = np.linspace(-16.0, 16.0, 41)
    grid = H.Grid(axis, axis, 6.0)
    positions = np.array([[-6.4, -3.2], [-1.7, -3.1], [3.2, -3.2], [6.3, 1.7]])
    normals = np.array([[0.0, 1.0], [0.2, 1.0], [-0.15, 1.0], [-0.7, 0.7]])
    w = H.source_weights(grid, positions, normals, np.array([0.8, 1.2, 0.6, 1.0]))
    xx, zz = np.meshgrid(axis, axis)

    def bump(x0, z0, a, b):
        return np.maximum(1 - ((xx - x0) / a) ** 2 - ((zz - z0) / b) ** 2, 0) ** 2

    basis = sparse.csr_matrix(np.column_stack([bump(-2.0, 1.0, 2.2, 1.7).ravel(), bump(2.0, 4.0, 1.6, 2.1).ravel(),
                                               bump(-1.5, 6.5, 2.4, 1.3).ravel()]))
    return grid, w, basis


def data(grid, w, basis, coef, f=0.15):
    mk, me = basis @ coef[:3], basis @ coef[3:]
    x, _ = H.solve(grid, f, w, mk, me)
    return H.sample(x, w)


def test_jacobian_matches_fd_and_born():
    grid, w, basis = fixture()
    f = 0.15
    u0, _ = H.solve(grid, f, w)
    tx, rx = np.array([3, 0]), np.array([2, 1, 0])
    jac = H.born_jacobian(grid, f, u0, tx, rx, basis)
    eps = 1e-4
    for cols in (slice(0, 3), slice(3, 6)):
        d = np.zeros(6)
        d[cols] = [0.7, -0.4, 0.55]
        fd = (data(grid, w, basis, eps * d) - data(grid, w, basis, -eps * d)) / (2 * eps)
        exp = fd[np.ix_(tx, rx)].ravel()
        assert np.linalg.norm(jac @ d - exp) / np.linalg.norm(exp) < 1e-6
    full = H.born_jacobian(grid, f, u0, np.arange(4), np.arange(4), basis)
    direction = np.array([0.7, -0.2, 0.4, -0.25, 0.55, 0.1])
    mk, me = basis @ (0.08 * direction[:3]), basis @ (0.08 * direction[3:])
    born = H.born_data(grid, f, w, mk, me, background_fields=u0)
    assert np.linalg.norm(born.ravel() - full @ (0.08 * direction)) / np.linalg.norm(born) < 1e-9
    ref = H.sample(u0, w)
    tangent = (full @ direction).reshape(4, 4)
    errs = [np.linalg.norm(data(grid, w, basis, a * direction) - ref - a * tangent) for a in (0.08, 0.04, 0.02)]
    for c, fi in zip(errs[:-1], errs[1:]):
        assert 0.15 < fi / c < 0.35
