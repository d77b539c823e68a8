"""CPU pins of the oracle's BiCGSTAB (oracle/helmholtz.py bicgstab), the
restatement of nsrno/dense_wave.py:118-148.

* Golden vectors of the reference's own function (tests/golden/make_golden_b.py,
  complex128): iterates after 1, 2, 3, 4 and 7 steps (each formula of the
  recurrence enters them), a converged solve, the half-step exit, the
  breakdown on <r~, v> = 0, a solve past several 40-step refreshes and one
  that ends 'diverged' at a refresh, each with the reason, the iteration
  count, the operator-application count and the final residual.
* Mathematics: at convergence the iterate equals the direct solve (a dense
  system and the theranostic Helmholtz matrix with splu), and the reported
  residual is the true one.
"""
from pathlib import Path

import numpy as np
import pytest
from scipy.sparse.linalg import splu

from oracle import helmholtz as O

G = np.load(Path(__file__).parent / "golden" / "bicgstab_small.npz")


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def mat(a):
    return lambda v: (a @ v.ravel()).reshape(v.shape)


def info_of(name):
    reason, it, steps, res = G[f"{name}_info"]
    return reason, int(it), int(steps), float(res)


def check_info(info, name, res_rtol=1e-6):
    reason, it, steps, res = info_of(name)
    assert info["reason"] == reason
    assert info["iterations"] == it
    assert info["steps"] == steps
    assert info["method"] == "BiCGSTAB"
    assert abs(info["residual"] - res) <= res_rtol * res + 1e-15


@pytest.mark.parametrize("name,tol,maxit", [("conv", 1e-12, 800), ("it1", 1e-30, 1), ("it2", 1e-30, 2),
                                            ("it3", 1e-30, 3), ("it7", 1e-30, 7)])
def test_iterates_match_reference(name, tol, maxit):
    a, b, pd = G["a"], G["b"], G["pdiag"]
    x, info = O.bicgstab(mat(a), lambda v: (pd * v.ravel()).reshape(v.shape), b, np.zeros_like(b),
                         tolerance=tol, max_iterations=maxit)
    assert x.shape == b.shape
    assert rel(x, G[f"{name}_x"]) < 1e-10
    check_info(info, name, res_rtol=1e-4 if name == "conv" else 1e-8)


def test_initial_guess_matches_reference():
    x, info = O.bicgstab(mat(G["a"]), lambda v: v, G["b"], G["x0"], tolerance=1e-30, max_iterations=4)
    assert rel(x, G["x0run_x"]) < 1e-10
    check_info(info, "x0run", res_rtol=1e-8)


def test_half_step_exit():
    c = complex(G["half_c"])
    x, info = O.bicgstab(lambda v: c * v, lambda v: v, G["half_b"], np.zeros(5, complex), tolerance=1e-8)
    assert rel(x, G["half_x"]) < 1e-14
    assert rel(x, G["half_b"] / c) < 1e-14
    check_info(info, "half", res_rtol=1.0)


def test_breakdown_on_shadow_projection():
    x, info = O.bicgstab(mat(G["brk_a"]), lambda v: v, G["brk_b"], np.zeros(2, complex), tolerance=1e-8)
    assert np.all(x == 0) and np.all(G["brk_x"] == 0)
    check_info(info, "brk")


def test_refresh_branch_and_convergence():
    x, info = O.bicgstab(mat(G["long_a"]), lambda v: v, G["long_b"], np.zeros(G["long_b"].size, complex),
                         tolerance=1e-12, max_iterations=400)
    reason, it, steps, _ = info_of("long")
    assert (info["reason"], info["iterations"], info["steps"]) == (reason, it, steps) == ("converged", it, steps)
    assert it > 40 and steps == 2 * it + it // 40 + 2  # two applies per step, one per refresh, initial + final
    assert rel(x, G["long_x"]) < 1e-10
    assert rel(x, np.linalg.solve(G["long_a"], G["long_b"])) < 1e-10


def test_diverged_at_refresh():
    """An operator whose output grows by 1e9 from its 82nd call (the refresh
    after 40 steps) makes the recomputed residual exceed 1e5: 'diverged'."""
    calls = [0]
    a = G["long_a"]

    def drifting(v):
        calls[0] += 1
        return (a @ v) * (1e9 if calls[0] >= 82 else 1.0)

    x, info = O.bicgstab(drifting, lambda v: v, G["long_b"], np.zeros(G["long_b"].size, complex), tolerance=1e-30,
                         max_iterations=200)
    check_info(info, "div", res_rtol=1e-6)
    assert info["reason"] == "diverged" and info["iterations"] == 40 and info["steps"] == 83
    assert rel(x, G["div_x"]) < 1e-10


def test_converged_equals_direct_solve():
    """A T-operator Helmholtz system (theranostic matrix) with a Jacobi right
    preconditioner: BiCGSTAB to 1e-11 equals splu, and the residual it reports
    is the true one."""
    axis = np.linspace(0.0, 3.0, 21)
    grid = O.Grid(axis, axis, 0.0)
    rng = np.random.default_rng(2)
    mk = 0.1 * rng.uniform(-1, 1, grid.shape)
    me = 0.1 * rng.uniform(-1, 1, grid.shape)
    a = O.assemble(grid, 0.4, mk, me, kappa_factor=np.full(grid.size, 1 + 1j), c0=1500.0, check_pml=False).tocsc()
    b = rng.normal(size=grid.size) + 1j * rng.normal(size=grid.size)
    d = 1 / a.diagonal()
    x, info = O.bicgstab(lambda v: a @ v, lambda v: d * v, b, np.zeros_like(b), tolerance=1e-11, max_iterations=2000)
    assert info["reason"] == "converged"
    ref = splu(a).solve(b)
    assert rel(x, ref) < 1e-9
    assert abs(info["residual"] - np.linalg.norm(a @ x - b) / np.linalg.norm(b)) < 1e-14
