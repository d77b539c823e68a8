"""GPU: per-right-hand-side termination reasons and the automatic fallback
(nsrno/dense_wave.py:118-148 reasons, :212-216 fallback to GMRES with the
fixed preconditioner; BASELINE's "plain GMRES as the automatic fallback" of
the learned preconditioner).

* BiCGSTAB with two right-hand sides, one of them so small that rho =
  <r~, r> is below the reference's absolute 1e-30 at the first step: that one
  ends 'breakdown' (as the oracle's restatement of the reference does on the
  same vector), the other converges unaffected; with the fallback on, the
  broken-down one is finished by multigrid FGMRES and reaches the direct
  solve.
* FGMRES with a deliberately useless learned preconditioner (random network
  weights with an output scale of 1e4, so its direction swamps the
  multigrid's): stagnation is detected over the window; with the fallback on
  the solve goes on with the multigrid and converges to the direct solve,
  with it off that right-hand side stops with reason 'stagnation'.
"""
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


def setup(nrhs=2):
    from nsrno_amd.helmholtz import HelmholtzPlan, Ports
    axis = GOLDEN["axis"]
    grid = O.Grid(axis, axis, 6.0)
    f = float(GOLDEN["frequency_mhz"])
    plan = HelmholtzPlan(grid.shape[0], grid.shape[1], grid.h_mm, f, 6.0, x0_mm=axis[0], z0_mm=axis[0],
                         nrhs=nrhs, restart=30)
    plan.set_medium(1 + GOLDEN["m_eta"], (1 + GOLDEN["m_kappa"]).astype(complex))
    plan.setup_multigrid(levels=3, shift=0.5, pre=2, post=2, coarse_sweeps=30)
    w = O.source_weights(grid, GOLDEN["positions"], GOLDEN["normals"], GOLDEN["widths"])
    rhs = plan.inject(Ports(w), np.eye(4)[:nrhs])
    a = O.assemble(grid, f, GOLDEN["m_kappa"], GOLDEN["m_eta"]).tocsc()
    return grid, plan, rhs, a


def test_bicgstab_reasons_per_rhs_and_fallback():
    grid, plan, rhs, a = setup()
    rhs = rhs.clone()
    rhs[0] *= 1e-16 / float(torch.linalg.vector_norm(rhs[0]))  # |<r~, r>| = 1e-32 < 1e-30
    b = rhs.cpu().numpy().reshape(2, -1)
    _, oinfo = O.bicgstab(lambda u: a @ u, lambda u: u, b[0], np.zeros_like(b[0]), tolerance=1e-9)
    assert oinfo["reason"] == "breakdown" and oinfo["iterations"] == 0
    x, infos = plan.solve(rhs, tolerance=1e-9, method="bicgstab", precond="mg", max_iterations=400, fallback=False)
    print([(i.reason, i.iterations, i.relative_residual) for i in infos])
    assert infos[0].reason == "breakdown" and infos[0].iterations == 0 and not infos[0].fallback
    assert infos[1].reason == "converged" and infos[1].iterations > 0
    lu = splu(a)
    assert rel(x[1].cpu().numpy().ravel(), lu.solve(b[1])) < 1e-7
    x, infos = plan.solve(rhs, tolerance=1e-9, method="bicgstab", precond="mg", max_iterations=400, fallback=True)
    print([(i.reason, i.first_reason, i.fallback, i.iterations) for i in infos])
    assert infos[0].reason == "converged" and infos[0].fallback and infos[0].first_reason == "breakdown"
    assert infos[1].reason == "converged" and not infos[1].fallback
    for r in range(2):
        assert rel(x[r].cpu().numpy().ravel(), lu.solve(b[r])) < 1e-7


def test_learned_stagnation_falls_back_to_multigrid():
    from nsrno_amd.learned import LearnedNet, attach, canvas, init_params
    grid, plan, rhs, a = setup()
    net = LearnedNet(canvas(grid.shape[0]), canvas(grid.shape[1]), 2, params=init_params(seed=3, out_scale=1e4))
    attach(plan, net)
    kw = dict(tolerance=1e-9, restart=6, max_iterations=600, precond="learned", stagnation_steps=12,
              stagnation_ratio=0.5)
    x, infos = plan.solve(rhs, fallback=True, **kw)
    print([(i.reason, i.first_reason, i.fallback, i.iterations) for i in infos])
    assert all(i.reason == "converged" and i.fallback and i.first_reason == "stagnation" for i in infos)
    b = rhs.cpu().numpy().reshape(2, -1)
    lu = splu(a)
    for r in range(2):
        assert rel(x[r].cpu().numpy().ravel(), lu.solve(b[r])) < 1e-7
    x, infos = plan.solve(rhs, fallback=False, **kw)
    assert all(i.reason == "stagnation" and not i.fallback for i in infos)
    net.close()
