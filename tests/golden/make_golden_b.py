"""Write golden vectors of the reference's BiCGSTAB (run here only).

Imports nsrno.dense_wave.bicgstab from /root/reference (never shipped) and
runs it in complex128 on small seeded systems, one per branch of
nsrno/dense_wave.py:118-148: a converged solve, the iterates after 1, 2, 3
and 7 steps, the half-step exit (||s|| below the tolerance), the breakdown
on <r~, v> = 0, a solve long enough to pass the 40-step residual refresh,
and a solve that ends 'diverged' at that refresh.  The recurrence's
iterates and its reason/step bookkeeping have no closed form that pins them.

Usage:  python tests/golden/make_golden_b.py  ->  tests/golden/bicgstab_small.npz
"""
import sys
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, "/root/reference")
from nsrno.dense_wave import bicgstab  # noqa: E402


def matmul(a):
    at = torch.tensor(a)
    return lambda v: (at @ v.reshape(-1)).reshape(v.shape)


def diag_precond(d):
    dt = torch.tensor(d)
    return lambda v: (dt * v.reshape(-1)).reshape(v.shape)


def well_conditioned(rng, n):
    return 4 * np.eye(n) + (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)


def main():
    torch.set_default_dtype(torch.float64)
    out = {}
    rng = np.random.default_rng(5)
    # shared-angle shape: 2 x 8 entries form one vector
    a = well_conditioned(rng, 16)
    b = rng.normal(size=(2, 8)) + 1j * rng.normal(size=(2, 8))
    pd = 1 / np.diag(a)
    out.update(a=a, b=b, pdiag=pd)
    cases = [("conv", dict(tolerance=1e-12, max_iterations=800)),
             ("it1", dict(tolerance=1e-30, max_iterations=1)), ("it2", dict(tolerance=1e-30, max_iterations=2)),
             ("it3", dict(tolerance=1e-30, max_iterations=3)), ("it7", dict(tolerance=1e-30, max_iterations=7))]
    for name, kw in cases:
        x, info = bicgstab(matmul(a), diag_precond(pd), torch.tensor(b), torch.zeros(2, 8, dtype=torch.complex128),
                           **kw)
        out[f"{name}_x"] = x.numpy()
        out[f"{name}_info"] = np.array([info["reason"], info["iterations"], info["steps"], info["residual"]],
                                       dtype=object).astype(str)
    # nonzero initial guess
    x0 = rng.normal(size=(2, 8)) + 1j * rng.normal(size=(2, 8))
    x, info = bicgstab(matmul(a), lambda v: v, torch.tensor(b), torch.tensor(x0), tolerance=1e-30, max_iterations=4)
    out.update(x0=x0, x0run_x=x.numpy(),
               x0run_info=np.array([info["reason"], info["iterations"], info["steps"], info["residual"]], dtype=object).astype(str))
    # half-step exit: A = c I makes s = 0 after the first half step
    c = 2.0 - 1.0j
    hb = rng.normal(size=5) + 1j * rng.normal(size=5)
    x, info = bicgstab(lambda v: c * v, lambda v: v, torch.tensor(hb), torch.zeros(5, dtype=torch.complex128),
                       tolerance=1e-8)
    out.update(half_c=c, half_b=hb, half_x=x.numpy(),
               half_info=np.array([info["reason"], info["iterations"], info["steps"], info["residual"]], dtype=object).astype(str))
    # breakdown: <r~, A p> = 0 for a swap operator and b = e0
    sw = np.array([[0, 1], [1, 0]], dtype=complex)
    eb = np.array([1.0 + 0j, 0.0])
    x, info = bicgstab(matmul(sw), lambda v: v, torch.tensor(eb), torch.zeros(2, dtype=torch.complex128),
                       tolerance=1e-8)
    out.update(brk_a=sw, brk_b=eb, brk_x=x.numpy(),
               brk_info=np.array([info["reason"], info["iterations"], info["steps"], info["residual"]], dtype=object).astype(str))
    # past the 40-step refresh, converging (well conditioned: the iterates are
    # insensitive to the order of floating-point sums)
    r2 = np.random.default_rng(23)
    n = 120
    big = np.eye(n) + 0.9 * (r2.normal(size=(n, n)) + 1j * r2.normal(size=(n, n))) / np.sqrt(2 * n)
    bb = r2.normal(size=n) + 1j * r2.normal(size=n)
    x, info = bicgstab(matmul(big), lambda v: v, torch.tensor(bb), torch.zeros(n, dtype=torch.complex128),
                       tolerance=1e-12, max_iterations=400)
    out.update(long_a=big, long_b=bb, long_x=x.numpy(),
               long_info=np.array([info["reason"], info["iterations"], info["steps"], info["residual"]], dtype=object).astype(str))
    # diverged at the first refresh: an operator whose output grows by 1e9 from
    # its 82nd call on (the refresh after 40 steps of two applications each),
    # so the recomputed residual exceeds 1e5 whatever the rounding
    calls = [0]
    at = torch.tensor(big)

    def drifting(v):
        calls[0] += 1
        return (at @ v) * (1e9 if calls[0] >= 82 else 1.0)

    x, info = bicgstab(drifting, lambda v: v, torch.tensor(bb), torch.zeros(n, dtype=torch.complex128),
                       tolerance=1e-30, max_iterations=200)
    out.update(div_x=x.numpy(),
               div_info=np.array([info["reason"], info["iterations"], info["steps"], info["residual"]], dtype=object).astype(str))
    path = Path(__file__).with_name("bicgstab_small.npz")
    np.savez_compressed(path, **out)
    for k in sorted(out):
        if k.endswith("_info"):
            print(k, out[k])


if __name__ == "__main__":
    main()
