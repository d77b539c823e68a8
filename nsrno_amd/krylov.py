"""Flexible GMRES over caller callables, driven by libhelm.so's Krylov kernels.

The reference's solver API is callable-based: ``gmres(operator, precondition,
rhs, x, callback, restart, max_cycles, tolerance)`` takes any tensor -> tensor
``operator`` and ``precondition`` (nsrno/wave_solver.py:9-36), and
``batched_gmres_cycle(operator, precondition, rhs, iterations)`` runs one
unrestarted cycle (nsrno/dense_wave.py:84-116).  Here the callables are
applied by the caller's own code (torch ops, this package's plans, anything
on torch's current stream) while the library's reverse-communication entry
points (helm_krylov_*) do the rest: delayed classical Gram-Schmidt with FP64
inner products, FP64 Givens least squares on the device, and x accumulated
in complex128 so that every restart is an FP64 refinement step when the
operator is applied in complex128.

``KrylovSpace`` owns the basis V (c64 [m+1, R, n]), the directions Z
(c64 [m, R, n]) and the product slot w (c64 [R, n]) as torch tensors and a
``helm_krylov`` handle with the small per-space state.  With ``batched=False``
(the reference's semantics) the whole ``rhs`` tensor is one vector and one
Krylov space; with ``batched=True`` its leading dimension indexes independent
spaces that run in lock step.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import check, lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class KrylovSpace:
    """Krylov state for R spaces of n complex entries (helm_krylov_create)."""

    def __init__(self, n, nrhs, restart, device="cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("KrylovSpace needs a GPU (no CPU fallback)")
        self.n, self.nrhs, self.m = int(n), int(nrhs), int(restart)
        self.device = torch.device(device)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().helm_krylov_create(self.n, self.nrhs, self.m, ctypes.byref(h)), "helm_krylov_create")
        self._h = h
        self.V = torch.empty((self.m + 1, self.nrhs, self.n), dtype=torch.complex64, device=self.device)
        self.Z = torch.empty((self.m, self.nrhs, self.n), dtype=torch.complex64, device=self.device)
        self.w = torch.empty((self.nrhs, self.n), dtype=torch.complex64, device=self.device)

    def close(self):
        if getattr(self, "_h", None):
            lib().helm_krylov_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def begin(self, r, b=None, tolerance=1e-6):
        """Start a cycle from the residual r (c128 [R, n]); b given on the first
        cycle sets ||b||.  Returns (relative residuals, active flags) per space."""
        r = r.to(torch.complex128).reshape(self.nrhs, self.n).contiguous()
        rel = (ctypes.c_double * self.nrhs)()
        act = (ctypes.c_int32 * self.nrhs)()
        bb = None if b is None else b.to(torch.complex128).reshape(self.nrhs, self.n).contiguous()
        check(lib().helm_krylov_begin(self._h, None if bb is None else _ptr(bb), _ptr(r), _ptr(self.V),
                                      float(tolerance), int(bb is not None), rel, act, _stream()),
              "helm_krylov_begin")
        return [rel[i] for i in range(self.nrhs)], [bool(act[i]) for i in range(self.nrhs)]

    def step(self, j, tolerance=1e-6, last=False):
        """Orthogonalise w (= A Z_j, written by the caller) into V_{j+1}."""
        check(lib().helm_krylov_step(self._h, int(j), int(bool(last)), _ptr(self.V), _ptr(self.w), float(tolerance),
                                     _stream()), "helm_krylov_step")

    def end(self, x):
        """x (c128 [R, n], updated in place) += Z y; returns the steps used per space."""
        if x.dtype != torch.complex128 or not x.is_contiguous() or x.numel() != self.nrhs * self.n:
            raise ValueError("x must be a contiguous complex128 tensor of R * n entries")
        used = (ctypes.c_int32 * self.nrhs)()
        check(lib().helm_krylov_end(self._h, _ptr(self.Z), _ptr(x), used, _stream()), "helm_krylov_end")
        return [used[i] for i in range(self.nrhs)]


def _as_field(v, shape, dtype):
    return v.reshape(shape).to(dtype)


def fgmres(operator, precondition, rhs, x, callback=None, restart=32, max_cycles=80, tolerance=1e-4, batched=False,
           space=None):
    """Right-preconditioned restarted flexible GMRES over callables.

    Semantics of nsrno/wave_solver.py:9-36: convergence is decided on the true
    residual ``rhs - operator(x)`` at each restart, ``callback(applications,
    x, relative)`` fires there (x the current iterate, relative a float, or a
    list of floats when batched), and the result is ``(x, reason)`` with
    reason 'converged' or 'iteration_budget' (a list of reasons when
    batched).  ``operator`` and ``precondition`` see tensors of ``rhs``'s
    shape and dtype.  The basis and directions are stored in complex64 and
    the inner products, least squares and x in FP64; the Arnoldi relation
    uses the stored directions, w = operator(Z_j), so the FGMRES update is
    exact for them.
    """
    if rhs.device.type != "cuda":
        raise ValueError("rhs must be a GPU tensor")
    shape, dtype = rhs.shape, rhs.dtype
    if not dtype.is_complex:
        raise ValueError("rhs must be complex")
    nrhs = shape[0] if batched else 1
    n = rhs.numel() // nrhs
    own = space is None
    space = space or KrylovSpace(n, nrhs, restart, device=rhs.device)
    if (space.n, space.nrhs) != (n, nrhs) or space.m < restart:
        raise ValueError("the Krylov space does not match rhs and restart")
    x64 = x.to(torch.complex128).reshape(nrhs, n).contiguous().clone()
    b = rhs.reshape(nrhs, n)
    applications = 0
    first = True
    try:
        for _ in range(max_cycles):
            xr = _as_field(x64, shape, dtype)
            r = (b.to(torch.complex128) - operator(xr).reshape(nrhs, n).to(torch.complex128)).contiguous()
            rel, act = space.begin(r, b if first else None, tolerance)
            first = False
            if callback is not None:
                callback(applications, xr, rel if batched else rel[0])
            if not any(act):
                return xr, (["converged"] * nrhs if batched else "converged")
            for j in range(restart):
                z = precondition(_as_field(space.V[j], shape, dtype))
                space.Z[j].copy_(z.reshape(nrhs, n))
                space.w.copy_(operator(_as_field(space.Z[j], shape, dtype)).reshape(nrhs, n))
                applications += 1
                space.step(j, tolerance, last=j + 1 == restart)
            space.end(x64)
        xr = _as_field(x64, shape, dtype)
        r = (b.to(torch.complex128) - operator(xr).reshape(nrhs, n).to(torch.complex128))
        rel = (torch.linalg.vector_norm(r, dim=1) /
               torch.linalg.vector_norm(b.to(torch.complex128), dim=1).clamp_min(1e-30)).tolist()
        if callback is not None:
            callback(applications, xr, rel if batched else rel[0])
        reasons = ["converged" if q < tolerance else "iteration_budget" for q in rel]
        return xr, (reasons if batched else reasons[0])
    finally:
        if own:
            space.close()


def gmres_cycle(operator, precondition, rhs, iterations=8, batched=False):
    """One unrestarted GMRES cycle of ``iterations`` steps from zero
    (nsrno/dense_wave.py:84-116, inference branch): returns the minimal-
    residual combination of the directions, the reference's learned warm
    start."""
    shape, dtype = rhs.shape, rhs.dtype
    nrhs = shape[0] if batched else 1
    n = rhs.numel() // nrhs
    space = KrylovSpace(n, nrhs, iterations, device=rhs.device)
    try:
        b = rhs.reshape(nrhs, n).to(torch.complex128).contiguous()
        space.begin(b, b, 1e-300)
        for j in range(iterations):
            z = precondition(_as_field(space.V[j], shape, dtype))
            space.Z[j].copy_(z.reshape(nrhs, n))
            space.w.copy_(operator(_as_field(space.Z[j], shape, dtype)).reshape(nrhs, n))
            space.step(j, 1e-300, last=j + 1 == iterations)
        x = torch.zeros_like(b)
        space.end(x)
        return x.reshape(shape).to(dtype)
    finally:
        space.close()
