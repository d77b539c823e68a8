"""Device-side Helmholtz plan: operator, ports and Krylov solves on one GPU.

``HelmholtzPlan`` owns one ``helm_plan`` of libhelm.so.  Fields are torch
tensors on the GPU in [rhs, z, x] layout; every call is enqueued on torch's
current HIP stream.  The operator is the flux-form stretched-PML Helmholtz
matrix of nsrno/theranostic/wave.py:174-188 (complex symmetric,
exp(-i omega t)), extended by a complex kappa_c for attenuation.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch
from scipy import sparse

from . import _lib
from ._lib import check, lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


PRECONDITIONERS = {"none": _lib.PC_NONE, "mg": _lib.PC_MG, "learned": _lib.PC_LEARNED, "spectral": _lib.PC_SPECTRAL,
                   "spectral_plain": _lib.PC_SPECTRAL_PLAIN}


@dataclass
class SolveInfo:
    """Per right-hand-side result of a solve (keys of wave_solver.py:65-67)."""
    reason: str
    iterations: int
    applications: int
    cycles: int
    relative_residual: float
    fallback: bool = False      # finished by the fallback solver (multigrid FGMRES)
    first_reason: str = ""      # reason of the first solver, before any fallback
    graph_fallbacks: int = 0    # restart cycles run directly because their graph capture failed


class Ports:
    """Sparse real port weights W (N x P) laid out for the device.

    Node-major CSR drives injection (rhs = -W a / h^2, wave.py:260), port-major
    CSC drives sampling (W^T u, wave.py:275-278).
    """

    def __init__(self, weights, device="cuda"):
        w = sparse.csc_matrix(weights, dtype=np.float64)
        if np.iscomplexobj(weights):
            raise ValueError("port weights must be real")
        self.n_nodes_total, self.n_ports = w.shape
        csr = w.tocsr()
        csr.sort_indices()
        counts = np.diff(csr.indptr)
        nodes = np.flatnonzero(counts)
        ptr = np.concatenate(([0], np.cumsum(counts[nodes]))).astype(np.int32)
        dev = torch.device(device)
        self.node_ptr = torch.from_numpy(ptr).to(dev)
        self.node_idx = torch.from_numpy(nodes.astype(np.int32)).to(dev)
        self.port_of_tap = torch.from_numpy(csr.indices.astype(np.int32)).to(dev)
        self.w_node = torch.from_numpy(csr.data.astype(np.float64)).to(dev)
        self.n_nodes = int(nodes.size)
        w.sort_indices()
        self.port_ptr = torch.from_numpy(w.indptr.astype(np.int32)).to(dev)
        self.node_of_tap = torch.from_numpy(w.indices.astype(np.int32)).to(dev)
        self.w_port = torch.from_numpy(w.data.astype(np.float64)).to(dev)


class HelmholtzPlan:
    """One grid + frequency + PML on the GPU (helm_plan_create / helm_set_medium)."""

    def __init__(self, nz, nx, h_mm, frequency_mhz, pml_mm, *, x0_mm=0.0, z0_mm=0.0, c0_m_s=1540.0,
                 pml_strength=12.0, nrhs=1, restart=30, device="cuda", op_kind=_lib.HELM_OP_FLUX5):
        if not torch.cuda.is_available():
            raise RuntimeError("HelmholtzPlan needs a GPU (no CPU fallback)")
        self.device = torch.device(device)
        self.nz, self.nx, self.n = int(nz), int(nx), int(nz) * int(nx)
        self.h_mm, self.h_m = float(h_mm), float(h_mm) * 1e-3
        self.frequency_mhz = float(frequency_mhz)
        self.omega = 2 * np.pi * self.frequency_mhz * 1e6
        self.c0 = float(c0_m_s)
        self.k0 = self.omega / self.c0
        self.nrhs_capacity, self.restart_capacity = int(nrhs), int(restart)
        desc = _lib.HelmDesc(self.nz, self.nx, int(nrhs), int(restart), float(x0_mm), float(z0_mm),
                             float(h_mm), float(pml_mm), float(frequency_mhz), float(c0_m_s), float(pml_strength),
                             int(op_kind))
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().helm_plan_create(ctypes.byref(desc), ctypes.byref(handle)), "helm_plan_create")
        self._h = handle
        self.mg_levels = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().helm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ medium
    def _field(self, a, dtype):
        t = torch.as_tensor(a, device=self.device, dtype=dtype)
        if t.numel() != self.n:
            raise ValueError(f"medium map must have {self.n} values, got {t.numel()}")
        return t.reshape(self.nz, self.nx).contiguous()

    def set_medium(self, eta_rel, kappa_c):
        """eta_rel = 1 + m_eta (real), kappa_c = (1 + m_kappa) * attenuation factor (complex)."""
        self._eta = self._field(eta_rel, torch.float64)
        self._kap = self._field(kappa_c, torch.complex128)
        check(lib().helm_set_medium(self._h, _ptr(self._eta), _ptr(self._kap), _stream()), "helm_set_medium")

    def set_medium_physical(self, c_m_s, rho_kg_m3, alpha_np_m=None, rho0=1000.0):
        c = self._field(c_m_s, torch.float64)
        rho = self._field(rho_kg_m3, torch.float64)
        alpha = None if alpha_np_m is None else self._field(alpha_np_m, torch.float64)
        self._phys = (c, rho, alpha)
        check(lib().helm_set_medium_physical(self._h, _ptr(c), _ptr(rho), None if alpha is None else _ptr(alpha),
                                              float(rho0), _stream()), "helm_set_medium_physical")

    # ------------------------------------------------------------ operator
    def _check_field(self, u):
        if u.device.type != "cuda" or not u.is_contiguous():
            raise ValueError("fields must be contiguous GPU tensors")
        if u.dtype not in (torch.complex64, torch.complex128):
            raise ValueError("fields must be complex64 or complex128")
        if u.numel() % self.n:
            raise ValueError(f"field size {u.numel()} is not a multiple of {self.n}")
        return u.numel() // self.n

    def apply(self, u, adjoint=False, out=None):
        """A u (or A^H u) for fields [..., nz, nx]: FP32 for complex64, FP64 for complex128."""
        nrhs = self._check_field(u)
        if out is None:
            out = torch.empty_like(u)
        else:
            self._check_field(out)
            if out.shape != u.shape or out.dtype != u.dtype or out.device != u.device:
                raise ValueError("out must match u in shape, dtype and device")
        prec = 32 if u.dtype == torch.complex64 else 64
        check(lib().helm_apply(self._h, _ptr(u), _ptr(out), nrhs, prec, int(bool(adjoint)), _stream()),
              "helm_apply")
        return out

    # ------------------------------------------------------------ ports
    def inject(self, ports: Ports, amplitudes):
        """rhs[r] = -(1/h^2) W @ amplitudes[r]  ->  complex128 [R, nz, nx]."""
        amp = torch.as_tensor(amplitudes, dtype=torch.complex128, device=self.device)
        if amp.ndim == 1:
            amp = amp[None]
        amp = amp.contiguous()
        if amp.shape[1] != ports.n_ports:
            raise ValueError("amplitudes must be [nrhs, n_ports]")
        nrhs = amp.shape[0]
        rhs = torch.empty((nrhs, self.nz, self.nx), dtype=torch.complex128, device=self.device)
        check(lib().helm_inject(self._h, ports.n_ports, ports.n_nodes, _ptr(ports.node_ptr), _ptr(ports.node_idx),
                                _ptr(ports.port_of_tap), _ptr(ports.w_node), _ptr(amp), _ptr(rhs), nrhs, _stream()),
              "helm_inject")
        return rhs

    def sample(self, ports: Ports, field):
        """W^T u for every field -> complex128 [R, n_ports]."""
        nrhs = self._check_field(field)
        traces = torch.empty((nrhs, ports.n_ports), dtype=torch.complex128, device=self.device)
        prec = 32 if field.dtype == torch.complex64 else 64
        check(lib().helm_sample(self._h, ports.n_ports, _ptr(ports.port_ptr), _ptr(ports.node_of_tap),
                                _ptr(ports.w_port), _ptr(field), prec, _ptr(traces), nrhs, _stream()), "helm_sample")
        return traces

    # ------------------------------------------------------------ preconditioners
    def setup_multigrid(self, levels=5, shift=0.5, pre=2, post=2, coarse_sweeps=10, omega=0.0):
        check(lib().helm_mg_setup(self._h, int(levels), float(shift), int(pre), int(post), int(coarse_sweeps),
                                  float(omega), _stream()), "helm_mg_setup")
        self.mg_levels = int(levels)

    def set_spectral_preconditioner(self, multiplier, gain=None):
        """Fourier-multiplier preconditioner gain * ifft2(multiplier * fft2(v))
        on the plan's periodic box (helm_spectral_set_preconditioner); then
        precond="spectral" / kind=PC_SPECTRAL use it, precond="spectral_plain" /
        PC_SPECTRAL_PLAIN the same without the gain."""
        m = torch.as_tensor(multiplier, device=self.device).to(torch.complex128).reshape(self.nz, self.nx).contiguous()
        g = None if gain is None else torch.as_tensor(gain, device=self.device).to(torch.complex128).reshape(
            self.nz, self.nx).contiguous()
        self._spectral = (m, g)
        check(lib().helm_spectral_set_preconditioner(self._h, _ptr(m), None if g is None else _ptr(g), _stream()),
              "helm_spectral_set_preconditioner")

    def precondition(self, v, kind=_lib.PC_MG, adjoint=False):
        """z = M^{-1} v (adjoint: M^{-H} v) for complex64 fields [..., nz, nx]."""
        nrhs = self._check_field(v)
        if v.dtype != torch.complex64:
            raise ValueError("preconditioner input must be complex64")
        z = torch.empty_like(v)
        check(lib().helm_precond_apply_ex(self._h, int(kind), _ptr(v), _ptr(z), nrhs, int(bool(adjoint)), _stream()),
              "helm_precond_apply_ex")
        return z

    # ------------------------------------------------------------ kernel timing
    def autotune(self, nrhs=None):
        """Pick the stencil and multigrid launch shapes by timing them on this
        plan (helm_autotune); returns the choice and the timings."""
        t = _lib.HelmTuning()
        n = self.nrhs_capacity if nrhs is None else int(nrhs)
        check(lib().helm_autotune(self._h, n, ctypes.byref(t), _stream()), "helm_autotune")
        return dict(stencil_rows=t.stencil_rows, stencil_rhs=t.stencil_rhs, mg_workgroups=t.mg_workgroups,
                    stencil_us=t.stencil_us, vcycle_us=t.vcycle_us, stencil_block=t.stencil_block)

    def set_stencil_shape(self, rows=0, rhs=4, form=1):
        """Fix the FP32 stencil's launch shape (helm_set_stencil_shape): form 0
        streaming, 1 block, 2 block in XCD-contiguous order, 3 / 4 the paired
        block form (hardware / XCD-contiguous order), 5 / 6 block / paired in
        XCD-contiguous order with the RHS groups of a run together; rhs 1, 2,
        4 (8 in forms 1-6); rows=0 restores the default."""
        check(lib().helm_set_stencil_shape(self._h, int(rows), int(rhs), int(form)), "helm_set_stencil_shape")

    PROFILE_KERNELS = ("operator_apply", "preconditioner", "dcgs_dots", "dcgs_update", "true_residual",
                       "mg_pre2_restrict", "mg_prolong_post2", "mg_pre2_restrict_finest", "mg_prolong_post2_finest")

    def profile(self, kernels=None, enable=True):
        """Turn HIP-event timing of kernels inside solves on (clearing totals) or off.

        kernels: names from PROFILE_KERNELS (None = all).  Every timed launch
        adds an event pair to the stream, so time only what is needed."""
        mask = 0
        if enable:
            names = self.PROFILE_KERNELS if kernels is None else kernels
            for k in names:
                mask |= 1 << self.PROFILE_KERNELS.index(k)
        check(lib().helm_profile(self._h, mask), "helm_profile")

    def profile_read(self):
        """{kernel: dict(launches, total_ms, avg_us, bytes, gb_s)} accumulated since profile(True)."""
        stats = (_lib.HelmKernelStat * 16)()
        count = ctypes.c_int32()
        check(lib().helm_profile_read(self._h, stats, 16, ctypes.byref(count)), "helm_profile_read")
        out = {}
        for st in stats[:count.value]:
            n = int(st.launches)
            out[st.name.decode()] = dict(launches=n, total_ms=st.total_ms,
                                         avg_us=1e3 * st.total_ms / n if n else 0.0, bytes=st.bytes,
                                         gb_s=st.bytes / (st.total_ms * 1e6) if st.total_ms > 0 else 0.0)
        return out

    # ------------------------------------------------------------ solve
    def warm_start(self, rhs, x, steps, *, precond="mg", adjoint=False):
        """`steps` minimal-residual preconditioned steps on x in place (complex128 [R, nz, nx])."""
        if x.dtype != torch.complex128 or rhs.dtype != torch.complex128 or not x.is_contiguous():
            raise ValueError("warm_start takes contiguous complex128 rhs and x")
        if x.device.type != "cuda":
            raise ValueError("warm_start takes GPU tensors")
        nrhs = self._check_field(rhs)
        if x.shape != rhs.shape:
            raise ValueError("x and rhs must have the same shape")
        pc = PRECONDITIONERS[precond]
        check(lib().helm_warm_start(self._h, pc, _ptr(rhs), _ptr(x), nrhs, int(steps), int(bool(adjoint)),
                                    _stream()), "helm_warm_start")
        return x

    def solve(self, rhs, x0=None, *, tolerance=1e-6, restart=None, max_iterations=3000, precond="none",
              method="gmres", adjoint=False, callback=None, graph=False, fallback=None, stagnation_steps=0,
              stagnation_ratio=0.0):
        """Krylov solve of A x = rhs (A^H x = rhs) for all right-hand sides.

        rhs: complex128 [R, nz, nx] (converted if needed).  Returns
        (x complex128 [R, nz, nx], [SolveInfo] * R).  graph: replay full restart
        cycles after the first as one captured HIP graph (GMRES).
        fallback (default: on for the learned preconditioner): a stagnating
        learned FGMRES goes on with the multigrid, and BiCGSTAB right-hand
        sides that end unconverged are finished by multigrid FGMRES
        (dense_wave.py:212-216); off: stagnation stops a right-hand side with
        reason "stagnation".  stagnation_steps / stagnation_ratio: the window
        (Krylov steps) and the residual reduction it must reach (0: 60, 0.5).
        """
        rhs = torch.as_tensor(rhs, device=self.device).to(torch.complex128).contiguous()
        nrhs = self._check_field(rhs)
        if nrhs > self.nrhs_capacity:
            raise ValueError(f"nrhs {nrhs} exceeds the plan capacity {self.nrhs_capacity}")
        restart = self.restart_capacity if restart is None else int(restart)
        if x0 is None:
            x = torch.empty_like(rhs)
        else:
            x = torch.as_tensor(x0, device=self.device).to(torch.complex128).contiguous().clone()
        pc = PRECONDITIONERS[precond]
        meth = {"gmres": _lib.HELM_GMRES, "bicgstab": _lib.HELM_BICGSTAB}[method]

        def _cb(user, applications, rel, n):
            if callback is not None:
                callback(int(applications), [rel[i] for i in range(n)])

        cfun = _lib.CALLBACK(_cb)
        if fallback is None:
            fallback = precond == "learned"
        opts = _lib.HelmSolveOpts(meth, pc, restart, int(max_iterations), float(tolerance), int(bool(adjoint)),
                                  int(x0 is not None), cfun, None, int(bool(graph)), int(bool(fallback)),
                                  int(stagnation_steps), float(stagnation_ratio))
        infos = (_lib.HelmSolveInfo * nrhs)()
        check(lib().helm_solve(self._h, _ptr(rhs), _ptr(x), nrhs, ctypes.byref(opts), infos, _stream()),
              "helm_solve")
        out = [SolveInfo(_lib.REASONS.get(i.reason, "unknown"), i.iterations, i.applications, i.cycles,
                         i.relative_residual, bool(i.fallback), _lib.REASONS.get(i.first_reason, "unknown"),
                         int(i.graph_fallbacks)) for i in infos]
        return x, out


def solve_batch(plans, rhs, x0=None, *, tolerance=1e-6, restart=3, max_iterations=3000, adjoint=False, graph=True):
    """Multigrid FGMRES on several plans at once (helm_batch_solve).

    plans: HelmholtzPlans of any grid sizes, each with its medium and
    multigrid set (pre = post = 2); rhs: one complex128 [R, nz_k, nx_k] per
    plan, the same R for all; x0: starting fields (None: zero).  Every kernel
    of a restart cycle is one segmented launch over all plans still running,
    the cycle a captured HIP graph.  Each plan's numerics are those of
    plan.solve(precond="mg") on its finest-level fused path.  Returns
    (list of x, list of [SolveInfo] * R)."""
    plans = list(plans)
    if len(plans) != len(rhs) or not plans:
        raise ValueError("solve_batch: one right-hand side batch per plan")
    if x0 is not None and len(x0) != len(plans):
        raise ValueError("solve_batch: one starting field batch per plan")
    bs, xs = [], []
    nrhs = None
    for k, (p, b) in enumerate(zip(plans, rhs)):
        b = torch.as_tensor(b, device=p.device).to(torch.complex128).contiguous()
        r = p._check_field(b)
        if nrhs is None:
            nrhs = r
        elif r != nrhs:
            raise ValueError("solve_batch: every plan needs the same number of right-hand sides")
        if r > p.nrhs_capacity:
            raise ValueError(f"nrhs {r} exceeds the capacity {p.nrhs_capacity} of plan {k}")
        bs.append(b)
        if x0 is None:
            xs.append(torch.empty_like(b))
        else:
            xs.append(torch.as_tensor(x0[k], device=p.device).to(torch.complex128).contiguous().clone())
    n = len(plans)
    hp = (ctypes.c_void_p * n)(*[p._h for p in plans])
    hb = (ctypes.c_void_p * n)(*[_ptr(b) for b in bs])
    hx = (ctypes.c_void_p * n)(*[_ptr(x) for x in xs])
    opts = _lib.HelmSolveOpts(_lib.HELM_GMRES, PRECONDITIONERS["mg"], int(restart), int(max_iterations),
                              float(tolerance), int(bool(adjoint)), int(x0 is not None), _lib.CALLBACK(), None,
                              int(bool(graph)), 0, 0, 0.0)
    infos = (_lib.HelmSolveInfo * (n * nrhs))()
    check(lib().helm_batch_solve(n, hp, hb, hx, nrhs, ctypes.byref(opts), infos, _stream()), "helm_batch_solve")
    out = [[SolveInfo(_lib.REASONS.get(i.reason, "unknown"), i.iterations, i.applications, i.cycles,
                      i.relative_residual, bool(i.fallback), _lib.REASONS.get(i.first_reason, "unknown"),
                      int(i.graph_fallbacks)) for i in infos[k * nrhs:(k + 1) * nrhs]] for k in range(n)]
    return xs, out
