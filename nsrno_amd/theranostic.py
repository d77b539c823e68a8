"""Drop-in for the theranostic CPU forward (nsrno/theranostic/wave.py).

Same classes, call signatures, return types and ValueError messages as the
reference ``Grid`` (wave.py:29-87) and ``Solver`` (wave.py:90-278); the sparse
LU factorisation is replaced by the GPU Krylov solve of libhelm.so with FP64
true-residual refinement, run to ``tolerance`` (default 1e-10, far below the
field accuracy the tests ask for).  ``matrix`` returns a matrix-free
``scipy.sparse.linalg.LinearOperator`` backed by the GPU FP64 stencil
instead of an assembled CSC matrix.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from time import perf_counter

import numpy as np
import torch
from scipy import sparse
from scipy.sparse.linalg import LinearOperator

from ._lib import check, lib
from .helmholtz import HelmholtzPlan, Ports


@dataclass(frozen=True)
class Grid:
    """Uniform square grid including its PML; arrays use [z, x] order (wave.py:29-87)."""

    x_mm: np.ndarray
    z_mm: np.ndarray
    pml_mm: float

    def __post_init__(self):
        for name in ("x_mm", "z_mm"):
            values = np.asarray(getattr(self, name), dtype=np.float64)
            if values.ndim != 1 or values.size < 3 or not np.isfinite(values).all():
                raise ValueError(f"{name} must be a finite one-dimensional axis with >=3 nodes")
            step = np.diff(values)
            if np.any(step <= 0):
                raise ValueError(f"{name} must be strictly increasing")
            if not np.allclose(step, step[0], rtol=1e-10, atol=1e-10):
                raise ValueError(f"{name} must be uniformly sampled")
            values = values.copy()
            values.setflags(write=False)
            object.__setattr__(self, name, values)
        if not np.isclose(self.x_mm[1] - self.x_mm[0], self.z_mm[1] - self.z_mm[0], rtol=1e-10, atol=1e-10):
            raise ValueError("The Helmholtz grid must have equal x and z spacing")
        if not np.isfinite(self.pml_mm) or self.pml_mm < 0:
            raise ValueError("pml_mm must be finite and nonnegative")
        if 2 * self.pml_mm >= min(np.ptp(self.x_mm), np.ptp(self.z_mm)):
            raise ValueError("PML leaves no physical domain")

    @property
    def shape(self):
        return (self.z_mm.size, self.x_mm.size)

    @property
    def size(self):
        return self.x_mm.size * self.z_mm.size

    @property
    def h_mm(self):
        return float(self.x_mm[1] - self.x_mm[0])

    @property
    def h_m(self):
        return self.h_mm * 1e-3

    @property
    def physical_bounds_mm(self):
        return (self.x_mm[0] + self.pml_mm, self.x_mm[-1] - self.pml_mm,
                self.z_mm[0] + self.pml_mm, self.z_mm[-1] - self.pml_mm)

    @property
    def pml_mask(self):
        x0, x1, z0, z1 = self.physical_bounds_mm
        inside_x = (self.x_mm >= x0 - 1e-9) & (self.x_mm <= x1 + 1e-9)
        inside_z = (self.z_mm >= z0 - 1e-9) & (self.z_mm <= z1 + 1e-9)
        return ~(inside_z[:, None] & inside_x[None, :])


class Solver:
    """Flux-form Helmholtz solver on a homogeneous exterior (wave.py:90-125).

    ``pml_strength`` defines sigma_max = pml_strength*c0/PML_thickness and the
    stretch s = 1 + i sigma/omega with a quadratic profile.
    """

    def __init__(self, grid, c0_m_s=1540.0, pml_strength=12.0, device="cuda"):
        self.grid = grid
        self.c0_m_s = float(c0_m_s)
        self.pml_strength = float(pml_strength)
        if not np.isfinite(self.c0_m_s) or self.c0_m_s <= 0:
            raise ValueError("c0_m_s must be finite and positive")
        if not np.isfinite(self.pml_strength) or self.pml_strength <= 0:
            raise ValueError("pml_strength must be finite and positive")
        self.device = device

    # ------------------------------------------------------------ validation
    def _contrast(self, value, name):
        """wave.py:158-172."""
        if value is None:
            return np.zeros(self.grid.size)
        raw = np.asarray(value)
        if np.iscomplexobj(raw):
            raise ValueError(f"{name} must be real")
        values = np.asarray(raw, dtype=np.float64)
        if values.shape not in (self.grid.shape, (self.grid.size,)):
            raise ValueError(f"{name} shape must be {self.grid.shape} or ({self.grid.size},)")
        values = values.ravel()
        if not np.isfinite(values).all() or np.any(values <= -1):
            raise ValueError(f"{name} must be finite and > -1")
        if np.any(np.abs(values[self.grid.pml_mask.ravel()]) > 1e-12):
            raise ValueError(f"{name} must vanish in the PML")
        return values

    def _frequency(self, frequency_mhz):
        if not np.isfinite(frequency_mhz) or frequency_mhz <= 0:
            raise ValueError("frequency_mhz must be finite and positive")
        return float(frequency_mhz)

    # ------------------------------------------------------------ coefficients
    def _stretch(self, coordinates, lower, upper, omega):
        """s = 1 + i (sigma_max/omega) clip(frac, 0, 1)^2 (wave.py:127-135)."""
        coords = np.asarray(coordinates, dtype=np.float64)
        if self.grid.pml_mm == 0:
            return np.ones_like(coords, dtype=np.complex128)
        width = self.grid.pml_mm
        frac = np.maximum((lower + width - coords) / width, (coords - upper + width) / width)
        sigma_max = self.pml_strength * self.c0_m_s / (width * 1e-3)
        return 1.0 + 1j * (sigma_max / omega) * np.clip(frac, 0.0, 1.0) ** 2

    def _edges(self):
        """Edge topology (wave.py:107-125): horizontal, vertical, then the four
        exterior ghost groups; edge_j = -1 marks a ghost end."""
        if not hasattr(self, "_edge_cache"):
            nodes = np.arange(self.grid.size).reshape(self.grid.shape)
            bnd = np.concatenate((nodes[:, 0], nodes[:, -1], nodes[0], nodes[-1]))
            ei = np.concatenate((nodes[:, :-1].ravel(), nodes[:-1, :].ravel(), bnd))
            ej = np.concatenate((nodes[:, 1:].ravel(), nodes[1:, :].ravel(), -np.ones(bnd.size, dtype=int)))
            valid = ej >= 0
            rows = np.concatenate((np.arange(ei.size), np.flatnonzero(valid)))
            cols = np.concatenate((ei, ej[valid]))
            avg = sparse.csr_matrix((np.full(rows.size, 0.5), (rows, cols)), shape=(ei.size, self.grid.size))
            self._edge_cache = (ei, ej, avg)
        return self._edge_cache

    def coefficients(self, frequency_mhz):
        """(k0, stretched cell mass, stretched edge flux weights), wave.py:137-156."""
        f = self._frequency(frequency_mhz)
        omega = 2 * np.pi * f * 1e6
        x, z, h = self.grid.x_mm, self.grid.z_mm, self.grid.h_mm
        st = lambda c, a: self._stretch(c, a[0], a[-1], omega)
        sx, sz = st(x, x), st(z, z)
        sxh, szh = st((x[:-1] + x[1:]) / 2, x), st((z[:-1] + z[1:]) / 2, z)
        sxo = st(np.array([x[0] - h / 2, x[-1] + h / 2]), x)
        szo = st(np.array([z[0] - h / 2, z[-1] + h / 2]), z)
        edge = np.concatenate(((sz[:, None] / sxh[None, :]).ravel(), (sx[None, :] / szh[:, None]).ravel(),
                               sz / sxo[0], sz / sxo[1], sx / szo[0], sx / szo[1]))
        return omega / self.c0_m_s, (sz[:, None] * sx[None, :]).ravel(), edge

    def plan(self, frequency_mhz, m_kappa=None, m_eta=None, *, nrhs=1, restart=30):
        """A HelmholtzPlan holding this grid, frequency and contrast on the GPU."""
        f = self._frequency(frequency_mhz)
        mk = self._contrast(m_kappa, "m_kappa")
        me = self._contrast(m_eta, "m_eta")
        g = self.grid
        p = HelmholtzPlan(g.shape[0], g.shape[1], g.h_mm, f, g.pml_mm, x0_mm=float(g.x_mm[0]),
                          z0_mm=float(g.z_mm[0]), c0_m_s=self.c0_m_s, pml_strength=self.pml_strength,
                          nrhs=nrhs, restart=restart, device=self.device)
        p.set_medium(1.0 + me, (1.0 + mk).astype(np.complex128))
        return p

    def matrix(self, frequency_mhz, m_kappa=None, m_eta=None):
        """Matrix-free complex-symmetric flux operator (FP64 on the GPU)."""
        p = self.plan(frequency_mhz, m_kappa, m_eta)
        n = self.grid.size

        def mv(v):
            v = np.asarray(v, dtype=np.complex128)
            cols = v.reshape(n, -1)
            u = torch.from_numpy(np.ascontiguousarray(cols.T)).to(p.device)
            out = p.apply(u.reshape(-1, *self.grid.shape)).reshape(cols.shape[1], n).T
            return out.cpu().numpy().reshape(v.shape)

        op = LinearOperator((n, n), matvec=mv, matmat=mv, dtype=np.complex128)
        op.plan = p
        return op

    # ------------------------------------------------------------ ports
    def source_weights(self, positions_mm, normals, widths_mm, quadrature=5):
        """Unit-strength finite-line element weights (CSC N x n_elements), wave.py:190-238.

        Gauss-Legendre points along the tangent of each element, each spread
        bilinearly over its four surrounding nodes with weight q_w/2.
        """
        positions = np.asarray(positions_mm, dtype=np.float64)
        normal = np.asarray(normals, dtype=np.float64)
        widths = np.asarray(widths_mm, dtype=np.float64)
        if positions.ndim != 2 or positions.shape[1] != 2 or not positions.shape[0]:
            raise ValueError("positions_mm must have shape [n_elements, 2] in [x,z] order")
        if normal.shape != positions.shape or widths.shape != (positions.shape[0],):
            raise ValueError("normals or widths_mm have inconsistent element dimensions")
        if not (np.isfinite(positions).all() and np.isfinite(normal).all() and np.isfinite(widths).all()):
            raise ValueError("Element geometry must be finite")
        length = np.hypot(normal[:, 0], normal[:, 1])
        if np.any(length <= 0) or np.any(widths < 0):
            raise ValueError("Normals must be nonzero and element widths nonnegative")
        if not isinstance(quadrature, (int, np.integer)) or quadrature < 1:
            raise ValueError("quadrature must be a positive integer")
        g = self.grid
        unit = normal / length[:, None]
        t = np.stack((-unit[:, 1], unit[:, 0]), axis=1)            # tangent
        q, qw = np.polynomial.legendre.leggauss(quadrature)
        offs = widths[:, None] * q[None, :] / 2                       # [E, Q]
        x = positions[:, 0, None] + t[:, 0, None] * offs
        z = positions[:, 1, None] + t[:, 1, None] * offs
        x0, x1, z0, z1 = g.physical_bounds_mm
        if np.any((x < x0 - 1e-9) | (x > x1 + 1e-9) | (z < z0 - 1e-9) | (z > z1 + 1e-9)):
            raise ValueError("All transducer quadrature points must be outside the PML")
        if np.any((x <= g.x_mm[0]) | (x >= g.x_mm[-1]) | (z <= g.z_mm[0]) | (z >= g.z_mm[-1])):
            raise ValueError("Transducers must be strictly inside the computational grid")
        fx = (x - g.x_mm[0]) / g.h_mm
        fz = (z - g.z_mm[0]) / g.h_mm
        ix = np.floor(fx).astype(np.int64)
        iz = np.floor(fz).astype(np.int64)
        ax, az = fx - ix, fz - iz
        half_w = qw[None, :] / 2
        nx = g.shape[1]
        taps = [(0, 0, (1 - ax) * (1 - az)), (1, 0, ax * (1 - az)), (0, 1, (1 - ax) * az), (1, 1, ax * az)]
        rows = np.concatenate([((iz + dz) * nx + ix + dx).ravel() for dx, dz, _ in taps])
        vals = np.concatenate([(c * half_w).ravel() for _, _, c in taps])
        cols = np.tile(np.repeat(np.arange(positions.shape[0]), quadrature), len(taps))
        w = sparse.csc_matrix((vals, (rows, cols)), shape=(g.size, positions.shape[0]))
        w.eliminate_zeros()
        return w

    @staticmethod
    def sample(fields, receiver_weights):
        """[transmitter, receiver] data W^T u (wave.py:275-278)."""
        return np.asarray(receiver_weights.T @ fields).T

    # ------------------------------------------------------------ solve
    def solve(self, frequency_mhz, source_weights, m_kappa=None, m_eta=None, *, return_diagnostics=False,
              source_chunk=8, tolerance=1e-10, max_iterations=200000, restart=30, precond="auto"):
        """Solve all port fields (N, n_src) complex128, wave.py:240-273.

        RHS = -W/h^2.  ``source_chunk`` right-hand sides share one batched
        GPU solve.  Diagnostics keep the reference keys; ``factor_seconds``
        is the GPU set-up time and ``lu_nonzeros`` is None (no factorisation).
        """
        if np.iscomplexobj(source_weights):
            raise ValueError("Reciprocal aperture weights must be real; apply transmit encoding to solved port fields")
        weights = sparse.csc_matrix(source_weights, dtype=np.float64)
        if weights.shape[0] != self.grid.size or weights.shape[1] == 0:
            raise ValueError("source_weights has incompatible shape")
        if not isinstance(source_chunk, (int, np.integer)) or source_chunk < 1:
            raise ValueError("source_chunk must be a positive integer")
        if not np.isfinite(weights.data).all():
            raise ValueError("source_weights must be finite")
        started = perf_counter()
        chunk = min(int(source_chunk), weights.shape[1])
        plan = self.plan(frequency_mhz, m_kappa, m_eta, nrhs=chunk, restart=restart)
        use_mg = precond == "mg" or (precond == "auto" and min(self.grid.shape) >= 64)
        if use_mg:
            plan.setup_multigrid()
        ports = Ports(weights, device=self.device)
        torch.cuda.synchronize()
        setup_seconds = perf_counter() - started
        fields = np.empty(weights.shape, dtype=np.complex128)
        err2 = rhs2 = 0.0
        infos = []
        for first in range(0, weights.shape[1], chunk):
            end = min(first + chunk, weights.shape[1])
            amp = np.zeros((end - first, weights.shape[1]), dtype=np.complex128)
            amp[np.arange(end - first), np.arange(first, end)] = 1.0
            rhs = plan.inject(ports, amp)
            x, info = plan.solve(rhs, tolerance=tolerance, restart=restart, max_iterations=max_iterations,
                                 precond="mg" if use_mg else "none")
            fields[:, first:end] = x.reshape(end - first, -1).T.cpu().numpy()
            bn2 = torch.linalg.vector_norm(rhs.reshape(end - first, -1), dim=1).square().cpu().numpy()
            for i, b2 in zip(info, bn2):
                err2 += (i.relative_residual ** 2) * b2
                rhs2 += b2
            infos.extend(info)
        if not return_diagnostics:
            return fields
        diagnostics = dict(grid_shape=list(self.grid.shape), nodes=self.grid.size,
                           frequency_mhz=float(frequency_mhz), sources=weights.shape[1],
                           relative_linear_residual=float(np.sqrt(err2 / max(rhs2, np.finfo(float).tiny))),
                           factor_seconds=float(setup_seconds), elapsed_seconds=float(perf_counter() - started),
                           lu_nonzeros=None,
                           krylov_iterations=[i.iterations for i in infos], reasons=[i.reason for i in infos])
        return fields, diagnostics

    # ------------------------------------------------------------ Born
    def born_data(self, frequency_mhz, source_weights, m_kappa, m_eta, background_fields=None, *,
                  return_diagnostics=False, source_chunk=8, tolerance=1e-10, restart=30, max_iterations=200000):
        """Born data delta_u = -A0^{-1} delta_A u0 at the ports, [tx, rx] (wave.py:280-329).

        delta_A u0 is formed matrix-free on the GPU as A(m) u0 - A0 u0 (FP64
        stencil applies); both solves use the GPU Krylov solver."""
        if np.iscomplexobj(source_weights):
            raise ValueError("Reciprocal aperture weights must be real; apply transmit encoding to solved port fields")
        weights = sparse.csc_matrix(source_weights, dtype=np.float64)
        if weights.shape[0] != self.grid.size or weights.shape[1] == 0:
            raise ValueError("source_weights has incompatible shape")
        if not isinstance(source_chunk, (int, np.integer)) or source_chunk < 1:
            raise ValueError("source_chunk must be a positive integer")
        if not np.isfinite(weights.data).all():
            raise ValueError("source_weights must be finite")
        started = perf_counter()
        n_src = weights.shape[1]
        chunk = min(int(source_chunk), n_src)
        p0 = self.plan(frequency_mhz, nrhs=chunk, restart=restart)
        p1 = self.plan(frequency_mhz, m_kappa, m_eta, nrhs=chunk, restart=restart)
        use_mg = min(self.grid.shape) >= 64
        if use_mg:
            p0.setup_multigrid()
        pc = "mg" if use_mg else "none"
        ports = Ports(weights, device=self.device)
        if background_fields is not None:
            background_fields = np.asarray(background_fields, dtype=np.complex128)
            if background_fields.shape != weights.shape or not np.isfinite(background_fields).all():
                raise ValueError("background_fields has incompatible shape or nonfinite values")
        torch.cuda.synchronize()
        setup = perf_counter() - started
        result = np.empty((n_src, n_src), dtype=np.complex128)
        err2 = rhs2 = 0.0
        shape = (-1,) + self.grid.shape
        for first in range(0, n_src, chunk):
            end = min(first + chunk, n_src)
            if background_fields is None:
                amp = np.zeros((end - first, n_src), dtype=np.complex128)
                amp[np.arange(end - first), np.arange(first, end)] = 1.0
                u0, _ = p0.solve(p0.inject(ports, amp), tolerance=tolerance, restart=restart,
                                 max_iterations=max_iterations, precond=pc)
            else:
                u0 = torch.from_numpy(np.ascontiguousarray(background_fields[:, first:end].T)).to(p0.device)
                u0 = u0.reshape(shape).contiguous()
            rhs = (p0.apply(u0) - p1.apply(u0)).contiguous()
            sc, info = p0.solve(rhs, tolerance=tolerance, restart=restart, max_iterations=max_iterations,
                                precond=pc)
            result[first:end] = p0.sample(ports, sc).cpu().numpy()
            b2 = torch.linalg.vector_norm(rhs.reshape(end - first, -1), dim=1).square().cpu().numpy()
            for i, bb in zip(info, b2):
                err2 += (i.relative_residual ** 2) * bb
                rhs2 += bb
        if not return_diagnostics:
            return result
        return result, dict(grid_shape=list(self.grid.shape), frequency_mhz=float(frequency_mhz), sources=n_src,
                            relative_linear_residual=float(np.sqrt(err2 / max(rhs2, np.finfo(float).tiny))),
                            factor_seconds=float(setup), elapsed_seconds=float(perf_counter() - started))

    def born_jacobian(self, frequency_mhz, background_fields, source_indices, receiver_indices, basis, *,
                      pair_chunk=16):
        """Exact derivative of this discretisation at zero contrast (wave.py:331-389).

        Rows [tx, rx] flattened, columns [m_kappa basis | m_eta basis].  The
        reciprocal field products over the basis support and the contraction
        with the (stretch- and mass-weighted) basis run in libhelm.so's
        helm_born_jacobian kernel (one thread per pair and basis function);
        no conjugate enters the products (inversion of real parameters uses
        Re(J^H r)).  ``pair_chunk`` is accepted for the reference's signature
        and validated; the kernel takes all pairs at once."""
        fields = np.asarray(background_fields, dtype=np.complex128)
        if fields.ndim != 2 or fields.shape[0] != self.grid.size:
            raise ValueError("background_fields has incompatible shape")
        indices = []
        for value in (source_indices, receiver_indices):
            index = np.asarray(value)
            if index.ndim != 1 or not np.issubdtype(index.dtype, np.integer) or index.size == 0:
                raise ValueError("Source/receiver indices must be nonempty integer vectors")
            if np.any(index < 0) or np.any(index >= fields.shape[1]):
                raise ValueError("Source/receiver index out of bounds")
            indices.append(index.astype(int))
        transmitters, receivers = indices
        if not isinstance(pair_chunk, (int, np.integer)) or pair_chunk < 1:
            raise ValueError("pair_chunk must be a positive integer")
        if np.iscomplexobj(basis):
            raise ValueError("Material basis must be real")
        basis = sparse.csr_matrix(basis, dtype=np.float64)
        basis.eliminate_zeros()
        if basis.shape[0] != self.grid.size or basis.shape[1] == 0 or not np.isfinite(basis.data).all():
            raise ValueError("Material basis has incompatible shape or nonfinite entries")
        active_nodes = np.flatnonzero(np.diff(basis.indptr))
        if np.any(self.grid.pml_mask.ravel()[active_nodes]):
            raise ValueError("Material basis must vanish in the PML")
        ei, ej, avg = self._edges()
        edge_basis = (avg @ basis).tocsc()
        edge_basis.eliminate_zeros()
        k0, mass, stretch = self.coefficients(frequency_mhz)
        h2 = self.grid.h_m ** 2
        node_basis = sparse.csc_matrix(sparse.diags(h2 * k0 ** 2 * mass) @ basis, dtype=np.complex128)
        node_basis.sort_indices()
        edge_w = sparse.csc_matrix(sparse.diags(-stretch) @ edge_basis, dtype=np.complex128)
        edge_w.sort_indices()
        tx = np.repeat(transmitters, receivers.size).astype(np.int32)
        rx = np.tile(receivers, transmitters.size).astype(np.int32)
        nb = basis.shape[1]
        dev = torch.device(self.device)

        def up(a, dtype):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(dev)

        # every index the kernel gathers with is checked here (node rows < N, sources < n_src)
        e_idx = edge_w.indices
        if (node_basis.indices.size and node_basis.indices.max() >= self.grid.size) or (
                e_idx.size and (ei[e_idx].max() >= self.grid.size or ej[e_idx].max() >= self.grid.size)):
            raise ValueError("material basis or edge topology indexes outside the grid")
        f_t = up(fields, np.complex128)
        out = torch.empty((tx.size, 2 * nb), dtype=torch.complex128, device=dev)
        args = [up(tx, np.int32), up(rx, np.int32), up(node_basis.indptr, np.int32), up(node_basis.indices, np.int32),
                up(node_basis.data, np.complex128), up(edge_w.indptr, np.int32), up(ei[e_idx], np.int32),
                up(ej[e_idx], np.int32), up(edge_w.data, np.complex128)]
        p = [ctypes.c_void_p(t.data_ptr()) if t.numel() else None for t in args]
        check(lib().helm_born_jacobian(self.grid.size, fields.shape[1], ctypes.c_void_p(f_t.data_ptr()), tx.size,
                                       p[0], p[1], nb, p[2], p[3], p[4], p[5], p[6], p[7], p[8],
                                       ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "helm_born_jacobian")
        return out.cpu().numpy()
