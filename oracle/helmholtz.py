"""CPU oracle for the 2-D variable-density viscoacoustic Helmholtz path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it is imported by the
tests, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, and never by the product package ``nsrno_amd``.  Everything is
plain NumPy/SciPy in float64/complex128.

It restates, function by function, the reference's CPU forward
(``nsrno/theranostic/wave.py``) and its restarted GMRES
(``nsrno/wave_solver.py``):

* ``stretch``            wave.py:127-135   (quadratic complex PML stretch)
* ``coefficients``       wave.py:137-156   (k0, cell mass, edge stretch in the
                                            reference's edge order)
* ``assemble``           wave.py:99-125 and 174-188 (edge topology, CSC matrix)
* ``apply``              matrix-free form of the same matrix (own restatement)
* ``source_weights``     wave.py:190-238   (finite-line Gauss-Legendre ports)
* ``solve``              wave.py:240-273   (splu, RHS = -W/h^2, diagnostics)
* ``sample``             wave.py:275-278
* ``gmres``              wave_solver.py:9-36 (restarted GMRES, CGS twice)

Viscoacoustic extension (the reference's T path has no attenuation): the
diagonal coefficient ``1 + m_kappa`` is replaced by a complex
``kappa_c = (1 + m_kappa) * (1 + 1j*alpha*c/omega)**2``, the T-convention
(exp(-i omega t)) form of the power-law attenuation the reference uses with
the opposite sign convention in nsrno/wave_operator.py:56,67.

How each function is pinned (the checks live in tests/):
* assemble / coefficients / source_weights: golden vectors written by
  tests/golden/make_golden.py from the reference's own ``Solver`` on small
  grids; plus the closed-form outgoing Green's function (i/4)H0(k r) with a
  second-order grid-refinement ratio (tests/test_theranostic.py:141-158),
  reciprocity A = A^T, unit column sums of port weights.
* apply: equality with ``assemble(...) @ u`` and brute-force columns.
* solve: relative residual < 1e-12 and the Green's-function test.
* gmres: equality with the direct solve at convergence on small systems,
  exact termination within N steps on tiny systems, reason semantics.
* attenuation (kappa_c): the 1-D closed form: a plane wave in a homogeneous
  attenuating medium decays as exp(-alpha z) (test_oracle_attenuation).
"""
from __future__ import annotations

from dataclasses import dataclass
from time import perf_counter

import numpy as np
from scipy import sparse
from scipy.sparse.linalg import splu


# --------------------------------------------------------------------------
# grid
# --------------------------------------------------------------------------
@dataclass(frozen=True)
class Grid:
    """Uniform square grid including its PML, arrays in [z, x] order.

    Restates nsrno/theranostic/wave.py:29-87 (validation reduced to what the
    oracle needs).
    """

    x_mm: np.ndarray
    z_mm: np.ndarray
    pml_mm: float

    @property
    def shape(self):
        return (len(self.z_mm), len(self.x_mm))

    @property
    def size(self):
        return len(self.z_mm) * len(self.x_mm)

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
        return ((self.x_mm[None, :] < x0 - 1e-9) | (self.x_mm[None, :] > x1 + 1e-9)
                | (self.z_mm[:, None] < z0 - 1e-9) | (self.z_mm[:, None] > z1 + 1e-9))


def grid_with_pml(n_interior_x, n_interior_z, pml_cells, h_mm, x0_mm=0.0, z0_mm=0.0):
    """Grid of ``n_interior + 2*pml_cells`` nodes per axis, PML of pml_cells*h.

    Same node/PML placement rule as nsrno/theranostic/setup.py:175-193 when
    the requested interval is an exact multiple of h.
    """
    x = x0_mm + h_mm * np.arange(-pml_cells, n_interior_x + pml_cells)
    z = z0_mm + h_mm * np.arange(-pml_cells, n_interior_z + pml_cells)
    return Grid(x, z, pml_cells * h_mm)


# --------------------------------------------------------------------------
# coefficients (wave.py:127-156)
# --------------------------------------------------------------------------
def stretch(coords_mm, lower_mm, upper_mm, pml_mm, omega, c0=1540.0, pml_strength=12.0):
    """s = 1 + i (sigma_max/omega) clip(frac,0,1)^2  (wave.py:127-135)."""
    coords = np.asarray(coords_mm, dtype=np.float64)
    if pml_mm == 0:
        return np.ones_like(coords, dtype=np.complex128)
    width = pml_mm
    frac = np.maximum((lower_mm + width - coords) / width, (coords - upper_mm + width) / width)
    profile = np.clip(frac, 0.0, 1.0) ** 2
    sigma_max = pml_strength * c0 / (width * 1e-3)
    return 1.0 + 1j * (sigma_max / omega) * profile


def profiles(grid, frequency_mhz, c0=1540.0, pml_strength=12.0):
    """1-D stretch profiles: node values and edge-midpoint values.

    Returns (omega, k0, sx[nx], sz[nz], sxe[nx+1], sze[nz+1]) where sxe[k] is
    the stretch at x_0 + (k - 1/2) h, i.e. the edge between nodes k-1 and k
    (k=0 and k=nx are the two exterior ghost edges of wave.py:148-149).
    """
    omega = 2 * np.pi * frequency_mhz * 1e6
    k0 = omega / c0
    x, z, h = grid.x_mm, grid.z_mm, grid.h_mm
    args = dict(pml_mm=grid.pml_mm, omega=omega, c0=c0, pml_strength=pml_strength)
    sx = stretch(x, x[0], x[-1], **args)
    sz = stretch(z, z[0], z[-1], **args)
    xe = x[0] + h * (np.arange(len(x) + 1) - 0.5)
    ze = z[0] + h * (np.arange(len(z) + 1) - 0.5)
    sxe = stretch(xe, x[0], x[-1], **args)
    sze = stretch(ze, z[0], z[-1], **args)
    return omega, k0, sx, sz, sxe, sze


def coefficients(grid, frequency_mhz, c0=1540.0, pml_strength=12.0):
    """(k0, cell mass, edge stretch) in the reference edge order (wave.py:137-156)."""
    omega = 2 * np.pi * frequency_mhz * 1e6
    k0 = omega / c0
    x, z, h = grid.x_mm, grid.z_mm, grid.h_mm
    args = dict(pml_mm=grid.pml_mm, omega=omega, c0=c0, pml_strength=pml_strength)
    sx = stretch(x, x[0], x[-1], **args)
    sz = stretch(z, z[0], z[-1], **args)
    sxh = stretch((x[:-1] + x[1:]) / 2, x[0], x[-1], **args)
    szh = stretch((z[:-1] + z[1:]) / 2, z[0], z[-1], **args)
    sx_out = stretch(np.array([x[0] - h / 2, x[-1] + h / 2]), x[0], x[-1], **args)
    sz_out = stretch(np.array([z[0] - h / 2, z[-1] + h / 2]), z[0], z[-1], **args)
    edge = np.concatenate(((sz[:, None] / sxh[None, :]).ravel(),
                           (sx[None, :] / szh[:, None]).ravel(),
                           sz / sx_out[0], sz / sx_out[1],
                           sx / sz_out[0], sx / sz_out[1]))
    return k0, (sz[:, None] * sx[None, :]).ravel(), edge


def edges(grid):
    """Edge topology of wave.py:107-125: (edge_i, edge_j, edge_average)."""
    nz, nx = grid.shape
    nodes = np.arange(grid.size).reshape(grid.shape)
    hi, hj = nodes[:, :-1].ravel(), nodes[:, 1:].ravel()
    vi, vj = nodes[:-1, :].ravel(), nodes[1:, :].ravel()
    bnd = np.concatenate((nodes[:, 0], nodes[:, -1], nodes[0], nodes[-1]))
    ei = np.concatenate((hi, vi, bnd))
    ej = np.concatenate((hj, vj, -np.ones(bnd.size, dtype=int)))
    valid = ej >= 0
    rows = np.concatenate((np.arange(ei.size), np.flatnonzero(valid)))
    cols = np.concatenate((ei, ej[valid]))
    avg = sparse.csr_matrix((np.full(rows.size, 0.5), (rows, cols)), shape=(ei.size, grid.size))
    return ei, ej, avg


def contrast(grid, value, name, check_pml=True):
    """Validation of wave.py:158-172."""
    if value is None:
        return np.zeros(grid.size)
    raw = np.asarray(value)
    if np.iscomplexobj(raw):
        raise ValueError(f"{name} must be real")
    v = np.asarray(raw, dtype=np.float64)
    if v.shape not in (grid.shape, (grid.size,)):
        raise ValueError(f"{name} shape must be {grid.shape} or ({grid.size},)")
    v = v.ravel()
    if not np.isfinite(v).all() or np.any(v <= -1):
        raise ValueError(f"{name} must be finite and > -1")
    if check_pml and np.any(np.abs(v[grid.pml_mask.ravel()]) > 1e-12):
        raise ValueError(f"{name} must vanish in the PML")
    return v


def attenuation_factor(c_m_s, alpha_np_m, omega):
    """(1 + i alpha c / omega)^2: exp(-i omega t) power-law damping factor."""
    return (1.0 + 1j * np.asarray(alpha_np_m) * np.asarray(c_m_s) / omega) ** 2


def alpha_np_per_m(alpha0_db_cm_mhz, frequency_mhz, power=1.1):
    """dB/cm/MHz^y -> Np/m at f (the unit conversion of wave_operator.py:56)."""
    return np.asarray(alpha0_db_cm_mhz) * 100 * np.log(10) / 20 * frequency_mhz ** power


def assemble(grid, frequency_mhz, m_kappa=None, m_eta=None, *, kappa_factor=None,
             c0=1540.0, pml_strength=12.0, check_pml=True):
    """Complex-symmetric flux matrix (CSC, complex128), wave.py:174-188.

    ``kappa_factor`` (complex per node, default 1) multiplies (1 + m_kappa)
    on the diagonal: the attenuation extension.
    """
    mk = contrast(grid, m_kappa, "m_kappa", check_pml)
    me = contrast(grid, m_eta, "m_eta", check_pml)
    ei, ej, avg = edges(grid)
    k0, mass, edge_stretch = coefficients(grid, frequency_mhz, c0, pml_strength)
    flux = edge_stretch * (1.0 + avg @ me) / grid.h_m ** 2
    kap = (1.0 + mk) if kappa_factor is None else (1.0 + mk) * np.asarray(kappa_factor).ravel()
    diag = (k0 ** 2 * mass * kap).astype(np.complex128)
    np.add.at(diag, ei, -flux)
    valid = ej >= 0
    np.add.at(diag, ej[valid], -flux[valid])
    nodes = np.arange(grid.size)
    row = np.concatenate((nodes, ei[valid], ej[valid]))
    col = np.concatenate((nodes, ej[valid], ei[valid]))
    val = np.concatenate((diag, flux[valid], flux[valid]))
    return sparse.csc_matrix((val, (row, col)), shape=(grid.size, grid.size))


def apply(grid, frequency_mhz, u, eta_rel=None, kappa_c=None, *, c0=1540.0, pml_strength=12.0):
    """Matrix-free A u on [..., nz, nx] arrays (complex128).

    (A u)[i,j] = k0^2 sz_i sx_j kappa_c[i,j] u[i,j]
                 + sum over the 4 edges e of F_e (u_neighbour - u[i,j]),
    F_e = stretch_e * (mean of eta_rel over the two ends; ghost end = 1) / h^2,
    horizontal stretch sz_i / sxe_k, vertical sx_j / sze_k, ghost u = 0.
    Identical to ``assemble(...) @ u`` with eta_rel = 1 + m_eta and
    kappa_c = (1 + m_kappa) * kappa_factor.
    """
    nz, nx = grid.shape
    _, k0, sx, sz, sxe, sze = profiles(grid, frequency_mhz, c0, pml_strength)
    u = np.asarray(u, dtype=np.complex128)
    er = np.ones((nz, nx)) if eta_rel is None else np.asarray(eta_rel, dtype=np.float64).reshape(nz, nx)
    kc = np.ones((nz, nx)) if kappa_c is None else np.asarray(kappa_c).reshape(nz, nx)
    h2 = grid.h_m ** 2
    # eta_rel on the edges, ghost contrast zero -> ghost eta_rel = 1 averaged
    erp = np.pad(er, 1, constant_values=1.0)
    ex = 0.5 * (erp[1:-1, :-1] + erp[1:-1, 1:])      # [nz, nx+1] x-edges
    ez = 0.5 * (erp[:-1, 1:-1] + erp[1:, 1:-1])      # [nz+1, nx] z-edges
    fx = sz[:, None] / sxe[None, :] * ex / h2
    fz = sx[None, :] / sze[:, None] * ez / h2
    up = np.pad(u, [(0, 0)] * (u.ndim - 2) + [(1, 1), (1, 1)])
    c = up[..., 1:-1, 1:-1]
    out = k0 ** 2 * (sz[:, None] * sx[None, :]) * kc * c
    out = out + fx[:, :-1] * (up[..., 1:-1, :-2] - c) + fx[:, 1:] * (up[..., 1:-1, 2:] - c)
    out = out + fz[:-1, :] * (up[..., :-2, 1:-1] - c) + fz[1:, :] * (up[..., 2:, 1:-1] - c)
    return out


# --------------------------------------------------------------------------
# ports (wave.py:190-278)
# --------------------------------------------------------------------------
def source_weights(grid, positions_mm, normals, widths_mm, quadrature=5):
    """Unit-sum finite-line element weights (CSC, N x n_elements), wave.py:190-238."""
    pos = np.asarray(positions_mm, dtype=np.float64)
    nrm = np.asarray(normals, dtype=np.float64)
    wid = np.asarray(widths_mm, dtype=np.float64)
    nrm = nrm / np.linalg.norm(nrm, axis=1)[:, None]
    tangent = np.column_stack((-nrm[:, 1], nrm[:, 0]))
    nodes, weights = np.polynomial.legendre.leggauss(quadrature)
    pts = pos[:, None, :] + tangent[:, None, :] * wid[:, None, None] * nodes[None, :, None] / 2
    x, z = pts[..., 0], pts[..., 1]
    x0, x1, z0, z1 = grid.physical_bounds_mm
    if np.any((x < x0 - 1e-9) | (x > x1 + 1e-9) | (z < z0 - 1e-9) | (z > z1 + 1e-9)):
        raise ValueError("All transducer quadrature points must be outside the PML")
    fx = (x - grid.x_mm[0]) / grid.h_mm
    fz = (z - grid.z_mm[0]) / grid.h_mm
    ix, iz = np.floor(fx).astype(int), np.floor(fz).astype(int)
    wx, wz = fx - ix, fz - iz
    elem = np.broadcast_to(np.arange(pos.shape[0])[:, None], x.shape).ravel()
    rows, cols, vals = [], [], []
    for ddx, ddz, coef in ((0, 0, (1 - wx) * (1 - wz)), (1, 0, wx * (1 - wz)),
                           (0, 1, (1 - wx) * wz), (1, 1, wx * wz)):
        rows.append(((iz + ddz) * grid.shape[1] + ix + ddx).ravel())
        cols.append(elem)
        vals.append((coef * weights[None, :] / 2).ravel())
    w = sparse.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(grid.size, pos.shape[0]))
    w.eliminate_zeros()
    return w


def sample(fields, receiver_weights):
    """[transmitter, receiver] data, wave.py:275-278."""
    return np.asarray(receiver_weights.T @ fields).T


def solve(grid, frequency_mhz, source_weights_csc, m_kappa=None, m_eta=None, *,
          kappa_factor=None, rhs=None, c0=1540.0, pml_strength=12.0, check_pml=True):
    """Direct complex128 solve, wave.py:240-273.

    RHS = -W / h^2 (or an explicit complex ``rhs`` [N, R]).  Returns
    (fields [N, R], diagnostics dict with relative_linear_residual,
    factor_seconds, elapsed_seconds, lu_nonzeros).
    """
    t0 = perf_counter()
    a = assemble(grid, frequency_mhz, m_kappa, m_eta, kappa_factor=kappa_factor,
                 c0=c0, pml_strength=pml_strength, check_pml=check_pml)
    lu = splu(a)
    t_factor = perf_counter() - t0
    if rhs is None:
        rhs = -np.asarray(source_weights_csc.toarray(), dtype=np.complex128) / grid.h_m ** 2
    rhs = np.asarray(rhs, dtype=np.complex128)
    if rhs.ndim == 1:
        rhs = rhs[:, None]
    x = lu.solve(rhs)
    res = np.linalg.norm(a @ x - rhs) / max(np.linalg.norm(rhs), np.finfo(float).tiny)
    return x, dict(relative_linear_residual=float(res), factor_seconds=t_factor,
                   elapsed_seconds=perf_counter() - t0, lu_nonzeros=int(lu.L.nnz + lu.U.nnz))


# --------------------------------------------------------------------------
# restarted GMRES (wave_solver.py:9-36)
# --------------------------------------------------------------------------
def gmres(operator, precondition, rhs, x, callback=None, restart=32, max_cycles=80, tolerance=1e-4):
    """Right-preconditioned restarted GMRES, classical Gram-Schmidt applied twice.

    Restates nsrno/wave_solver.py:9-36: convergence is tested only at restarts
    on the true residual; happy breakdown at |h| < 1e-12; least squares in
    complex128.  Returns (x, reason) with reason 'converged' or
    'iteration_budget'.
    """
    shape = rhs.shape
    n = rhs.size
    scale = max(np.linalg.norm(rhs), 1e-30)
    basis = np.empty((restart + 1, n), dtype=np.complex128)
    directions = np.empty((restart, n), dtype=np.complex128)
    applications = 0
    for _ in range(max_cycles):
        r = rhs - operator(x)
        beta = np.linalg.norm(r)
        rel = beta / scale
        if callback is not None:
            callback(applications, x, rel)
        if rel < tolerance:
            return x, "converged"
        basis[0] = r.ravel() / beta
        h = np.zeros((restart + 1, restart), dtype=np.complex128)
        used = restart
        for j in range(restart):
            z = precondition(basis[j].reshape(shape))
            directions[j] = z.ravel()
            w = operator(z).ravel()
            applications += 1
            for _ in range(2):
                c = basis[:j + 1].conj() @ w
                w = w - basis[:j + 1].T @ c
                h[:j + 1, j] += c
            h[j + 1, j] = np.linalg.norm(w)
            if abs(h[j + 1, j]) < 1e-12:
                used = j + 1
                break
            basis[j + 1] = w / h[j + 1, j]
        q = np.zeros(used + 1, dtype=np.complex128)
        q[0] = beta
        y = np.linalg.lstsq(h[:used + 1, :used], q, rcond=None)[0]
        x = x + (directions[:used].T @ y).reshape(shape)
    if callback is not None:
        callback(applications, x, np.linalg.norm(rhs - operator(x)) / scale)
    return x, "iteration_budget"


def bicgstab(operator, precondition, rhs, x, tolerance=1e-3, max_iterations=800):
    """Right-preconditioned BiCGSTAB, restating nsrno/dense_wave.py:118-148.

    One Krylov recurrence over the whole array ``rhs`` (all of its entries form
    one vector, as the reference's shared-angle batch does).  Branches, in the
    reference's order: convergence test on the recursive ||r|| (:126);
    breakdown when rho = <r~, r> is not finite or |rho| or |omega| < 1e-30
    (:127-129); breakdown when |<r~, v>| < 1e-30 (:132-133); the half-step
    exit when ||s|| is below the tolerance (:135-136); omega with the
    denominator clamped at 1e-30 (:138); every 40th iteration the residual is
    recomputed from x and a non-finite or > 1e5 relative residual ends the
    solve as 'diverged' (:140-142); the final true residual, which upgrades
    any reason to 'converged' when it meets the tolerance (:143-144).
    Returns (x, info) with the reference's keys reason, iterations, steps
    (operator applications), residual, method.
    """
    def norm(a):
        return float(np.linalg.norm(np.ravel(a)))

    def dot(a, b):  # sum(conj(a) * b)
        return complex(np.vdot(np.ravel(a), np.ravel(b)))

    rhs = np.asarray(rhs)
    x = np.array(x, dtype=np.result_type(x, rhs, np.complex64), copy=True)
    scale = max(norm(rhs), 1e-30)
    r = rhs - operator(x)
    shadow = r.copy()
    p = np.zeros_like(r)
    v = np.zeros_like(r)
    rho_old = alpha = omega = 1.0 + 0.0j
    calls, reason, used = 1, "iteration_budget", 0
    for it in range(max_iterations):
        if norm(r) / scale < tolerance:
            reason = "converged"
            break
        rho = dot(shadow, r)
        if not np.isfinite(rho) or abs(rho) < 1e-30 or abs(omega) < 1e-30:
            reason = "breakdown"
            break
        beta = (rho / rho_old) * (alpha / omega)
        p = r + beta * (p - omega * v)
        phat = precondition(p)
        v = operator(phat)
        calls += 1
        den = dot(shadow, v)
        if abs(den) < 1e-30:
            reason = "breakdown"
            break
        alpha = rho / den
        s = r - alpha * v
        if norm(s) / scale < tolerance:
            x = x + alpha * phat
            r = s
            used = it + 1
            reason = "converged"
            break
        shat = precondition(s)
        t = operator(shat)
        calls += 1
        omega = dot(t, s) / max(float(np.sum(np.abs(t) ** 2)), 1e-30)
        x = x + alpha * phat + omega * shat
        r = s - omega * t
        rho_old = rho
        used = it + 1
        if it % 40 == 39:
            r = rhs - operator(x)
            calls += 1
            if not np.all(np.isfinite(r)) or norm(r) / scale > 1e5:
                reason = "diverged"
                break
    actual = norm(operator(x) - rhs) / scale
    calls += 1
    if actual < tolerance:
        reason = "converged"
    return x, dict(reason=reason, iterations=used, steps=calls, residual=actual, method="BiCGSTAB")


# --------------------------------------------------------------------------
# Born data and Jacobian (wave.py:280-389)
# --------------------------------------------------------------------------
def born_data(grid, frequency_mhz, weights, m_kappa, m_eta, background_fields=None, c0=1540.0, pml_strength=12.0):
    """delta_u = -A0^{-1} (A(m) - A0) u0 sampled at the ports -> [tx, rx] (wave.py:280-329)."""
    a0 = assemble(grid, frequency_mhz, c0=c0, pml_strength=pml_strength)
    da = assemble(grid, frequency_mhz, m_kappa, m_eta, c0=c0, pml_strength=pml_strength) - a0
    lu = splu(a0)
    w = weights.tocsc()
    if background_fields is None:
        u0 = lu.solve(-w.toarray().astype(np.complex128) / grid.h_m ** 2)
    else:
        u0 = np.asarray(background_fields, dtype=np.complex128)
    scattered = lu.solve(-(da @ u0))
    return sample(scattered, w)


def born_jacobian(grid, frequency_mhz, background_fields, source_indices, receiver_indices, basis, c0=1540.0,
                  pml_strength=12.0):
    """Exact derivative of the port data at zero contrast (wave.py:331-389):
    rows [tx, rx] flattened, columns [m_kappa basis | m_eta basis]."""
    fields = np.asarray(background_fields, dtype=np.complex128)
    tx = np.repeat(np.asarray(source_indices), len(receiver_indices))
    rx = np.tile(np.asarray(receiver_indices), len(source_indices))
    basis = sparse.csr_matrix(basis, dtype=np.float64)
    ei, ej, avg = edges(grid)
    k0, mass, stretch_e = coefficients(grid, frequency_mhz, c0, pml_strength)
    edge_basis = avg @ basis
    grads = -fields[ei].copy()
    valid = ej >= 0
    grads[valid] += fields[ej[valid]]
    grads /= grid.h_m
    h2 = grid.h_m ** 2
    mass_b = sparse.diags(k0 ** 2 * mass) @ basis
    flux_b = sparse.diags(-stretch_e) @ edge_basis
    jk = (mass_b.T @ (fields[:, rx] * fields[:, tx])).T * h2
    je = (flux_b.T @ (grads[:, rx] * grads[:, tx])).T * h2
    return np.hstack((jk, je))
