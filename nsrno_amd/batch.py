"""Batched frequency sweep: many (frequency, angle) solves of uneven grid sizes.

The reference solves one frequency at a time (reconstruct_dense.py:91 shards
frequencies over processes; solve_system batches the angles of one
frequency, dense_wave.py:167-225).  Here every frequency's angles are one
batched solve (right-hand sides sharing the coefficients), and several
frequencies are in flight at once on separate HIP streams, each driven by its
own host thread and plan, so that small grids overlap and keep the GPU busy.
The medium is one physical map resampled onto each frequency's grid, which is
sized for a fixed number of points per wavelength.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch
from scipy.ndimage import map_coordinates

from . import media
from .helmholtz import HelmholtzPlan, Ports
from .theranostic import Grid, Solver


@dataclass
class FrequencyCase:
    frequency_mhz: float
    interior: int
    h_mm: float
    nodes: int = 0
    iterations: list = field(default_factory=list)
    relative_residual: float = 0.0
    reasons: tuple = ()
    seconds: float = 0.0
    traces: np.ndarray | None = None
    krylov_bytes: float = 0.0
    tuning: dict | None = None
    graph_fallbacks: int = 0    # restart cycles run directly because their graph capture failed
    fields: np.ndarray | None = None  # complex128 [R, n, n] when asked for (keep_fields)
    setup: dict | None = None   # grid, medium and acquisition of the solve (keep_fields)


class PhysicalMedium:
    """An inclusion medium on a fixed physical square, sampled on any grid.

    Built by ``media.inclusion_medium`` at a reference frequency and interior
    size; ``sample`` interpolates c and rho bilinearly at another spacing
    (water in the PML)."""

    def __init__(self, seed=0, f_ref_mhz=4.0, interior_ref=384, pml=32, c_min=1440.0):
        w = media.inclusion_medium(seed=seed, interior=interior_ref, pml=pml, frequency_mhz=f_ref_mhz)
        self.c0, self.rho0 = w.c0, w.rho0
        self.h_ref = w.h_mm
        self.size_mm = interior_ref * w.h_mm
        self.c = w.c[pml:-pml, pml:-pml]
        self.rho = w.rho[pml:-pml, pml:-pml]
        self.pml = pml
        self.c_min = c_min

    def case(self, frequency_mhz, ppw=10):
        h = media.spacing_mm(self.c_min, frequency_mhz, ppw)
        return FrequencyCase(frequency_mhz, int(round(self.size_mm / h)), h)

    def sample(self, case):
        n, p = case.interior, self.pml
        # interior node i sits at (i + 0.5) * h on [0, size]; the reference map likewise at h_ref
        s = ((np.arange(n) + 0.5) * case.h_mm) / self.h_ref - 0.5
        zz, xx = np.meshgrid(s, s, indexing="ij")
        out = []
        for a, bg in ((self.c, self.c0), (self.rho, self.rho0)):
            v = map_coordinates(a, [zz, xx], order=1, mode="nearest")
            full = np.full((n + 2 * p, n + 2 * p), bg)
            full[p:p + n, p:p + n] = v
            out.append(full)
        return out


def _mg_levels(n, coarsest=40):
    levels = 1
    while n > coarsest and levels < 7:
        n = (n + 1) // 2
        levels += 1
    return levels


def solve_case(medium, case, angles, tol, restart, n_elements=192, stream=None, autotune=True, graph=False,
               mg_shift=0.4, keep_fields=False, plans=None):
    """One frequency: all angles as one batched multigrid-FGMRES solve.

    plans: a list that receives the plan instead of closing it here (run_sweep
    closes them after every worker has finished, so that no plan's memory is
    freed while another thread captures a restart cycle)."""
    p = medium.pml
    n = case.interior + 2 * p
    c, rho = medium.sample(case)
    work = media.Workload(n, n, case.h_mm, p, case.frequency_mhz, medium.c0, medium.rho0, c, rho, None)
    pos, nrm, wid = media.line_array(work, n_elements)
    weights = Solver(Grid(work.x_mm, work.z_mm, work.pml_mm), c0_m_s=work.c0).source_weights(pos, nrm, wid)
    amp = media.plane_wave_amplitudes(pos, angles, case.frequency_mhz, work.c0)
    t0 = time.perf_counter()
    plan = HelmholtzPlan(n, n, case.h_mm, case.frequency_mhz, work.pml_mm, c0_m_s=work.c0, nrhs=len(angles),
                         restart=restart)
    ports = Ports(weights)
    plan.set_medium_physical(torch.from_numpy(c).cuda(), torch.from_numpy(rho).cuda(), None, rho0=work.rho0)
    plan.setup_multigrid(levels=_mg_levels(n), shift=mg_shift, pre=2, post=2, coarse_sweeps=10)
    if autotune:
        case.tuning = plan.autotune(len(angles))
    rhs = plan.inject(ports, torch.from_numpy(amp).cuda())
    x, info = plan.solve(rhs, tolerance=tol, restart=restart, max_iterations=6000, precond="mg", graph=graph)
    traces = plan.sample(ports, x)
    if stream is not None:
        stream.synchronize()
    case.seconds = time.perf_counter() - t0
    case.nodes = n * n
    case.iterations = [i.iterations for i in info]
    case.relative_residual = max(i.relative_residual for i in info)
    case.reasons = tuple(sorted(set(i.reason for i in info)))
    case.traces = traces.cpu().numpy()
    case.graph_fallbacks = max(i.graph_fallbacks for i in info)
    if keep_fields:
        case.fields = x.cpu().numpy()
        case.setup = dict(work=work, weights=weights, amplitudes=amp)
    # algorithmic Arnoldi + operator bytes (DCGS2 step j: 8(j+2) + 8(j+4); stencil 16 + 12/R per node per RHS)
    R = len(angles)
    steps = max(case.iterations)
    js = np.arange(steps) % restart
    case.krylov_bytes = float(np.sum(16.0 * js + 48.0 + 16.0 + 12.0 / R) * n * n * R)
    if plans is None:
        plan.close()
    else:
        plans.append(plan)
    return case


def run_sweep(frequencies, angles, tol=1e-6, restart=3, workers=4, seed=0, log=None, autotune=True, graph=False,
              mg_shift=0.4, keep_fields=()):
    """Solve every frequency (largest grids first) on `workers` concurrent
    streams.  Returns (cases in input order, wall seconds).  keep_fields:
    indices of the frequencies whose fields are copied to the host.  The
    plans stay allocated until every worker has finished (a few GB for a
    64-frequency sweep) and are closed after the timed region."""
    medium = PhysicalMedium(seed=seed)
    cases = [medium.case(f) for f in frequencies]
    order = sorted(range(len(cases)), key=lambda k: -cases[k].interior)
    work = queue.Queue()
    for k in order:
        work.put(k)
    errors = []
    plans = []
    keep = set(keep_fields)

    def worker():
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            while True:
                try:
                    k = work.get_nowait()
                except queue.Empty:
                    return
                try:
                    solve_case(medium, cases[k], angles, tol, restart, stream=stream, autotune=autotune, graph=graph,
                               mg_shift=mg_shift, keep_fields=k in keep, plans=plans)
                    if log:
                        c = cases[k]
                        log(f"{c.frequency_mhz:.3f} MHz {c.interior}^2 steps {max(c.iterations)} "
                            f"{c.seconds * 1e3:.0f} ms")
                except Exception as e:  # surfaced after the join
                    errors.append(e)
                    return

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker) for _ in range(workers)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for p in plans:
        p.close()
    if errors:
        raise errors[0]
    return cases, wall
