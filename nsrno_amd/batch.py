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
from .helmholtz import HelmholtzPlan, Ports, solve_batch
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


def setup_case(medium, case, angles, restart, n_elements=192, mg_shift=0.4, autotune=True):
    """One frequency's plan on the current stream: medium resampled to the
    case's grid, multigrid, ports and right-hand sides.  Returns (plan,
    ports, rhs, setup) with setup = dict(work, weights, amplitudes)."""
    p = medium.pml
    n = case.interior + 2 * p
    c, rho = medium.sample(case)
    work = media.Workload(n, n, case.h_mm, p, case.frequency_mhz, medium.c0, medium.rho0, c, rho, None)
    pos, nrm, wid = media.line_array(work, n_elements)
    weights = Solver(Grid(work.x_mm, work.z_mm, work.pml_mm), c0_m_s=work.c0).source_weights(pos, nrm, wid)
    amp = media.plane_wave_amplitudes(pos, angles, case.frequency_mhz, work.c0)
    plan = HelmholtzPlan(n, n, case.h_mm, case.frequency_mhz, work.pml_mm, c0_m_s=work.c0, nrhs=len(angles),
                         restart=restart)
    ports = Ports(weights)
    plan.set_medium_physical(torch.from_numpy(c).cuda(), torch.from_numpy(rho).cuda(), None, rho0=work.rho0)
    plan.setup_multigrid(levels=_mg_levels(n), shift=mg_shift, pre=2, post=2, coarse_sweeps=10)
    if autotune:
        case.tuning = plan.autotune(len(angles))
    rhs = plan.inject(ports, torch.from_numpy(amp).cuda())
    return plan, ports, rhs, dict(work=work, weights=weights, amplitudes=amp)


def finish_case(case, plan, ports, x, info, restart, keep_fields=False, setup=None):
    """Traces, convergence summary and algorithmic bytes of a solved case."""
    n = plan.nz * plan.nx
    traces = plan.sample(ports, x)
    case.nodes = n
    case.iterations = [i.iterations for i in info]
    case.relative_residual = max(i.relative_residual for i in info)
    case.reasons = tuple(sorted(set(i.reason for i in info)))
    case.traces = traces.cpu().numpy()
    case.graph_fallbacks = max(i.graph_fallbacks for i in info)
    if keep_fields:
        case.fields = x.cpu().numpy()
        case.setup = setup
    # algorithmic Arnoldi + operator bytes (DCGS2 step j: 8(j+2) + 8(j+4); stencil 16 + 12/R per node per RHS)
    R = len(info)
    steps = max(case.iterations)
    js = np.arange(steps) % restart
    case.krylov_bytes = float(np.sum(16.0 * js + 48.0 + 16.0 + 12.0 / R) * n * R)
    return case


def solve_case(medium, case, angles, tol, restart, n_elements=192, stream=None, autotune=True, graph=False,
               mg_shift=0.4, keep_fields=False, plans=None):
    """One frequency: all angles as one batched multigrid-FGMRES solve.

    plans: a list that receives the plan instead of closing it here (run_sweep
    closes them after every worker has finished, so that no plan's memory is
    freed while another thread captures a restart cycle)."""
    t0 = time.perf_counter()
    plan, ports, rhs, setup = setup_case(medium, case, angles, restart, n_elements, mg_shift, autotune)
    x, info = plan.solve(rhs, tolerance=tol, restart=restart, max_iterations=6000, precond="mg", graph=graph)
    finish_case(case, plan, ports, x, info, restart, keep_fields, setup)
    if stream is not None:
        stream.synchronize()
    case.seconds = time.perf_counter() - t0
    if plans is None:
        plan.close()
    else:
        plans.append(plan)
    return case


def run_sweep(frequencies, angles, tol=1e-6, restart=3, workers=4, seed=0, log=None, autotune=True, graph=False,
              mg_shift=0.4, keep_fields=(), mode="threads", timing=None):
    """Solve every frequency.  mode "batched": all frequencies in one
    helm_batch_solve (every kernel of a restart cycle one segmented launch
    over the frequencies still running; set-up on `workers` host threads);
    mode "threads": each frequency its own solve, `workers` of them in
    flight on concurrent streams (largest grids first).  Returns (cases in
    input order, wall seconds).  keep_fields: indices of the frequencies
    whose fields are copied to the host.  The plans stay allocated until
    every frequency is solved (a few GB for a 64-frequency sweep) and are
    closed after the timed region.  timing (a dict, batched mode): receives
    the set-up and solve seconds."""
    if mode == "batched":
        return run_sweep_batched(frequencies, angles, tol, restart, workers, seed, log, graph, mg_shift, keep_fields,
                                 timing)
    if mode != "threads":
        raise ValueError(f"unknown sweep mode {mode!r}")
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


def run_sweep_batched(frequencies, angles, tol=1e-6, restart=3, workers=4, seed=0, log=None, graph=True,
                      mg_shift=0.4, keep_fields=(), timing=None):
    """run_sweep(mode="batched"): set-up of every frequency (medium
    resampling, plan, multigrid, ports; `workers` host threads, each on its
    own stream), then ONE helm_batch_solve over all of them, then traces.
    The launch shapes are the library defaults (one segmented launch serves
    every grid, so there is no per-frequency autotuning).  Each case's
    seconds is its set-up time; the wall covers set-up, solve and traces."""
    medium = PhysicalMedium(seed=seed)
    cases = [medium.case(f) for f in frequencies]
    order = sorted(range(len(cases)), key=lambda k: -cases[k].interior)
    built = [None] * len(cases)
    errors = []
    work = queue.Queue()
    for k in order:
        work.put(k)

    def worker():
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            while True:
                try:
                    k = work.get_nowait()
                except queue.Empty:
                    break
                try:
                    t0 = time.perf_counter()
                    built[k] = setup_case(medium, cases[k], angles, restart, mg_shift=mg_shift, autotune=False)
                    cases[k].seconds = time.perf_counter() - t0
                except Exception as e:  # surfaced after the join
                    errors.append(e)
                    break
            stream.synchronize()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker) for _ in range(max(1, workers))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    try:
        if errors:
            raise errors[0]
        t_setup = time.perf_counter() - t0
        plans = [b[0] for b in built]
        xs, infos = solve_batch(plans, [b[2] for b in built], tolerance=tol, restart=restart, max_iterations=6000,
                                graph=graph)
        keep = set(keep_fields)
        for k, (plan, ports, rhs, setup) in enumerate(built):
            finish_case(cases[k], plan, ports, xs[k], infos[k], restart, k in keep, setup)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if timing is not None:
            timing.update(setup_seconds=t_setup, solve_seconds=wall - t_setup)
        if log:
            log(f"batched sweep: set-up {t_setup:.2f} s, solve + traces {wall - t_setup:.2f} s, "
                f"steps {max(max(c.iterations) for c in cases)}")
    finally:
        for b in built:
            if b is not None:
                b[0].close()
    return cases, wall
