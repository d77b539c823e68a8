"""Benchmark: batched frequency-domain Helmholtz solves on one MI355X.

Workload (default): 512^2 interior + 32-cell stretched PML (576^2 nodes),
3 MHz, dx = lambda_min/10, seeded smooth random medium (c, rho, power-law
alpha), 11 plane-wave angles in [-15, +15] deg from a 192-element line array,
192 receivers; complex64 flexible GMRES(6) with FP64 inner products and
FP64 true-residual refinement at every restart to relative residual 1e-6 on
every angle; multigrid shift 0.4.

"learned FGMRES" (the timed leg): a warm start of 4 minimal-residual steps
along the learned directions MG(r) + N(r) (the BF16 U-Net trained by
nsrno_amd/train_preconditioner.py, weights in nsrno_amd/data/), then
multigrid-preconditioned FGMRES.  Multigrid FGMRES without the net and plain
GMRES(30) (no preconditioner) are timed beside it.

A "step" = one full batched solve of the 11 angles from a zero start:
coefficients + multigrid set-up on the device, source injection, warm start,
the Krylov solve to 1e-6, receiver sampling.  Inputs (c, rho, alpha maps) are resident
in HBM before the timed region.  ``value`` = solves/s = 11 * K / seconds.

Prints ONE JSON line (rank 0, N=1).  --gpus N>1 exits with a message: this
build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

# profiled single kernels of the headline solve -> kernel names (the
# "preconditioner" entry is the whole V-cycle, a group, so it is not one)
names = {"dcgs_dots": "k_dcgs_dots", "dcgs_update": "k_dcgs_reduce_update", "operator_apply": "k_apply_block<float>",
         "true_residual": "k_apply<double> (residual)", "mg_pre2_restrict": "k_mgt_pre2_rr",
         "mg_prolong_post2": "k_mgt_post2"}
# the multigrid kernels are timed on the finest level only inside the timed
# steps (one event pair per V-cycle instead of one per level)
finest = {"mg_pre2_restrict": "mg_pre2_restrict_finest", "mg_prolong_post2": "mg_prolong_post2_finest"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precond", default="mg", choices=["none", "mg", "learned"])
    ap.add_argument("--restart", type=int, default=3,
                    help="FGMRES restart length (short cycles: each restart is an FP64 refinement step)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iterations", type=int, default=6000)
    ap.add_argument("--graph", type=int, default=1, help="1: replay restart cycles as captured HIP graphs")
    ap.add_argument("--autotune", type=int, default=1, help="1: time launch shapes on the plan first (helm_autotune)")
    ap.add_argument("--warm", default="learned", choices=["learned", "mg", "none"],
                    help="warm-start directions before the FGMRES solve")
    ap.add_argument("--warm-steps", type=int, default=4)
    ap.add_argument("--learned-params", default=os.path.join(ROOT, "nsrno_amd", "data",
                                                             "learned_warm_576x576_3mhz.npz"))
    ap.add_argument("--plain-max-iterations", type=int, default=6000,
                    help="iteration cap of the plain (unpreconditioned) GMRES comparison leg")
    ap.add_argument("--plain-restart", type=int, default=30,
                    help="restart length of the plain GMRES comparison leg (GMRES(30))")
    ap.add_argument("--interior", type=int, default=512)
    ap.add_argument("--pml", type=int, default=32)
    ap.add_argument("--freq", type=float, default=3.0)
    ap.add_argument("--angles", type=int, default=11)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mg-levels", type=int, default=5)
    ap.add_argument("--mg-shift", type=float, default=0.4)
    ap.add_argument("--mg-omega", type=float, default=0.0)
    ap.add_argument("--mg-pre", type=int, default=2)
    ap.add_argument("--mg-post", type=int, default=2)
    ap.add_argument("--mg-coarse", type=int, default=10)
    ap.add_argument("--roofline-kernel", default="auto",
                    choices=["auto", "dcgs_dots", "dcgs_update", "operator_apply", "true_residual",
                             "mg_pre2_restrict", "mg_prolong_post2"],
                    help="kernel timed with HIP events inside the timed solves for the roofline object "
                         "(auto: the single kernel with the most time in an all-kernel profiled solve)")
    ap.add_argument("--cond-shape", type=int, nargs=2, default=[960, 1216],
                    help="native medium of the conditioned workload (box 1024 x 1280)")
    ap.add_argument("--cond-tol", type=float, default=1e-5)
    ap.add_argument("--cond-restart", type=int, default=6, help="FGMRES restart length of the conditioned workload")
    ap.add_argument("--workload", default="angles", choices=["angles", "frequencies", "conditioned", "spectral", "net", "learned", "aux"],
                    help="angles: the 576^2 x 11-angle batch (default); frequencies: the batched sweep of "
                         "64 frequencies x 11 angles on uneven grids; conditioned: the split-PML system; "
                         "spectral: the S operator on a 1056x1216 box; net: the learned U-Net alone; "
                         "learned: config 5 (learned vs multigrid vs plain on held-out media, 3-9 MHz); "
                         "aux: capture, band RF and Born Jacobian kernels")
    ap.add_argument("--media", type=int, default=3, help="held-out media per frequency (learned workload)")
    ap.add_argument("--frequencies", type=int, default=64)
    ap.add_argument("--spectral-restart", type=int, default=10,
                    help="FGMRES restart length of the spectral workload (the reference's Fourier preconditioner)")
    ap.add_argument("--workers", type=int, default=4,
                    help="frequencies in flight, or set-up threads with --sweep-mode batched (frequencies workload)")
    ap.add_argument("--sweep-mode", default="batched", choices=["batched", "threads"],
                    help="frequencies workload: one segmented batched solve, or one solve per frequency on streams")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of known size (PMC counter calibration) and exit")
    ap.add_argument("--pmc-summary", metavar="DIR",
                    help="summarise rocprofv3 --pmc runs under DIR (fetch_bench, write_bench, fetch_copy, "
                         "write_copy) into profiles/pmc_traffic.json and exit")
    ap.add_argument("--pmc-labels", default="",
                    help="with --pmc-summary: comma-separated kernel labels to update (default: every kernel seen)")
    ap.add_argument("--pmc-tag", default="",
                    help="with --pmc-summary: stored as '<label> [TAG]' (a kernel measured on another workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-GMRES comparison leg")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched frequency-sweep leg")
    ap.add_argument("--cpu-interior", type=int, default=None,
                    help="interior size of the CPU baseline sample (default: the full workload)")
    return ap.parse_args()


class Bench:
    def __init__(self, a):
        import numpy as np
        import torch
        from nsrno_amd import media
        from nsrno_amd.helmholtz import HelmholtzPlan, Ports
        from nsrno_amd.theranostic import Grid, Solver

        self.a = a
        self.torch = torch
        self.work = w = media.smooth_random_medium(seed=a.seed, interior=a.interior, pml=a.pml,
                                                   frequency_mhz=a.freq)
        self.angles = np.linspace(-15.0, 15.0, a.angles)
        pos, nrm, wid = media.line_array(w, 192)
        grid = Grid(w.x_mm, w.z_mm, w.pml_mm)
        self.weights = Solver(grid, c0_m_s=w.c0).source_weights(pos, nrm, wid)
        self.amp = media.plane_wave_amplitudes(pos, self.angles, w.frequency_mhz, w.c0)
        dev = torch.device("cuda")
        self.ports = Ports(self.weights, device="cuda")
        self.amp_d = torch.from_numpy(self.amp).to(dev)
        self.c = torch.from_numpy(w.c).to(dev)
        self.rho = torch.from_numpy(w.rho).to(dev)
        self.alpha = torch.from_numpy(w.alpha_np_m).to(dev)
        self.plan = HelmholtzPlan(w.nz, w.nx, w.h_mm, w.frequency_mhz, w.pml_mm, c0_m_s=w.c0,
                                  nrhs=a.angles, restart=max(a.restart, a.plain_restart))
        self.R = a.angles
        self.n = w.nz * w.nx
        if a.warm == "learned":
            from nsrno_amd.learned import LearnedNet, attach, canvas
            self.net = LearnedNet(canvas(w.nz), canvas(w.nx), a.angles,
                                  params=LearnedNet.load_params(a.learned_params),
                                  linear=LearnedNet.load_linear(a.learned_params))
            attach(self.plan, self.net)

    def step(self, precond, max_iterations, warm="none"):
        a, p = self.a, self.plan
        restart = a.restart if precond != "none" else a.plain_restart
        p.set_medium_physical(self.c, self.rho, self.alpha, rho0=self.work.rho0)
        if precond != "none":
            p.setup_multigrid(levels=a.mg_levels, shift=a.mg_shift, pre=a.mg_pre, post=a.mg_post,
                              coarse_sweeps=a.mg_coarse, omega=a.mg_omega)
        rhs = p.inject(self.ports, self.amp_d)
        x0 = None
        if warm != "none":
            x0 = self.torch.zeros_like(rhs)
            p.warm_start(rhs, x0, a.warm_steps, precond=warm)
        x, infos = p.solve(rhs, x0=x0, tolerance=a.tol, restart=restart, max_iterations=max_iterations,
                           precond=precond, graph=bool(a.graph))
        traces = p.sample(self.ports, x)
        self.warm_applications = 2 * a.warm_steps if warm != "none" else 0
        return x, infos, traces

    def timed(self, precond, steps, warmup, max_iterations, profile=None, warm="none"):
        """Time `steps` solves; `profile` names kernels whose launches the library
        times with HIP events over the same timed steps (helm_profile)."""
        torch = self.torch
        for _ in range(warmup):
            self.step(precond, max_iterations, warm)
        torch.cuda.synchronize()
        if profile:
            self.plan.profile(profile)
        t0 = time.perf_counter()
        for _ in range(steps):
            x, infos, traces = self.step(precond, max_iterations, warm)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        self.kernel_stats = self.plan.profile_read() if profile else None
        if profile:
            self.plan.profile(enable=False)
        return dt, x, infos, traces

    def stencil_roofline(self, reps=48):
        """Average duration of the FP32 stencil kernel on the bench's batch,
        timed with HIP events on the stream it is launched on.  Cold: the
        launches rotate through enough field pairs (>= 512 MB) that no
        input is still in the 256 MB last-level cache, as in the solve.
        Warm: the same pair again and again (inputs served by that cache)."""
        torch = self.torch
        p = self.plan
        p.set_medium_physical(self.c, self.rho, self.alpha, rho0=self.work.rho0)
        shape = (self.R, self.work.nz, self.work.nx)
        nbuf = max(2, min(16, (512 << 20) // (2 * 8 * self.R * self.n) + 1))
        us = [torch.randn(shape, dtype=torch.complex64, device="cuda") for _ in range(nbuf)]
        outs = [torch.empty_like(us[0]) for _ in range(nbuf)]
        s = torch.cuda.current_stream()

        def timed(pairs):
            for k in range(5):
                p.apply(pairs[k % len(pairs)][0], out=pairs[k % len(pairs)][1])
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for k in range(reps):
                p.apply(pairs[k % len(pairs)][0], out=pairs[k % len(pairs)][1])
            e1.record(s)
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / 1e3 / reps

        t_cold = timed(list(zip(us, outs)))
        t_warm = timed([(us[0], outs[0])])
        bytes_alg = (16.0 + 12.0 / self.R) * self.n * self.R
        return t_cold, t_warm, bytes_alg


def stencil_kernel_name(tuning):
    """The FP32 flux stencil's kernel and launch shape as the autotuner left them."""
    if not tuning:
        return "k_apply_stream<4> (FP32 stencil, default 8-row runs)"
    rows, rhs, blk = tuning["stencil_rows"], tuning["stencil_rhs"], tuning.get("stencil_block", 0)
    if blk:
        order = {2: ", XCD-contiguous order", 4: ", XCD-contiguous order",
                 5: ", XCD-contiguous order, RHS groups of a run together",
                 6: ", XCD-contiguous order, RHS groups of a run together"}.get(blk, "")
        if blk in (3, 4, 6):
            return f"k_apply_pair<{rhs},{rows}> (FP32 stencil, two columns per lane, {rows}-row runs{order})"
        return f"k_apply_block<{rhs},{rows}> (FP32 stencil, {rows}-row runs loaded up front{order})"
    return f"k_apply_stream<{rhs}> (FP32 stencil, {rows}-row runs)"


COPY_BYTES = 1 << 28  # bytes read (and written) by one calibration copy
COPY_REPS = 5


def copy_calibration():
    """A device-to-device copy of COPY_BYTES, COPY_REPS times: the PMC counters
    of these dispatches give the counter-to-bytes factor."""
    import torch
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(COPY_REPS):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[bench] copy calibration: {COPY_REPS} copies of {COPY_BYTES} B", flush=True)


def _pmc_rows(directory, counter):
    """(kernel name, counter value) per dispatch from a rocprofv3 --pmc run."""
    import csv
    import glob
    out = []
    for path in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    out.append((r["Kernel_Name"], float(r["Counter_Value"]), int(r.get("Grid_Size") or 0)))
    return out


def pmc_summary(directory, labels=None, tag=""):
    """FETCH_SIZE/WRITE_SIZE (KiB per dispatch) of the bench kernels, each
    scaled by the factor that maps the same counter to the known bytes of the
    calibration copy -> profiles/pmc_traffic.json (only the entries named in
    `labels` when given, so that a run of another workload updates its own
    kernels and leaves the rest)."""
    import collections
    scale = {}
    for counter, sub in (("FETCH_SIZE", "fetch_copy"), ("WRITE_SIZE", "write_copy")):
        rows = _pmc_rows(os.path.join(directory, sub), counter)
        vals = sorted(v for _, v, _ in rows)
        if not vals:
            raise SystemExit(f"no {counter} rows under {directory}/{sub}")
        big = vals[-COPY_REPS:]  # the calibration copies are the largest dispatches
        scale[counter] = COPY_BYTES / (sum(big) / len(big))
    # matched in this order (a name that contains another comes first)
    short = {"k_dcgs_dots": "k_dcgs_dots", "k_dcgs_reduce_update": "k_dcgs_reduce_update",
             "k_dcgs_norm_scale": "k_dcgs_norm_scale", "k_dcgs_update": "k_dcgs_update", "k_tail_lds": "k_tail_lds",
             "k_apply<float": "k_apply<float>",
             "k_apply_stream": "k_apply_stream<float>", "k_apply_block": "k_apply_block<float>",
             "k_apply_pair": "k_apply_pair<float>",
             "k_mgt_pre2_rr": "k_mgt_pre2_rr", "k_mgt_post2": "k_mgt_post2", "k_mgt_jacobi": "k_mgt_jacobi",
             "k_mgt_pre2": "k_mgt_pre2", "k_mgt_prolong_smooth": "k_mgt_prolong_smooth",
             "k_mgt_restrict_res": "k_mgt_restrict_res", "k_tail(": "k_tail", "k_cond8s<float": "k_cond8s<float>",
             "k_stag8<float": "k_stag8<float>"}
    totals = collections.defaultdict(lambda: {"FETCH_SIZE": [0.0, 0], "WRITE_SIZE": [0.0, 0]})
    # multigrid kernels run on every level: their finest-level launches (the
    # largest grid of the kernel, over all its template instances: the finest
    # level forms its stencils from the medium, the others read a table) are
    # also summarised on their own ("<name> (finest level)")
    def label_of(name):
        return next((label for key, label in short.items() if key in name), None)

    biggest = collections.defaultdict(int)
    rows_bench = {c: _pmc_rows(os.path.join(directory, sub), c)
                  for c, sub in (("FETCH_SIZE", "fetch_bench"), ("WRITE_SIZE", "write_bench"))}
    for name, _, grid in rows_bench["FETCH_SIZE"]:
        biggest[label_of(name)] = max(biggest[label_of(name)], grid)
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        for name, v, grid in rows_bench[counter]:
            for key, label in short.items():
                if key in name:
                    for lab in ([label, label + " (finest level)"] if key.startswith("k_mgt") and grid == biggest[label]
                                else [label]):
                        t = totals[lab][counter]
                        t[0] += v * scale[counter]
                        t[1] += 1
                    break
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:  # kernels measured by earlier runs (other workloads) are kept
        with open(path) as fh:
            kept = json.load(fh).get("kernels", {})
    except (OSError, ValueError):
        kept = {}
    out = {"unit": "bytes per launch (mean over the launches of one solve of the workload that ran the kernel)",
           "scale_fetch_bytes_per_count": scale["FETCH_SIZE"], "scale_write_bytes_per_count": scale["WRITE_SIZE"],
           "kernels": kept}
    for label, t in totals.items():
        if labels and label not in labels:
            continue
        f, w = t["FETCH_SIZE"], t["WRITE_SIZE"]
        if f[1] and w[1]:
            fb, wb = f[0] / f[1], w[0] / w[1]
            out["kernels"][label + (f" [{tag}]" if tag else "")] = {"launches": f[1], "fetch_bytes_per_launch": round(fb),
                                     "write_bytes_per_launch": round(wb), "hbm_bytes_per_launch": round(fb + wb)}
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out), flush=True)


SWEEP_PMC = os.path.join(ROOT, "profiles", "pmc_frequency_sweep.json")


def sweep_hbm(a, n_freq, n_angles, timing):
    """HBM bytes the batched sweep moves (PMC FETCH_SIZE + WRITE_SIZE of its
    segmented kernels, measured once on the default workload and kept in
    profiles/pmc_frequency_sweep.json; the solve is deterministic, so the
    bytes are those of every run of the same workload) over this run's solve
    seconds.  None for another workload or mode."""
    if a.sweep_mode != "batched" or not timing or not os.path.exists(SWEEP_PMC):
        return None
    with open(SWEEP_PMC) as fh:
        pmc = json.load(fh)
    w = pmc["workload"]
    if (w["frequencies"], w["angles"], w["tolerance"], w["restart"]) != (n_freq, n_angles, a.tol, a.restart):
        return None
    return {"bytes": pmc["segmented_kernel_bytes"], "solve_seconds": round(timing["solve_seconds"], 3),
            "gb_s": round(pmc["segmented_kernel_bytes"] / timing["solve_seconds"] / 1e9, 1),
            "frac": round(pmc["segmented_kernel_bytes"] / timing["solve_seconds"] / 1e9 / PEAK_HBM_GBS, 4),
            "source": "profiles/frequency_sweep_batched_64x11.md"}


def sweep_summary(a):
    """The batched 64-frequency x 11-angle sweep (config 3) as a compact
    object for the default bench line."""
    import numpy as np
    from nsrno_amd.batch import run_sweep
    freqs = np.linspace(4.0, 8.9, a.frequencies)
    angles = np.linspace(-15.0, 15.0, a.angles)
    run_sweep(freqs[:2], angles, tol=a.tol, restart=a.restart, workers=1, mg_shift=a.mg_shift, mode="batched")
    timing = {}
    cases, wall = run_sweep(freqs, angles, tol=a.tol, restart=a.restart, workers=a.workers, graph=True,
                            mg_shift=a.mg_shift, mode="batched", timing=timing)
    sweep_args = argparse.Namespace(**{**vars(a), "sweep_mode": "batched"})
    return {"workload": f"{len(freqs)} frequencies 4.0-8.9 MHz x {len(angles)} angles, interiors "
                        f"{cases[0].interior}^2-{cases[-1].interior}^2 + 32 PML, one helm_batch_solve",
            "solves_per_s": round(len(freqs) * len(angles) / wall, 3), "seconds": round(wall, 3),
            "setup_seconds": round(timing["setup_seconds"], 3), "solve_seconds": round(timing["solve_seconds"], 3),
            "hbm": sweep_hbm(sweep_args, len(freqs), len(angles), timing),
            "all_converged": all(c.reasons == ("converged",) for c in cases),
            "max_relative_residual": max(c.relative_residual for c in cases),
            "steps_range": [min(max(c.iterations) for c in cases), max(max(c.iterations) for c in cases)]}


def frequency_sweep(a):
    """Batched uneven solves: `a.frequencies` frequencies evenly spaced in
    4.0-8.9 MHz x 11 angles; each grid sized for 10 points per wavelength at
    1440 m/s (interior 384^2 at 4 MHz to 854^2 at 8.9 MHz, + 32-cell PML) on
    one inclusion medium resampled.  --sweep-mode batched (default): one
    helm_batch_solve over all frequencies; threads: `a.workers` frequencies
    in flight on HIP streams.  One JSON line."""
    import numpy as np
    from nsrno_amd.batch import run_sweep
    freqs = np.linspace(4.0, 8.9, a.frequencies)
    angles = np.linspace(-15.0, 15.0, a.angles)
    batched = a.sweep_mode == "batched"
    for _ in range(a.warmup):  # small frequencies warm the kernels and allocator
        run_sweep(freqs[:2], angles, tol=a.tol, restart=a.restart, workers=1, mg_shift=a.mg_shift,
                  mode=a.sweep_mode)
    timing = {}
    cases, wall = run_sweep(freqs, angles, tol=a.tol, restart=a.restart, workers=a.workers,
                            log=lambda m: print("[bench] " + m, flush=True), autotune=bool(a.autotune) and not batched,
                            graph=bool(a.graph), mg_shift=a.mg_shift, mode=a.sweep_mode, timing=timing)
    solves = len(freqs) * len(angles)
    kbytes = sum(c.krylov_bytes for c in cases)
    line = {
        "metric": f"solves/s, {len(freqs)} frequencies (4.0-8.9 MHz) x {len(angles)} angles, uneven grids",
        "value": round(solves / wall, 3), "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (FP64 reductions/refinement)", "data": "synthetic (seeded inclusion medium, resampled)",
        "config": {"workload": "frequency sweep", "frequencies": len(freqs), "angles": len(angles),
                   "interiors": [cases[0].interior, cases[-1].interior], "pml": 32, "ppw": 10,
                   "tolerance": a.tol, "method": f"fgmres({a.restart}) + multigrid",
                   "mode": ("one helm_batch_solve: every kernel of a restart cycle one segmented launch over "
                            "the frequencies still running, set-up on %d host threads" % a.workers) if batched
                           else f"{a.workers} frequencies in flight on {a.workers} streams",
                   "autotune": bool(a.autotune) and not batched, "graph": bool(a.graph),
                   "parallelism": "single GPU"},
        "krylov_operator_gb_s": round(kbytes / wall / 1e9, 1),
        "hbm": sweep_hbm(a, len(freqs), len(angles), timing),
        "all_converged": all(c.reasons == ("converged",) for c in cases),
        "max_relative_residual": max(c.relative_residual for c in cases),
        "steps_range": [min(max(c.iterations) for c in cases), max(max(c.iterations) for c in cases)],
    }
    print(json.dumps(line), flush=True)


def conditioned_workload(a):
    """The dense path's conditioned split-PML system (conditioned_wave.py,
    dense_wave.solve_system): native medium 960 x 1216 at dx = 50 um in a
    1024 x 1280 box, 24-cell split PML, dt 2.5 ns, bin 360 (5.997 MHz),
    seeded smooth random c, rho and alpha, 11 plane-wave source lines;
    multigrid FGMRES(8) with FP64 refinement to 1e-5 on every angle.  A step
    = coefficients + multigrid set-up + source term + solve.  One JSON line,
    with the FP32 forward/adjoint operator kernel against the HBM roofline and
    the oracle's Fourier-form operator (the reference's algorithm, 11 FFTs per
    apply, numpy complex128, 1 thread) timed beside it."""
    import numpy as np
    import torch
    from scipy.ndimage import gaussian_filter
    from nsrno_amd.conditioned import ConditionedFrequencySystem, FrequencyGeometry, solve_system, source_rhs
    h, w = a.cond_shape
    dx = 50e-6
    spec = {"medium_shape": [2, h, w], "dx_m": dx, "x0_m": -w * dx / 2, "z0_m": -30 * dx, "native_dt_s": 2.5e-9,
            "pml": 24, "eta_per_unit_alpha": [.015, .025], "taus_s": [12e-9, 35e-9], "source_c_lookup_zx": [30, 24]}
    rng = np.random.default_rng(a.seed)

    def smooth():
        f = gaussian_filter(rng.normal(size=(h, w)), 8.0)
        return f / np.abs(f).max()

    c = 1540.0 + 40.0 * smooth()
    rho = 1000.0 + 50.0 * smooth()
    alpha = 0.5 + 0.2 * smooth()
    f = 360 * 40e6 / 2401
    R = a.angles
    g = FrequencyGeometry(spec)
    medium = torch.tensor(np.stack((c, rho)), dtype=torch.float64, device="cuda")
    alpha_d = torch.tensor(alpha, dtype=torch.float64, device="cuda")
    system = ConditionedFrequencySystem(g, medium, alpha_d, f, nrhs=R, restart=a.cond_restart)
    x = (np.arange(w) - (w - 1) / 2) * dx
    taper = np.hanning(w)
    lines = np.stack([taper * np.exp(-2j * np.pi * f * x * np.sin(np.deg2rad(t)) / 1540.0)
                      for t in np.linspace(-15.0, 15.0, R)])
    lines_d = torch.from_numpy(lines).cuda()
    print(f"[bench] conditioned box {g.H}x{g.W} medium {h}x{w} R={R} f={f / 1e6:.3f} MHz", flush=True)
    tuning = system.autotune(R) if a.autotune else None
    print(f"[bench] autotune {tuning}", flush=True)

    def step():
        system.update(medium, alpha_d)
        b = source_rhs(system, lines_d)
        return solve_system(system, b, tolerance=a.cond_tol, mixed=True, max_cycles=1200, restart=a.cond_restart)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        u, info = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"[bench] conditioned solve {dt / a.steps * 1e3:.1f} ms/step, {info}", flush=True)
    # operator kernel alone: FP32 forward and adjoint over the R fields, on
    # inputs outside the last-level cache (cold, the roofline figure) and on
    # one field pair again and again (warm)
    n = g.H * g.W
    alg = (16.0 + 16.0 / R) * n * R
    kern = {}
    for name, adj in (("forward", False), ("adjoint", True)):
        t, tw = _cold_warm_time(lambda u, o: system.plan.apply(u, adjoint=adj, out=o), (R, g.H, g.W))
        kern[name] = {"avg_launch_us": round(t * 1e6, 2), "gb_s": round(alg / t / 1e9, 1),
                      "frac": round(alg / t / 1e9 / PEAK_HBM_GBS, 4), "warm_us": round(tw * 1e6, 2),
                      "warm_frac": round(alg / tw / 1e9 / PEAK_HBM_GBS, 4)}
    cpu = None
    if not a.no_cpu:
        from oracle import conditioned as oc
        o = oc.System(oc.Geometry(spec), np.stack((c, rho)), alpha, f)
        field = (rng.normal(size=(g.H, g.W)) + 1j * rng.normal(size=(g.H, g.W))) * o.g.mask
        o.apply_fft(field)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < 10.0:
            o.apply_fft(field)
            k += 1
        tc = (time.perf_counter() - t0) / k
        cpu = {"value": round(1.0 / tc, 4), "unit": "operator applies/s (one field)", "cores": 1, "kind": "port",
               "sample": f"{k} applies of the Fourier-form operator (conditioned_wave.py:71-80, 11 FFTs, numpy "
                         f"complex128) on the same {g.H}x{g.W} box, {tc * 1e3:.0f} ms each",
               "gpu_same_unit": round(R / (kern["forward"]["avg_launch_us"] * 1e-6), 1)}
    line = {
        "metric": f"solves/s, conditioned split-PML system {g.H}x{g.W} x {R} angles to rel-res {a.cond_tol:g}",
        "value": round(R * a.steps / dt, 4), "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(dt / a.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64 (FP64 reductions/refinement)",
        "data": "synthetic (seeded smooth random c, rho, alpha)",
        "config": {"workload": f"conditioned {h}x{w} medium, box {g.H}x{g.W}, 24-cell split PML, "
                               f"{f / 1e6:.3f} MHz x{R} angles", "tolerance": a.cond_tol,
                   "method": f"fgmres({a.cond_restart}) + multigrid", "parallelism": "single GPU"},
        "steps_per_solve": info["steps"], "relative_residual": info["residual"],
        "max_angle_residual": info["max_angle_residual"], "reason": info["reason"],
        "roofline": {"bound": "hbm", "kernel": "k_cond8s (FP32 streaming) forward, all RHS",
                     "achieved": kern["forward"]["gb_s"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": kern["forward"]["frac"], "traffic": traffic_from_profile("k_cond8s<float>"),
                     "avg_launch_us": kern["forward"]["avg_launch_us"], "algorithmic_bytes_per_launch": int(alg),
                     "inputs": "cold (rotating field pairs, >= 512 MB, none left in the 256 MB last-level cache)",
                     "warm_us": kern["forward"]["warm_us"], "warm_frac": kern["forward"]["warm_frac"]},
        "adjoint_kernel": kern["adjoint"], "cpu_baseline": cpu, "autotune": tuning,
    }
    print(json.dumps(line), flush=True)
    system.close()


def _cold_warm_time(apply, shape, reps=30):
    """Average seconds per call of apply(u, out) on complex64 fields of
    `shape`, HIP events on torch's current stream.  Cold: the calls rotate
    through enough field pairs (>= 512 MB) that no input is still in the
    256 MB last-level cache, as in a solve; warm: one pair again and again."""
    import torch
    elems = 1
    for d in shape:
        elems *= d
    nbuf = max(2, min(16, (512 << 20) // (2 * 8 * elems) + 1))
    us = [torch.randn(shape, dtype=torch.complex64, device="cuda") for _ in range(nbuf)]
    outs = [torch.empty_like(us[0]) for _ in range(nbuf)]
    s = torch.cuda.current_stream()

    def timed(pairs):
        for k in range(3):
            apply(*pairs[k % len(pairs)])
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for k in range(reps):
            apply(*pairs[k % len(pairs)])
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    cold = timed(list(zip(us, outs)))
    warm = timed([(us[0], outs[0])])
    del us, outs
    return cold, warm


def _event_time(fn, reps):
    """Average seconds of fn() over reps calls, HIP events on torch's current stream."""
    import torch
    for _ in range(3):
        fn()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def spectral_workload(a):
    """The spectral (S) path of solve_frequency (wave_operator.py, wave_solver.py)
    on the reference's own box size: a 961 x 1121 medium at dx = 50 um in a
    1056 x 1216 periodic box with the 48-cell sponge, 5.9975 MHz (bin 360),
    power-law attenuation, seeded smooth random c and rho, 11 plane-wave angles
    in [-8, 8] deg as right-hand sides of one batched solve to rel-res a.tol
    (the reference's default tolerance is 1e-4).  A step = set-up + incident
    fields + right-hand sides + warm start + FGMRES with the reference's
    Fourier preconditioner + receiver traces.  The 8-tap stencil (FP32, all
    11 fields; the streaming kernel k_cond8s with unit split-PML factors) is
    timed alone against the HBM roofline: 16 + 16/R bytes per node per field."""
    import numpy as np
    import torch
    from scipy.ndimage import gaussian_filter
    from nsrno_amd.wave_operator import WaveResidualOperator
    from nsrno_amd.wave_solver import solve_frequency_batch
    h, w = 961, 1121
    dx = 50e-6
    R = a.angles
    spec = dict(medium_shape=[h, w], dx_m=dx, x0_m=-w * dx / 2, z0_m=-3 * dx,
                angles_deg=list(np.linspace(-8.0, 8.0, R)), source_tref_s=[0.0] * R)
    rng = np.random.default_rng(a.seed)

    def smooth():
        f = gaussian_filter(rng.normal(size=(h, w)), 8.0)
        return f / np.abs(f).max()

    medium = torch.tensor(np.stack((1540.0 + 50.0 * smooth(), 1000.0 + 60.0 * smooth()))[None],
                          dtype=torch.float32, device="cuda")
    f = 360 * 40e6 / 2401
    model = WaveResidualOperator(spec, iterations=a.warm_steps, restart=a.spectral_restart, nrhs=R)
    print(f"[bench] spectral box {model.H}x{model.W} medium {h}x{w} R={R} f={f / 1e6:.4f} MHz", flush=True)

    rs = a.spectral_restart

    def step():
        return solve_frequency_batch(model, medium, f, range(R), tolerance=a.tol,
                                     max_iterations=rs * (a.max_iterations // rs), restart=rs)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        preds, x, infos = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    data, _, _ = model.setup(medium, f)
    t, tw = _cold_warm_time(lambda u, o: data.plan.apply(u, adjoint=True, out=o), (R, model.H, model.W))
    n = model.H * model.W
    alg = (16.0 + 16.0 / R) * n * R
    line = {
        "metric": f"solves/s, spectral operator {model.H}x{model.W} box x {R} angles to rel-res {a.tol:g}",
        "value": round(R * a.steps / dt, 4), "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(dt / a.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64 (FP64 reductions/refinement)",
        "data": "synthetic (seeded smooth random c, rho)",
        "config": {"workload": f"spectral {h}x{w} medium, box {model.H}x{model.W}, {f / 1e6:.4f} MHz x{R} angles",
                   "tolerance": a.tol,
                   "method": f"fgmres({a.spectral_restart}) preconditioned by ifft2(multiplier fft2 r) "
                             f"(the reference's solve_frequency preconditioner), warm start x{a.warm_steps}",
                   "parallelism": "single GPU"},
        "solver_iterations": [i["solver_iterations"] for i in infos],
        "relative_residual": max(i["relative_residual"] for i in infos),
        "reasons": sorted(set(i["reason"] for i in infos)),
        "roofline": {"bound": "hbm",
                     "kernel": "k_cond8s<float> (the S operator's 8-tap periodic stencil as the streaming kernel "
                               "with unit split-PML factors, all RHS)",
                     "achieved": round(alg / t / 1e9, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(alg / t / 1e9 / PEAK_HBM_GBS, 4), "traffic": traffic_from_profile("k_cond8s<float> [spectral]"),
                     "avg_launch_us": round(t * 1e6, 2), "algorithmic_bytes_per_launch": int(alg),
                     "inputs": "cold (rotating field pairs, >= 512 MB, none left in the 256 MB last-level cache)",
                     "warm_us": round(tw * 1e6, 2), "warm_frac": round(alg / tw / 1e9 / PEAK_HBM_GBS, 4)},
    }
    print(json.dumps(line), flush=True)


def net_workload(a):
    """The learned preconditioner's BF16 U-Net (helm_net: 12 3x3 convolutions
    of 32 channels over 4 levels, MFMA 32x32x16 bf16, FP32 accumulation) on the
    576^2 canvas x 11 residuals: forward and backward passes timed with HIP
    events; FLOPs counted as 2 * 9 * 32 * 32 per pixel per convolution
    (backward: weight gradients of every layer, data gradients of all but
    the first)."""
    import numpy as np
    import torch
    from nsrno_amd.learned import LearnedNet, canvas
    n = canvas(a.interior + 2 * a.pml)
    B = a.angles
    net = LearnedNet(n, n, B, training=True)
    feats = (torch.randn(B, n, n, 32, device="cuda") * 0.5).to(torch.bfloat16)
    out = net.forward(feats)
    dout = (torch.randn_like(out, dtype=torch.float32) * 1e-2).to(torch.bfloat16)
    tf = _event_time(lambda: net.forward(feats), 20)
    tb = _event_time(lambda: net.backward(dout), 10)
    # pixels per level and convolutions per level (encoder 2, decoder 1, output 1 on level 0)
    px = [n * n, (n // 2) ** 2, (n // 4) ** 2, (n // 8) ** 2]
    convs = [4, 3, 3, 2]  # 12 layers: the layers of each level of the U-Net
    per_px = 2 * 9 * 32 * 32 * B
    flop = sum(p * c for p, c in zip(px, convs)) * per_px
    # backward: weight gradients of all 12 layers, data gradients of all but the first
    flop_b = 2 * flop - px[0] * per_px
    line = {
        "metric": f"U-Net forward+backward TFLOP/s, {n}x{n} x {B}",
        "value": round((flop + flop_b) / (tf + tb) / 1e12, 2),
        "unit": "TFLOP/s", "n_gpus": 1, "steps": 30, "warmup": 3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (fp32 accumulation)", "data": "synthetic (random features)",
        "config": {"workload": f"learned preconditioner U-Net, canvas {n}^2, batch {B}", "parallelism": "single GPU"},
        "forward_ms": round(tf * 1e3, 3), "backward_ms": round(tb * 1e3, 3),
        "roofline": {"bound": "mfma", "kernel": "k_conv_fwd (whole forward pass)",
                     "achieved": round(flop / tf / 1e12, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(flop / tf / 1e12 / 2500.0, 4), "traffic": None},
        "backward_tflop_s": round(flop_b / tb / 1e12, 2),
    }
    print(json.dumps(line), flush=True)
    net.close()


def learned_workload(a):
    """Config 5: learned warm start + multigrid FGMRES against multigrid FGMRES
    alone and plain GMRES(30), on held-out seeded media (seeds 1000 + k; the
    net was trained on seeds 0-199 at 3 MHz) at 3, 6 and 9 MHz on the fixed
    512^2 + 32-PML grid at 10 points per wavelength (dx = lambda_min / 10 at
    each frequency), 11 angles, tol 1e-6.  Per frequency: mean operator
    applications and milliseconds per 11-angle solve of each method (a solve =
    set-up + injection + warm start + solve, timed after one untimed run)."""
    import numpy as np
    import torch
    from nsrno_amd.learned import LearnedNet, attach, canvas
    from nsrno_amd.train_preconditioner import Family, set_medium
    mg = dict(levels=a.mg_levels, shift=a.mg_shift, pre=a.mg_pre, post=a.mg_post, coarse_sweeps=a.mg_coarse)
    freqs = [3.0, 6.0, 9.0]
    seeds = [1000 + k for k in range(a.media)]
    rows = {}
    total = {"learned": [0.0, 0], "mg": [0.0, 0], "plain": [0.0, 0]}
    for f in freqs:
        fam = Family(interior=a.interior, frequency_mhz=f)
        plan = fam.plan(restart=max(a.restart, a.plain_restart))
        net = LearnedNet(canvas(fam.nz), canvas(fam.nx), fam.R, params=LearnedNet.load_params(a.learned_params),
                         linear=LearnedNet.load_linear(a.learned_params))
        attach(plan, net)
        stats = {k: {"applications": [], "ms": [], "converged": True} for k in total}
        for seed in seeds:
            med = fam.medium(seed)
            for kind in ("learned", "mg", "plain"):
                for rep in range(2 if kind != "plain" else 1):
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    set_medium(plan, med, mg)
                    b = plan.inject(fam.ports, fam.amp)
                    x0, warm = None, 0
                    if kind == "learned":
                        x0 = torch.zeros_like(b)
                        plan.warm_start(b, x0, a.warm_steps, precond="learned")
                        warm = 2 * a.warm_steps
                    if kind == "plain":
                        x, info = plan.solve(b, tolerance=a.tol, restart=a.plain_restart,
                                             max_iterations=a.plain_max_iterations, precond="none")
                    else:
                        x, info = plan.solve(b, x0=x0, tolerance=a.tol, restart=a.restart,
                                             max_iterations=a.max_iterations, precond="mg", graph=bool(a.graph))
                    torch.cuda.synchronize()
                    dt = time.perf_counter() - t0
                st = stats[kind]
                st["applications"].append(max(i.applications for i in info) + warm)
                st["ms"].append(dt * 1e3)
                st["converged"] = st["converged"] and all(i.reason == "converged" for i in info)
                total[kind][0] += dt
                total[kind][1] += fam.R
        rows[f"{f:g} MHz"] = {k: {"applications": round(float(np.mean(v["applications"])), 1),
                                  "ms_per_solve": round(float(np.mean(v["ms"])), 2), "converged": v["converged"]}
                              for k, v in stats.items()}
        print(f"[bench] {f} MHz {rows[f'{f:g} MHz']}", flush=True)
        net.close()
        plan.close()
    sps = {k: total[k][1] / total[k][0] for k in total}
    line = {
        "metric": "solves/s to rel-res 1e-6, 512^2 x 11 angles, 3-9 MHz held-out media: learned warm start + "
                  "multigrid FGMRES", "value": round(sps["learned"], 3), "unit": "solves/s", "n_gpus": 1,
        "steps": len(seeds) * len(freqs), "warmup": 1, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (FP64 reductions/refinement); net bf16",
        "data": f"synthetic (held-out seeded media {seeds[0]}-{seeds[-1]})",
        "config": {"workload": f"learned preconditioner, {a.interior}^2+{a.pml} PML, 10 PPW, 3/6/9 MHz x 11 angles",
                   "tolerance": a.tol, "method": f"fgmres({a.restart})", "warm_start": f"learned x{a.warm_steps}",
                   "parallelism": "single GPU"},
        "solves_per_s": {k: round(v, 3) for k, v in sps.items()},
        "per_frequency": rows,
    }
    print(json.dumps(line), flush=True)


def aux_workload(a):
    """The kernels around the solve, timed with HIP events on their stream:
    * capture (helm_capture_accumulate): 64 frequencies of a 1056 x 1216 box
      (the reference's V5 box) accumulating 32 snapshots per launch; HBM
      bytes per launch n (4 * 32 + 32 * 64) against 8 TB/s;
    * band RF (helm_band_rf): 192 receivers x 1200 bins -> 2401 samples;
    * Born Jacobian (helm_born_jacobian): the theranostic quick grid
      (185 x 254 nodes), 64 ports, all 4096 pairs, a 12 x 16 bilinear basis."""
    import numpy as np
    import torch
    from scipy import sparse
    from nsrno_amd.spectra import FrequencyCapture, band_rf
    from nsrno_amd import theranostic as T
    out = {}
    H, W, F, K = 1056, 1216, 64, 32
    cap = FrequencyCapture((H, W), np.arange(1, F + 1) * 40e6 / 2401 * 18, 24010)
    snaps = torch.randn(K, H, W, device="cuda")
    t = _event_time(lambda: cap.add(snaps, 0), 20)
    alg = H * W * (4.0 * K + 32.0 * F)
    out["capture"] = {"kernel": "k_capture (64 frequencies, 32 snapshots per launch)", "avg_launch_us": round(t * 1e6, 1),
                      "algorithmic_bytes": int(alg), "gb_s": round(alg / t / 1e9, 1),
                      "frac": round(alg / t / 1e9 / PEAK_HBM_GBS, 4)}
    bins = np.arange(1, 1201)
    spec = torch.randn(192, 1200, dtype=torch.complex64, device="cuda")
    t = _event_time(lambda: band_rf(spec, bins), 10)
    out["band_rf"] = {"kernel": "k_band_rf (192 x 1200 bins -> 2401 samples)", "avg_launch_us": round(t * 1e6, 1),
                      "complex_macs_per_s": round(192 * 2401 * 1200 / t / 1e9, 1)}
    ax = np.arange(254) * 0.1
    az = np.arange(185) * 0.1
    grid = T.Grid(ax, az, pml_mm=2.0)
    solver = T.Solver(grid)
    xs = np.linspace(3.0, 22.0, 64)
    w = solver.source_weights(np.stack((xs, np.full(64, 2.6)), axis=1), np.tile([0.0, 1.0], (64, 1)), np.full(64, .25))
    u0 = solver.solve(1.0, w)
    xx, zz = np.meshgrid(ax, az)
    cols = [(np.maximum(1 - np.abs(xx - x0) / 1.2, 0) * np.maximum(1 - np.abs(zz - z0) / 1.2, 0)).ravel()
            for z0 in np.linspace(4.0, 15.0, 12) for x0 in np.linspace(4.0, 21.0, 16)]
    basis = sparse.csr_matrix(np.column_stack(cols))
    solver.born_jacobian(1.0, u0, np.arange(64), np.arange(64), basis)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        solver.born_jacobian(1.0, u0, np.arange(64), np.arange(64), basis)
    torch.cuda.synchronize()
    out["born_jacobian"] = {"case": "185x254 nodes, 64 x 64 pairs, 192 bilinear basis functions",
                            "ms_per_call_incl_host_prep_and_copies": round((time.perf_counter() - t0) / reps * 1e3, 2)}
    line = {"metric": "aux kernels (capture, band RF, Born Jacobian)", "value": out["capture"]["gb_s"],
            "unit": "GB/s (capture)", "n_gpus": 1, "higher_is_better": True, "data": "synthetic",
            "roofline": {"bound": "hbm", "kernel": out["capture"]["kernel"], "achieved": out["capture"]["gb_s"],
                         "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": out["capture"]["frac"], "traffic": None},
            "kernels": out}
    print(json.dumps(line), flush=True)


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (profiles/pmc_traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled
    by a copy kernel of known size), or None when it has not been measured."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh)["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median_run(fn, reps=3):
    """One untimed warm-up, then the median wall time of `reps` runs."""
    out = fn()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return float(sorted(times)[len(times) // 2]), out


def cpu_baseline(bench, a, plain_iterations=None):
    """The reference's CPU algorithms on this host, on the bench's workload
    (nsrno/theranostic/wave.py:240-273 and nsrno/wave_solver.py:9-36 as
    restated by the oracle):

    * complex128 sparse LU (assembly + splu factorisation + all 11 solves) of
      the same operator and right-hand sides, one warm-up then the median of
      3 runs, with the BLAS/OpenMP pools limited to 1 thread (the reference's
      default, theranostic/__main__.py:23) and to every host core;
    * plain restarted GMRES(30) (the reference's solver) in numpy on the same
      system: a bounded sample of two restart cycles of one right-hand side
      gives seconds per Krylov step, and the full solve is extrapolated with
      the step counts the GPU's plain GMRES(30) needed (`plain_iterations`)."""
    import numpy as np
    from scipy.sparse.linalg import splu
    from threadpoolctl import threadpool_limits
    from oracle import helmholtz as O
    from nsrno_amd import media
    w = bench.work
    if a.cpu_interior is not None and a.cpu_interior != a.interior:
        w = media.smooth_random_medium(seed=a.seed, interior=a.cpu_interior, pml=a.pml, frequency_mhz=a.freq)
    grid = O.Grid(w.x_mm, w.z_mm, w.pml_mm)
    pos, nrm, wid = media.line_array(w, 192)
    W = O.source_weights(grid, pos, nrm, wid)
    amp = media.plane_wave_amplitudes(pos, bench.angles, w.frequency_mhz, w.c0)
    rhs = -(W @ amp.T) / grid.h_m ** 2
    mk = w.rho0 * w.c0 ** 2 / (w.rho * w.c ** 2) - 1
    me = w.rho0 / w.rho - 1
    kf = (1 + 1j * w.alpha_np_m * w.c / (2 * np.pi * w.frequency_mhz * 1e6)) ** 2

    def direct():
        return O.solve(grid, w.frequency_mhz, None, mk, me, kappa_factor=kf, rhs=rhs, c0=w.c0, check_pml=False)

    cores = os.cpu_count() or 1
    legs = {}
    for n in (1, cores):
        with threadpool_limits(limits=n):
            t, (x, diag) = _median_run(direct)
        legs[n] = dict(seconds=t, x=x, residual=diag["relative_linear_residual"], lu_nonzeros=diag["lu_nonzeros"])
    # plain GMRES(30), numpy complex128, one thread: seconds per Krylov step
    mat = O.assemble(grid, w.frequency_mhz, mk, me, kappa_factor=kf, c0=w.c0, check_pml=False).tocsr()
    b0 = rhs[:, 0]
    with threadpool_limits(limits=1):
        O.gmres(lambda u: mat @ u, lambda u: u, b0, np.zeros_like(b0), restart=30, max_cycles=1, tolerance=1e-30)
        t0 = time.perf_counter()
        O.gmres(lambda u: mat @ u, lambda u: u, b0, np.zeros_like(b0), restart=30, max_cycles=2, tolerance=1e-30)
        per_step = (time.perf_counter() - t0) / 60
    gm = None
    if plain_iterations:
        total = per_step * float(sum(plain_iterations))
        gm = {"value": round(len(plain_iterations) / total, 5), "unit": "solves/s (extrapolated)", "cores": 1,
              "seconds_per_step": round(per_step, 5),
              "sample": f"numpy restatement of wave_solver.py:9-36, 2 GMRES(30) cycles of one right-hand side "
                        f"({per_step * 1e3:.1f} ms per step), scaled to the {sum(plain_iterations)} steps the GPU's "
                        f"plain GMRES(30) needed for the {len(plain_iterations)} angles"}
    return dict(legs=legs, solves=rhs.shape[1], nodes=grid.size, grid_shape=grid.shape, cores=cores,
                cpu_model=_cpu_model(), gmres=gm)


def main():
    a = parse()
    if a.pmc_summary:
        pmc_summary(a.pmc_summary, [x for x in a.pmc_labels.split(",") if x] if a.pmc_labels else None, a.pmc_tag)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if a.copy_calibration:
        copy_calibration()
        return
    if a.workload == "frequencies":
        frequency_sweep(a)
        return
    if a.workload == "conditioned":
        conditioned_workload(a)
        return
    if a.workload == "spectral":
        spectral_workload(a)
        return
    if a.workload == "net":
        net_workload(a)
        return
    if a.workload == "learned":
        learned_workload(a)
        return
    if a.workload == "aux":
        aux_workload(a)
        return
    b = Bench(a)
    print(f"[bench] workload {b.work.nz}x{b.work.nx} R={b.R} f={a.freq} MHz h={b.work.h_mm:.5f} mm", flush=True)
    tuning = None
    if a.autotune:
        b.plan.set_medium_physical(b.c, b.rho, b.alpha, rho0=b.work.rho0)
        b.plan.setup_multigrid(levels=a.mg_levels, shift=a.mg_shift, pre=a.mg_pre, post=a.mg_post,
                               coarse_sweeps=a.mg_coarse, omega=a.mg_omega)
        tuning = b.plan.autotune(b.R)
        print(f"[bench] autotune {tuning}", flush=True)
    # one all-kernel profiled solve (outside the timed region) gives the split
    # and names the dominant single kernel; that kernel is then timed with HIP
    # events inside the timed steps
    b.timed(a.precond, 1, 0, a.max_iterations,
            profile=[k for k in b.plan.PROFILE_KERNELS if not k.endswith("_finest")], warm=a.warm)
    kstats = b.kernel_stats
    singles = [k for k in names if kstats.get(k, {}).get("launches")]
    roof_name = max(singles, key=lambda k: kstats[k]["total_ms"]) if a.roofline_kernel == "auto" else a.roofline_kernel
    timed_name = finest.get(roof_name, roof_name)
    dt, x, infos, traces = b.timed(a.precond, a.steps, a.warmup, a.max_iterations, profile=[timed_name],
                                   warm=a.warm)
    kroof = b.kernel_stats[timed_name]
    applications = [i.applications + b.warm_applications for i in infos]
    solves_per_s = b.R * a.steps / dt
    iters = [i.iterations for i in infos]
    rel = [i.relative_residual for i in infos]
    print(f"[bench] {a.precond}: {dt / a.steps * 1e3:.1f} ms/step, iterations {iters}, max rel {max(rel):.2e}",
          flush=True)
    result = {}
    if a.warm != "none" and a.precond != "none":
        mdt, _, minf, _ = b.timed(a.precond, a.steps, 1, a.max_iterations)
        result["mg_fgmres_without_net"] = dict(
            seconds_per_step=round(mdt / a.steps, 5), solves_per_s=round(b.R * a.steps / mdt, 4),
            iterations=[i.iterations for i in minf], applications=[i.applications for i in minf],
            max_relative_residual=max(i.relative_residual for i in minf))
        print(f"[bench] mg-fgmres without the net: {mdt / a.steps * 1e3:.1f} ms/step, applications "
              f"{max(result['mg_fgmres_without_net']['applications'])}", flush=True)
    if not a.no_sweep:
        # config 3 beside the headline: the 64-frequency batched sweep (its own line: --workload frequencies)
        result["frequency_sweep"] = sweep_summary(a)
        print(f"[bench] frequency sweep: {result['frequency_sweep']['solves_per_s']} solves/s", flush=True)
    if not a.no_plain and a.precond != "none":
        pdt, _, pinf, _ = b.timed("none", 1, 0, a.plain_max_iterations)
        result["plain_gmres"] = dict(
            seconds_per_step=pdt, solves_per_s=b.R / pdt, iterations=[i.iterations for i in pinf],
            applications=[i.applications for i in pinf],
            max_relative_residual=max(i.relative_residual for i in pinf),
            reasons=sorted(set(i.reason for i in pinf)), max_iterations=a.plain_max_iterations,
            restart=a.plain_restart)
        print(f"[bench] plain: {pdt * 1e3:.1f} ms, iterations {result['plain_gmres']['iterations'][:3]}..., "
              f"max rel {result['plain_gmres']['max_relative_residual']:.2e}", flush=True)
    ks = kroof
    achieved = ks["gb_s"]
    kname = names[roof_name] + (" (finest level)" if timed_name != roof_name else "")
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": traffic_from_profile(kname),
                "avg_launch_us": round(ks["avg_us"], 2),
                "algorithmic_bytes_per_launch": int(ks["bytes"] / ks["launches"]), "launches": ks["launches"],
                "share_of_kernel_time": round(kstats[roof_name]["total_ms"] / sum(
                    v["total_ms"] for k, v in kstats.items() if k in names), 3)}
    result["kernels"] = {k: dict(launches=v["launches"], total_ms=round(v["total_ms"], 3),
                                 avg_us=round(v["avg_us"], 2), gb_s=round(v["gb_s"], 1))
                         for k, v in kstats.items() if v["launches"]}
    t_st, t_warm, bytes_alg = b.stencil_roofline()
    result["stencil"] = {"kernel": stencil_kernel_name(tuning),
                         "avg_launch_us": round(t_st * 1e6, 2), "algorithmic_bytes": int(bytes_alg),
                         "gb_s": round(bytes_alg / t_st / 1e9, 1),
                         "frac": round(bytes_alg / t_st / 1e9 / PEAK_HBM_GBS, 4),
                         "inputs": "cold (rotating field pairs, >= 512 MB, none left in the 256 MB last-level cache)",
                         "warm_cache": {"avg_launch_us": round(t_warm * 1e6, 2),
                                        "gb_s": round(bytes_alg / t_warm / 1e9, 1),
                                        "note": "one field pair again and again: inputs served by the cache"}}
    cpu = None
    if not a.no_cpu:
        c = cpu_baseline(b, a, result.get("plain_gmres", {}).get("iterations"))
        one, allc = c["legs"][1], c["legs"][c["cores"]]
        cpu = {"value": round(c["solves"] / one["seconds"], 5), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"complex128 splu (oracle restatement of wave.py:240-273) of the same "
                         f"{c['grid_shape'][0]}x{c['grid_shape'][1]} operator, {c['solves']} RHS, "
                         f"assembly+factor+solve, median of 3 after a warm-up: {one['seconds']:.2f} s at 1 thread, "
                         f"residual {one['residual']:.1e}, LU nonzeros {one['lu_nonzeros']}",
               "cpu_model": c["cpu_model"], "cpu_count": c["cores"],
               "all_cores": {"value": round(c["solves"] / allc["seconds"], 5), "unit": "solves/s",
                             "cores": c["cores"], "seconds": round(allc["seconds"], 3)},
               "plain_gmres30": c["gmres"]}
        if c["grid_shape"] == (b.work.nz, b.work.nx):
            xf = x.reshape(b.R, -1).T.cpu().numpy()
            err = np.linalg.norm(xf - one["x"], axis=0) / np.linalg.norm(one["x"], axis=0)
            result["field_error_vs_c128_direct"] = [float(f"{e:.3e}") for e in err]
        print(f"[bench] cpu baseline {one['seconds']:.2f} s (1 thread), {allc['seconds']:.2f} s ({c['cores']} threads)",
              flush=True)
    line = {
        "metric": "solves/s to rel-res 1e-6, 512²×11 angles: learned FGMRES vs plain GMRES",
        "value": round(solves_per_s, 4), "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(dt / a.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64 (FP64 reductions/refinement)",
        "data": "synthetic (seeded smooth random medium)",
        "config": {"workload": f"helmholtz {a.interior}^2+{a.pml}PML {a.freq}MHz x{a.angles} angles",
                   "nodes": b.n, "rhs": b.R, "tolerance": a.tol, "method": f"fgmres({a.restart})",
                   "precond": a.precond, "warm_start": f"{a.warm} x{a.warm_steps}" if a.warm != "none" else "none",
                   "parallelism": "single GPU"},
        "iterations": iters, "applications": applications, "max_relative_residual": max(rel),
        "reasons": sorted(set(i.reason for i in infos)),
        "roofline": roofline, "cpu_baseline": cpu, "autotune": tuning,
    }
    line.update(result)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
