"""Train the learned preconditioner on residuals logged from plain solves.

Mirrors the reference harness nsrno/train_preconditioner.py:45-109 (Adam,
lr 2e-3 with cosine decay to 2e-4, gradient clipping 2.0, residual-based
loss), with the data and the loss the FGMRES direction needs here:

* data: for every seeded medium, the 11 plane-wave right-hand sides are
  solved with multigrid-preconditioned FGMRES(30) one restart cycle at a
  time; the true residual b - A x at every restart is a training sample
  (its error is A^{-1} r; the loss needs only r);
* loss: the residual left after the preconditioned direction,
      L = || A (MG(r) + N(r)) - r ||^2 / ||r||^2,
  so the network learns exactly what the multigrid V-cycle misses; its
  gradient with respect to the network output is 2 A^H (A z - r)/||r||^2
  (A^H = conj(A) on the GPU, helm_apply adjoint);
* everything heavy (operator, adjoint, V-cycle, BF16 MFMA network forward
  and backward, Adam) runs in libhelm.so; torch holds device memory and does
  the few elementwise complex operations of the loss.

Usage (GPU):  python -m nsrno_amd.train_preconditioner --out params.npz
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np
import torch

from . import media
from .helmholtz import HelmholtzPlan, Ports
from .learned import LearnedNet, attach, canvas, features
from .theranostic import Grid, Solver


class Family:
    """Seeded media of one grid/frequency family and their acquisition."""

    def __init__(self, interior=512, pml=32, frequency_mhz=3.0, angles=11, n_elements=192):
        self.interior, self.pml, self.f = interior, pml, frequency_mhz
        w0 = media.smooth_random_medium(seed=0, interior=interior, pml=pml, frequency_mhz=frequency_mhz)
        self.nz, self.nx, self.h_mm, self.c0 = w0.nz, w0.nx, w0.h_mm, w0.c0
        pos, nrm, wid = media.line_array(w0, n_elements)
        W = Solver(Grid(w0.x_mm, w0.z_mm, w0.pml_mm), c0_m_s=w0.c0).source_weights(pos, nrm, wid)
        self.ports = Ports(W)
        self.amp = torch.from_numpy(media.plane_wave_amplitudes(pos, np.linspace(-15, 15, angles), frequency_mhz,
                                                                w0.c0)).cuda()
        self.R = angles

    def medium(self, seed):
        w = media.smooth_random_medium(seed=seed, interior=self.interior, pml=self.pml, frequency_mhz=self.f)
        return tuple(torch.from_numpy(a).cuda() for a in (w.c, w.rho, w.alpha_np_m))

    def plan(self, restart=30):
        return HelmholtzPlan(self.nz, self.nx, self.h_mm, self.f, self.pml * self.h_mm, c0_m_s=self.c0,
                             nrhs=self.R, restart=restart)


def set_medium(plan, med, mg):
    plan.set_medium_physical(*med, rho0=1000.0)
    plan.setup_multigrid(**mg)


def log_residuals(fam, plan, med, mg, cycles, restart=30):
    """Residuals b - A x at the first `cycles` restarts of MG-FGMRES, unit RMS, complex64."""
    set_medium(plan, med, mg)
    b = plan.inject(fam.ports, fam.amp)
    x = torch.zeros_like(b)
    out = []
    for _ in range(cycles):
        r = (b - plan.apply(x)).to(torch.complex64)
        rms = r.abs().square().mean(dim=(1, 2), keepdim=True).sqrt().clamp_min(1e-30)
        out.append(r / rms)
        x, _ = plan.solve(b, x0=x, tolerance=1e-12, restart=restart, max_iterations=restart, precond="mg")
    return torch.cat(out)


def log_krylov(fam, plan, med, mg, steps, restart=30):
    """Arnoldi vectors v_j of MG-preconditioned FGMRES(restart) on the 11
    right-hand sides (the inputs the preconditioner really sees), unit RMS.
    A plain torch Arnoldi (CGS2) on the device operator, for data logging."""
    set_medium(plan, med, mg)
    b = plan.inject(fam.ports, fam.amp).to(torch.complex64)
    R = b.shape[0]
    x = torch.zeros_like(b)
    out = []
    done = 0
    while done < steps:
        r = b - plan.apply(x)
        beta = torch.linalg.vector_norm(r.reshape(R, -1), dim=1)
        V = [r / beta[:, None, None]]
        Z, H = [], torch.zeros(R, restart + 1, restart, dtype=torch.complex128, device=b.device)
        m = min(restart, steps - done)
        for j in range(m):
            out.append(V[j] / V[j].abs().square().mean(dim=(1, 2), keepdim=True).sqrt().clamp_min(1e-30))
            z = plan.precondition(V[j].contiguous())
            Z.append(z)
            w = plan.apply(z)
            for _ in range(2):
                for k in range(j + 1):
                    c = (V[k].conj() * w).sum(dim=(1, 2))
                    w = w - c[:, None, None] * V[k]
                    H[:, k, j] += c.to(torch.complex128)
            hn = torch.linalg.vector_norm(w.reshape(R, -1), dim=1)
            H[:, j + 1, j] = hn.to(torch.complex128)
            V.append(w / hn[:, None, None].clamp_min(1e-30))
        e1 = torch.zeros(R, m + 1, dtype=torch.complex128, device=b.device)
        e1[:, 0] = beta.to(torch.complex128)
        y = torch.linalg.lstsq(H[:, :m + 1, :m], e1[..., None]).solution[..., 0].to(torch.complex64)
        x = x + sum(y[:, k, None, None] * Z[k] for k in range(m))
        done += m
    return torch.cat(out)


def loss_and_grad(plan, net, r, kind="residual"):
    """L and dL/d(net output) for residuals r (complex64 [B, nz, nx]).

    kind "residual": L = ||A z - r||^2 / ||r||^2 (the direction as a solve);
    kind "angle":    L = 1 - |<A z, r>|^2 / (||A z||^2 ||r||^2), the residual
                     left after the best scalar step along z (the minimal-
                     residual step of a Krylov method), blind to the scale of z.
    z = MG(r) + h^2 RMS(r) N(r).  The gradient reaches the network output
    through A^H (the adjoint apply on the GPU)."""
    B, nz, nx = r.shape
    zmg = plan.precondition(r)
    feats = features(plan, r)
    out = net.forward(feats)
    s = r.abs().square().mean(dim=(1, 2)).sqrt() * plan.h_m ** 2       # output unit h^2 RMS(r)
    o = out[:, :nz, :nx, :2].float()
    z = zmg + torch.complex(o[..., 0], o[..., 1]) * s[:, None, None]
    w = plan.apply(z.contiguous())
    r2 = r.abs().square().sum(dim=(1, 2))
    if kind == "angle":
        a = (w.conj() * r).sum(dim=(1, 2))                       # <A z, r>
        q = w.abs().square().sum(dim=(1, 2))
        loss = (1.0 - a.abs().square() / (q * r2)).mean()
        # dL/d conj(w) = -(conj(a) r - |a|^2/q w) / (q p)
        dw = -(a.conj()[:, None, None] * r - (a.abs().square() / q)[:, None, None] * w) / (q * r2)[:, None, None]
        g = plan.apply(dw.to(torch.complex64).contiguous(), adjoint=True)
        coef = (2.0 * s / B)[:, None, None]
    else:
        res = w - r
        loss = (res.abs().square().sum(dim=(1, 2)) / r2).mean()
        g = plan.apply(res.contiguous(), adjoint=True)
        coef = (2.0 * s / r2 / B)[:, None, None]
    dout = torch.zeros_like(out, dtype=torch.float32)
    dout[:, :nz, :nx, 0] = g.real * coef
    dout[:, :nz, :nx, 1] = g.imag * coef
    return float(loss), dout.to(torch.bfloat16)


def arnoldi(plan, r0, k, kind):
    """k steps of flexible Arnoldi (CGS2, torch on the device operator) from
    r0 (complex64 [1, nz, nx]) with the plan's preconditioner `kind`.
    Returns V [k, nz, nx] (unit vectors), W = A Z [k, nz, nx] and the
    least-squares coefficients y of min ||r0 - W y||, plus that residual."""
    beta = torch.linalg.vector_norm(r0)
    V = [r0[0] / beta]
    Wl = []
    H = torch.zeros(k + 1, k, dtype=torch.complex128, device=r0.device)
    for j in range(k):
        z = plan.precondition(V[j][None].contiguous(), kind=kind)
        w = plan.apply(z)[0]
        Wl.append(w)
        for _ in range(2):
            for i in range(j + 1):
                c = (V[i].conj() * w).sum()
                w = w - c * V[i]
                H[i, j] += c.to(torch.complex128)
        hn = torch.linalg.vector_norm(w)
        H[j + 1, j] = hn.to(torch.complex128)
        V.append(w / hn.clamp_min(1e-30))
    e1 = torch.zeros(k + 1, 1, dtype=torch.complex128, device=r0.device)
    e1[0, 0] = beta.to(torch.complex128)
    y = torch.linalg.lstsq(H, e1).solution[:, 0]
    W = torch.stack(Wl)
    rho = r0[0] - (y.to(torch.complex64)[:, None, None] * W).sum(dim=0)
    return torch.stack(V[:k]), W, y, rho


def subspace_loss_and_grad(plan, net, r0, k):
    """Residual left by one k-step flexible GMRES cycle with the learned
    directions z_j = MG(v_j) + h^2 RMS(v_j) N(v_j):

        L = min_y ||r0 - sum_j y_j A z_j||^2 / ||r0||^2.

    The Arnoldi vectors v_j come from the current network (on-policy) and are
    held fixed in the gradient; by optimality of y,
        dL/d conj(z_j) = -conj(y_j) A^H rho / ||r0||^2,  rho = r0 - A Z y,
    one adjoint apply for all k directions.  The network is re-run on the k
    vectors as one batch for its backward pass."""
    _, nz, nx = r0.shape
    V, W, y, rho = arnoldi(plan, r0, k, _lib_pc_learned())
    p = r0.abs().square().sum()
    loss = rho.abs().square().sum() / p
    feats = features(plan, V.contiguous())
    out = net.forward(feats)  # the same outputs the Arnoldi used, kept for the backward pass
    s = V.abs().square().mean(dim=(1, 2)).sqrt() * plan.h_m ** 2
    g = plan.apply(rho[None].contiguous(), adjoint=True)[0]
    dz = -(y.conj().to(torch.complex64)[:, None, None] * g[None]) / p
    dout = torch.zeros_like(out, dtype=torch.float32)
    coef = (2.0 * s)[:, None, None]
    dout[:, :nz, :nx, 0] = dz.real * coef
    dout[:, :nz, :nx, 1] = dz.imag * coef
    return float(loss), dout.to(torch.bfloat16)


def warm_residuals(plan, r, steps, kind):
    """Residuals after `steps` minimal-residual steps r <- r - alpha A P(r)
    (the warm start of helm_warm_start) with preconditioner `kind`."""
    for _ in range(steps):
        z = plan.precondition(r.contiguous(), kind=kind)
        w = plan.apply(z)
        a = (w.conj() * r).sum(dim=(1, 2)) / w.abs().square().sum(dim=(1, 2)).clamp_min(1e-30)
        r = r - a[:, None, None] * w
    return r


def subspace_baseline(plan, r0, k):
    """The same cycle residual with the multigrid directions alone."""
    from ._lib import PC_MG
    _, _, _, rho = arnoldi(plan, r0, k, PC_MG)
    return float(rho.abs().square().sum() / r0.abs().square().sum())


def _lib_pc_learned():
    from ._lib import PC_LEARNED
    return PC_LEARNED


def baseline_loss(plan, r, kind="residual"):
    """The same loss with the multigrid direction alone."""
    zmg = plan.precondition(r)
    w = plan.apply(zmg)
    r2 = r.abs().square().sum(dim=(1, 2))
    if kind == "angle":
        a = (w.conj() * r).sum(dim=(1, 2))
        return float((1.0 - a.abs().square() / (w.abs().square().sum(dim=(1, 2)) * r2)).mean())
    return float(((w - r).abs().square().sum(dim=(1, 2)) / r2).mean())


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="learned_params.npz")
    ap.add_argument("--media", type=int, default=200)
    ap.add_argument("--seed0", type=int, default=100)
    ap.add_argument("--cycles", type=int, default=8)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-min", type=float, default=2e-4)
    ap.add_argument("--clip", type=float, default=2.0)
    ap.add_argument("--interior", type=int, default=512)
    ap.add_argument("--freq", type=float, default=3.0)
    ap.add_argument("--freqs", default=None,
                    help="comma-separated frequencies in MHz (warmstart loss): media spread over them in turn, one "
                         "plan per frequency on the same grid (10 points per wavelength at each), one shared net")
    ap.add_argument("--init", default=None, help="start from these parameters")
    ap.add_argument("--log-every", type=int, default=50)
    ap.add_argument("--eval", default=None, help="only evaluate these parameters on held-out media")
    ap.add_argument("--eval-seeds", default="0,1000,1001,1002")
    ap.add_argument("--restart", type=int, default=20, help="FGMRES restart length of the evaluation")
    ap.add_argument("--eval-warm", action="store_true", help="evaluate the net as a warm start before MG-FGMRES")
    ap.add_argument("--linear", action="store_true", help="linear network (identity activations, no biases)")
    ap.add_argument("--loss", default="subspace", choices=["subspace", "warmstart", "angle", "residual"],
                    help="residual after an on-policy k-step FGMRES cycle (subspace), the scale-invariant "
                         "one-step minimal-residual loss (angle) or the one-step solve residual")
    ap.add_argument("--unroll", type=int, default=16, help="cycle length k of the subspace loss")
    ap.add_argument("--base-every", type=int, default=5, help="steps between multigrid-only reference cycles")
    ap.add_argument("--data-restart", type=int, default=30,
                    help="restart length of the multigrid FGMRES whose restart residuals are the data")
    ap.add_argument("--out-scale", type=float, default=0.1,
                    help="initial scale of the output layer (0: the net starts as the multigrid direction)")
    ap.add_argument("--data", default="krylov", choices=["krylov", "restart"],
                    help="train on Arnoldi vectors (what FGMRES feeds the preconditioner) or restart residuals")
    a = ap.parse_args(argv)
    if a.eval and a.eval_warm:
        evaluate_warm(a.eval, seeds=tuple(int(v) for v in a.eval_seeds.split(",")), interior=a.interior,
                      freq=a.freq, restart=a.restart)
        return
    if a.eval:
        evaluate(a.eval, seeds=tuple(int(v) for v in a.eval_seeds.split(",")), interior=a.interior, freq=a.freq,
                 restart=a.restart)
        return
    torch.manual_seed(0)
    mg = dict(levels=5, shift=0.5, pre=2, post=2, coarse_sweeps=10)
    freqs = [float(v) for v in a.freqs.split(",")] if a.freqs else [a.freq]
    if len(freqs) > 1 and a.loss != "warmstart":
        raise ValueError("--freqs needs --loss warmstart")
    fams = [Family(interior=a.interior, frequency_mhz=f) for f in freqs]
    plans = [f.plan() for f in fams]
    if len({(f.nz, f.nx) for f in fams}) != 1:
        raise ValueError("--freqs: every frequency must give the same grid (one canvas for the net)")
    fam, plan = fams[0], plans[0]
    t0 = time.perf_counter()
    meds, data, which = [], [], []
    for k in range(a.media):
        fi = k % len(fams)  # medium k belongs to frequency k mod len(freqs)
        fam, plan = fams[fi], plans[fi]
        med = fam.medium(a.seed0 + k)
        meds.append(med)
        which.append(fi)
        if a.loss == "warmstart":  # the source fields themselves (unit RMS)
            set_medium(plan, med, mg)
            b = plan.inject(fam.ports, fam.amp).to(torch.complex64)
            data.append(b / b.abs().square().mean(dim=(1, 2), keepdim=True).sqrt())
        elif a.data == "krylov":
            data.append(log_krylov(fam, plan, med, mg, a.cycles * 30)[:: max(1, (a.cycles * 30) // 8)])
        else:
            data.append(log_residuals(fam, plan, med, mg, a.cycles, restart=a.data_restart))
        if k % 20 == 0:
            print(json.dumps(dict(phase="data", media=k + 1, samples=sum(d.shape[0] for d in data),
                                  seconds=round(time.perf_counter() - t0, 1))), flush=True)
    n_samples = sum(d.shape[0] for d in data)
    print(json.dumps(dict(phase="data", samples=n_samples, seconds=round(time.perf_counter() - t0, 1))), flush=True)
    params = None if a.init is None else LearnedNet.load_params(a.init)
    if params is None and not a.linear:
        from .learned import init_params
        params = init_params(out_scale=a.out_scale)
    if params is None and a.linear:
        from .learned import init_params, NL as _NL, PSZ as _PSZ, WSZ as _WSZ
        params = init_params().reshape(_NL, _PSZ)
        params[:, :_WSZ] *= 0.5  # keep the linear cascade's gain moderate
        params = params.ravel()
    batch = a.unroll if a.loss == "subspace" else a.batch
    net = LearnedNet(canvas(fam.nz), canvas(fam.nx), batch, training=True, params=params, linear=a.linear)
    for p in plans:
        attach(p, net)
    rng = np.random.default_rng(0)
    hist, ratios = [], []
    t1 = time.perf_counter()
    for step in range(a.steps):
        k = int(rng.integers(len(meds)))
        plan = plans[which[k]]
        set_medium(plan, meds[k], mg)
        if a.loss == "subspace":
            idx = int(rng.integers(data[k].shape[0]))
            r = data[k][idx:idx + 1].contiguous()
            loss, dout = subspace_loss_and_grad(plan, net, r, a.unroll)
        elif a.loss == "warmstart":
            # on-policy: the residual after t learned warm-start steps from the source field
            idx = torch.from_numpy(rng.choice(data[k].shape[0], size=a.batch, replace=False)).cuda()
            t = int(rng.integers(a.unroll + 1))
            with torch.no_grad():
                r = warm_residuals(plan, data[k][idx].contiguous(), t, _lib_pc_learned())
                r = r / r.abs().square().mean(dim=(1, 2), keepdim=True).sqrt().clamp_min(1e-30)
            loss, dout = loss_and_grad(plan, net, r.to(torch.complex64).contiguous(), "angle")
        else:
            idx = torch.from_numpy(rng.choice(data[k].shape[0], size=a.batch, replace=False)).cuda()
            r = data[k][idx].contiguous()
            loss, dout = loss_and_grad(plan, net, r, a.loss)
        if not math.isfinite(loss):
            raise RuntimeError(f"non-finite loss at step {step}")  # as train_wave.py:80-82
        net.backward(dout)
        lr = a.lr_min + 0.5 * (a.lr - a.lr_min) * (1 + math.cos(math.pi * step / max(1, a.steps)))
        net.adam(lr=lr, clip=a.clip)
        hist.append(loss)
        if a.loss == "subspace" and step % a.base_every == 0:
            # log(learned / multigrid-only) on the same starting residual
            ratios.append(math.log(max(loss, 1e-30) / max(subspace_baseline(plan, r, a.unroll), 1e-30)))
        if step % a.log_every == 0 or step == a.steps - 1:
            rec = dict(phase="train", step=step, loss=round(float(np.mean(hist[-a.log_every:])), 5), lr=round(lr, 6),
                       seconds=round(time.perf_counter() - t1, 1))
            if a.loss == "subspace":
                rec["geomean_ratio_vs_mg"] = round(math.exp(float(np.mean(ratios[-20:]))), 4) if ratios else None
            else:
                rec["mg_only"] = round(baseline_loss(plan, r, a.loss), 5)
            print(json.dumps(rec), flush=True)
    net.save(a.out)
    print(json.dumps(dict(phase="done", out=a.out, steps=a.steps, final_loss=float(np.mean(hist[-50:])))),
          flush=True)
    del data
    torch.cuda.empty_cache()




def evaluate_warm(params_path, seeds=(0, 1000, 1001), interior=512, freq=3.0, tol=1e-6, restart=20,
                  steps=(0, 4, 8, 16)):
    """Warm start (learned or multigrid directions, `steps` minimal-residual
    steps) followed by multigrid FGMRES: residual after the warm start,
    total operator applications and wall seconds per 11-angle batch."""
    mg = dict(levels=5, shift=0.5, pre=2, post=2, coarse_sweeps=10)
    fam = Family(interior=interior, frequency_mhz=freq)
    plan = fam.plan(restart)
    net = LearnedNet(canvas(fam.nz), canvas(fam.nx), fam.R, params=LearnedNet.load_params(params_path),
                     linear=LearnedNet.load_linear(params_path))
    attach(plan, net)
    rows = []
    for seed in seeds:
        med = fam.medium(seed)
        row = dict(seed=seed)
        for kind in ("mg", "learned"):
            for n in steps:
                if kind == "mg" and n == 0 and "mg0" in row:
                    continue
                for rep in range(2):
                    torch.cuda.synchronize()
                    t = time.perf_counter()
                    set_medium(plan, med, mg)
                    b = plan.inject(fam.ports, fam.amp)
                    x = torch.zeros_like(b)
                    if n:
                        plan.warm_start(b, x, n, precond=kind)
                    r0 = float((torch.linalg.vector_norm((b - plan.apply(x)).reshape(fam.R, -1), dim=1)
                                / torch.linalg.vector_norm(b.reshape(fam.R, -1), dim=1)).max())
                    x, info = plan.solve(b, x0=x, tolerance=tol, restart=restart, max_iterations=4000, precond="mg")
                    torch.cuda.synchronize()
                    dt = time.perf_counter() - t
                row[f"{kind}{n}"] = dict(seconds=round(dt, 4), warm_residual=round(r0, 5),
                                         iterations=max(i.iterations for i in info),
                                         applications=max(i.applications for i in info) + 2 * n,
                                         reasons=sorted(set(i.reason for i in info)))
        rows.append(row)
        print(json.dumps(row), flush=True)
    return rows


def evaluate(params_path, seeds=(0, 1000, 1001, 1002), interior=512, freq=3.0, tol=1e-6, restart=30):
    """MG-FGMRES vs learned FGMRES (z = MG(v) + N(v)) on held-out media:
    Krylov steps, operator applies and wall seconds per 11-angle batch."""
    mg = dict(levels=5, shift=0.5, pre=2, post=2, coarse_sweeps=10)
    fam = Family(interior=interior, frequency_mhz=freq)
    plan = fam.plan(restart)
    net = LearnedNet(canvas(fam.nz), canvas(fam.nx), fam.R, params=LearnedNet.load_params(params_path),
                     linear=LearnedNet.load_linear(params_path))
    attach(plan, net)
    rows = []
    for seed in seeds:
        med = fam.medium(seed)
        row = dict(seed=seed)
        for pc in ("mg", "learned"):
            for rep in range(2):  # second run is timed
                torch.cuda.synchronize()
                t = time.perf_counter()
                set_medium(plan, med, mg)
                b = plan.inject(fam.ports, fam.amp)
                trace = []
                x, info = plan.solve(b, tolerance=tol, restart=restart, max_iterations=4000, precond=pc,
                                     callback=lambda n, r: trace.append((n, float("%.2e" % max(r)))))
                torch.cuda.synchronize()
                dt = time.perf_counter() - t
            row[pc] = dict(seconds=round(dt, 4), iterations=max(i.iterations for i in info), trace=trace[:60],
                           applications=max(i.applications for i in info),
                           residual=max(i.relative_residual for i in info),
                           reasons=sorted(set(i.reason for i in info)))
        rows.append(row)
        print(json.dumps(row), flush=True)
    return rows


if __name__ == "__main__":
    main()
