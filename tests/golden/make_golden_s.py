"""Write golden vectors of the reference's spectral (S) path (run here only).

Imports nsrno.wave_operator / nsrno.wave_solver from /root/reference (never
shipped) with a small default-initialised model (no checkpoint exists: the
zero-initialised last conv layer makes the learned gain 1 and the spectral
modulation 0), and stores setup coefficients (power-law and relaxation
attenuation), the incident field, the operator applied to a seeded field,
the source factor, a full solve_frequency prediction, and the warm-start-only
forward prediction of the default model and of a model with perturbed context
net, spectral modulation and shift (its gain map and multiplier too).  These functions
have no closed form that pins them (discrete-dispersion Newton, tissue mask,
sponge, relaxation modulus, source spectrum).

Usage:  python tests/golden/make_golden_s.py  ->  tests/golden/spectral_small.npz
"""
import math
import sys
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, "/root/reference")
from nsrno.wave_operator import WaveResidualOperator  # noqa: E402
from nsrno.wave_solver import solve_frequency  # noqa: E402


def main():
    torch.manual_seed(0)
    h, w = 40, 200
    dx = 2.0e-4
    spec = dict(medium_shape=[h, w], dx_m=dx, x0_m=-0.0191 - 2 * dx, z0_m=-3 * dx,
                angles_deg=[-8.0, 0.0, 8.0], source_tref_s=[1e-6, 0.0, 2e-6])
    model = WaveResidualOperator(spec, iterations=4).eval()
    rng = np.random.default_rng(7)
    zz, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    blob = np.exp(-((zz - 20) ** 2 + (xx - 100) ** 2) / 40.0)
    c = 1540.0 + 25.0 * blob
    rho = 1000.0 - 30.0 * blob
    medium = torch.tensor(np.stack((c, rho))[None], dtype=torch.float32)
    f = 2.5e6
    with torch.no_grad():
        (ax, az, kc, k0), gain, mult = model.setup(medium, f)
        inc = model.incident(f, 1)
        u = torch.tensor(rng.normal(size=(model.H, model.W)) + 1j * rng.normal(size=(model.H, model.W)),
                         dtype=torch.complex64)
        lu = model.apply_operator(u, (ax, az, kc, k0))
        sf = model.source_factor(f, 2)
    # discrete relaxation modulus (wave_operator.py:55,59-65) on a tissue medium
    rspec = dict(spec, attenuation="relaxation", eta_per_unit_alpha=[0.03, 0.015], taus_s=[8e-8, 2e-8])
    rmodel = WaveResidualOperator(rspec, iterations=4).eval()
    tissue = torch.tensor(np.stack((c - 60.0 * blob, rho + 40.0 * blob))[None], dtype=torch.float32)
    with torch.no_grad():
        (rax, raz, rkc, rk0), _, _ = rmodel.setup(tissue, f)
        pkc = model.setup(tissue, f)[0][2]
    pred, info = solve_frequency(model, medium, f, 1, tolerance=1e-6, max_iterations=4096, restart=32)
    # the warm-start-only forward (wave_operator.py:137-154) of the default model
    # (zero last context layer: gain 1, no spectral modulation) ...
    with torch.no_grad():
        fpred, floss, ftrace = model.forward(medium, f, 1, return_trace=True)
        # its right-hand side, first direction and final iterate (wave_operator.py:129-150)
        fdata, fgain, fmult = model.setup(medium, f)
        finc = model.incident(f, 1)
        fax, faz, fb, fk0 = fdata
        frhs = -model.apply_operator(finc, (fax - 1, faz - 1, fb - (1 - 1j * model.sponge), fk0))
        fdir0 = fgain * torch.fft.ifft2(fmult * torch.fft.fft2(frhs, s=fmult.shape))[:model.H, :model.W]
        fu = torch.zeros_like(finc)
        for _ in range(model.iterations):
            fu = model._iteration(fu, frhs, *fdata, fgain, fmult)
    # ... and of a model whose context net, spectral modulation and shift are
    # away from their initial values (what a trained checkpoint changes)
    lmodel = WaveResidualOperator(spec, iterations=4).eval()
    gen = torch.Generator().manual_seed(5)
    with torch.no_grad():
        for prm in lmodel.context.parameters():
            prm.copy_(torch.randn(prm.shape, generator=gen) * 0.3)
        lmodel.spectral.copy_(torch.randn(2, 4, generator=gen) * 0.5)
        lmodel.log_shift.fill_(math.log(0.25))
        _, lgain, lmult = lmodel.setup(medium, f)
        lpred, lloss, ltrace = lmodel.forward(medium, f, 1, return_trace=True)
    lstate = {"learned_" + k.replace(".", "_"): v.numpy() for k, v in lmodel.state_dict().items()
              if k.startswith(("context", "spectral", "log_shift"))}
    out = dict(h=h, w=w, dx=dx, x0=spec["x0_m"], z0=spec["z0_m"], angles=np.array(spec["angles_deg"]),
               trefs=np.array(spec["source_tref_s"]), c=c, rho=rho, f=f, H=model.H, W=model.W, pz=model.pz,
               px=model.px, face=model.face, rx_cols=model.rx_cols.numpy(),
               ax=ax.numpy(), az=az.numpy(), kc=kc.numpy(), k0=float(k0), gain=np.asarray(gain.numpy()),
               inc=inc.numpy(), u=u.numpy(), lu=lu.numpy(), sf=complex(sf), pred=pred.numpy(),
               pred_residual=info["relative_residual"], pred_reason=info["reason"],
               tissue_c=tissue[0, 0].numpy().astype(np.float64), tissue_rho=tissue[0, 1].numpy().astype(np.float64),
               relax_epa=np.array(rspec["eta_per_unit_alpha"]), relax_taus=np.array(rspec["taus_s"]),
               relax_kc=rkc.numpy(), relax_k0=float(rk0), power_kc_tissue=pkc.numpy(),
               fwd_pred=fpred.numpy(), fwd_loss=float(floss), fwd_trace=np.array(ftrace),
               fwd_rhs=frhs.numpy(), fwd_dir0=fdir0.numpy(), fwd_u=fu.numpy(),
               learned_gain=lgain.numpy(), learned_mult=lmult.numpy(), learned_pred=lpred.numpy(),
               learned_loss=float(lloss), learned_trace=np.array(ltrace), **lstate)
    path = Path(__file__).with_name("spectral_small.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, path.stat().st_size, "bytes", info["reason"], info["relative_residual"],
          info["solver_iterations"])


if __name__ == "__main__":
    main()
