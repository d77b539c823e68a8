"""Write golden vectors of the reference's conditioned split-PML operator (run here only).

Imports nsrno.conditioned_wave / nsrno.dense_wave from /root/reference (never
shipped), builds FrequencyGeometry and ConditionedFrequencySystem on a
seeded 64 x 64 medium with the geometry of the reference's own algebra check
(nsrno/check_pressure_joint_kernels.py:17-19), evaluated in float64 /
complex128 (torch default dtype float64), and stores the geometry factors,
the system coefficients, the operator (__call__), projected_operator,
adjoint_operator and source_rhs on seeded inputs, and the fixed
preconditioner's output.  The split-PML weights, the relaxation modulus and
the coefficient cropping have no closed form that pins them.

Usage:  python tests/golden/make_golden_c.py  ->  tests/golden/conditioned_small.npz
"""
import sys
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, "/root/reference")
from nsrno.conditioned_wave import ConditionedFrequencySystem, FrequencyGeometry  # noqa: E402
from nsrno.dense_wave import adjoint_operator, projected_operator, source_rhs  # noqa: E402

SPEC = {"medium_shape": [2, 64, 64], "dx_m": 50e-6, "x0_m": -.0016, "z0_m": -.0016, "native_dt_s": 5e-9,
        "pml": 24, "eta_per_unit_alpha": [.015, .025], "taus_s": [12e-9, 35e-9], "source_c_lookup_zx": [6, 3]}


def smooth(rng, h, w, sigma=4.0):
    from scipy.ndimage import gaussian_filter
    f = gaussian_filter(rng.normal(size=(h, w)), sigma)
    return f / np.abs(f).max()


def main():
    torch.set_default_dtype(torch.float64)
    rng = np.random.default_rng(11)
    g = FrequencyGeometry(SPEC)
    h, w = g.h, g.w
    c = 1540.0 + 40.0 * smooth(rng, h, w)
    rho = 1000.0 + 60.0 * smooth(rng, h, w)
    alpha = 0.4 + 0.3 * smooth(rng, h, w)
    medium = torch.tensor(np.stack((c, rho)))
    alpha_t = torch.tensor(alpha)
    f = 360 * 40e6 / 2401
    s = ConditionedFrequencySystem(g, medium, alpha_t, f)
    H, W = g.H, g.W
    # small-integer inputs (exact in every precision, and they compress)
    p = torch.tensor(rng.integers(-8, 9, size=(H, W)) + 1j * rng.integers(-8, 9, size=(H, W)), dtype=torch.complex128)
    line = torch.tensor(rng.integers(-8, 9, size=(2, w)) + 1j * rng.integers(-8, 9, size=(2, w)),
                        dtype=torch.complex128)
    r = torch.tensor(rng.integers(-8, 9, size=(H, W)) + 1j * rng.integers(-8, 9, size=(H, W)), dtype=torch.complex128)
    with torch.no_grad():
        call = s(p)
        proj = projected_operator(s, p)
        adj = adjoint_operator(s, p)
        src = source_rhs(s, line)
        fixed = s.fixed_preconditioner()(r)
    out = dict(spec_json=np.array(__import__("json").dumps(SPEC)), c=c, rho=rho, alpha=alpha, f=f, H=H, W=W,
               pz=g.pz, px=g.px, face=g.face, rx_cols=g.rx_cols.numpy(), mask=g.mask.numpy(),
               arx=g.arx.numpy(), arz=g.arz.numpy(), avx=g.avx.numpy(), avz=g.avz.numpy(),
               sc=s.c.numpy(), srho=s.rho.numpy(), salpha=s.alpha.numpy(), D=complex(s.D), k0=float(s.k0),
               ax=s.ax.numpy(), az=s.az.numpy(), mass=s.mass.numpy(), svx=s.svx.numpy(), svz=s.svz.numpy(),
               gvx=s.gvx.numpy(), gvz=s.gvz.numpy(), mx=s.mx.numpy(), mz=s.mz.numpy(),
               p=p.numpy(), line=line.numpy(), r=r.numpy(), call=call.numpy(), proj=proj.numpy(),
               adj=adj.numpy(), src=src.numpy(), fixed=fixed.numpy())
    path = Path(__file__).with_name("conditioned_small.npz")
    np.savez_compressed(path, **out)
    print(path, {k: getattr(v, "shape", v) for k, v in out.items() if k != "spec_json"})


if __name__ == "__main__":
    main()
