"""The solver's alternate kernel paths give the reference's fields too.

Each fast path of the library keeps its predecessor behind an environment
switch that is read once per process (HELM_DCGS_FUSED, HELM_LAZY_BASIS,
HELM_RESID_BLOCK, HELM_STENCIL_BLOCK, HELM_MG_TAIL_LDS, HELM_XCD_ORDER,
HELM_RESTART_AHEAD, HELM_COND64_TILED, HELM_STAG8_TILED),
so a test process only ever runs one side.  The tests below run child
processes with the switches off and with the defaults and check both against
the reference's golden outputs: the direct-solve fields of
nsrno/theranostic/wave.py Solver.solve (tests/golden/theranostic_small.npz)
the conditioned operator of nsrno/dense_wave.py projected_operator /
adjoint_operator in FP64 (tests/golden/conditioned_small.npz), and the S
operator of nsrno/wave_operator.py apply_operator in FP32 against the oracle's
restatement of it (oracle/spectral.py apply_stencil).
"""
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
GOLDEN_PATH = Path(__file__).parent / "golden" / "theranostic_small.npz"

# multigrid FGMRES on the golden grid, 4 right-hand sides (one per port), to 1e-10
SOLVE = r'''
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import helmholtz as O
from nsrno_amd.helmholtz import HelmholtzPlan, Ports

g = np.load(sys.argv[2])
axis = g["axis"]
grid = O.Grid(axis, axis, 6.0)
plan = HelmholtzPlan(grid.shape[0], grid.shape[1], grid.h_mm, float(g["frequency_mhz"]), 6.0, x0_mm=axis[0],
                     z0_mm=axis[0], nrhs=4, restart=8)
plan.set_medium(1 + g["m_eta"], (1 + g["m_kappa"]).astype(complex))
plan.setup_multigrid(levels=3, shift=0.5, pre=2, post=2, coarse_sweeps=30)
w = O.source_weights(grid, g["positions"], g["normals"], g["widths"])
rhs = plan.inject(Ports(w), np.eye(4))
x, infos = plan.solve(rhs, tolerance=1e-10, restart=4, precond="mg", max_iterations=20000, graph=True)
assert all(i.reason == "converged" for i in infos), infos
np.save(sys.argv[3], x.cpu().numpy().reshape(4, -1).T)
'''

SWITCHES = {"HELM_DCGS_FUSED": "0", "HELM_LAZY_BASIS": "0", "HELM_STENCIL_BLOCK": "0", "HELM_MG_TAIL_LDS": "0",
            "HELM_XCD_ORDER": "0"}
# the per-node restart residual (the lazy basis stays on, so that kernel runs)
SWITCHES_RESID = {"HELM_RESID_BLOCK": "0"}
# the next restart cycle queued only after the host has read the status: the
# same kernels on the same data, so the same bits
SWITCHES_AHEAD = {"HELM_RESTART_AHEAD": "0"}


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_switched_off_paths_match_default_and_reference(tmp_path):
    golden = np.load(GOLDEN_PATH)
    fields = {}
    defaults = {k: v for k, v in os.environ.items()
                if k not in SWITCHES and k not in SWITCHES_RESID and k not in SWITCHES_AHEAD}
    for name, env in (("alternate", dict(defaults, **SWITCHES)), ("per-node residual", dict(defaults, **SWITCHES_RESID)),
                      ("restart after status", dict(defaults, **SWITCHES_AHEAD)), ("default", defaults)):
        out = tmp_path / f"{name.replace(' ', '_')}.npy"
        print(f"child solve: {name} paths", flush=True)
        subprocess.run([sys.executable, "-c", SOLVE, str(ROOT), str(GOLDEN_PATH), str(out)], env=env, check=True,
                       timeout=240)
        fields[name] = np.load(out)
    for name, f in fields.items():
        assert rel(f, golden["fields"]) < 1e-7, name
    assert rel(fields["alternate"], fields["default"]) < 1e-7
    assert rel(fields["per-node residual"], fields["default"]) < 1e-7
    assert np.array_equal(fields["restart after status"], fields["default"])


COND_GOLDEN = Path(__file__).parent / "golden" / "conditioned_small.npz"

# FP64 projected and adjoint applies of the conditioned operator on the golden input
COND_APPLY = r'''
import json, sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
from nsrno_amd.conditioned import ConditionedFrequencySystem, FrequencyGeometry, adjoint_operator, projected_operator

g = dict(np.load(sys.argv[2]))
geo = FrequencyGeometry(json.loads(str(g["spec_json"])))
medium = torch.tensor(np.stack((g["c"], g["rho"])), dtype=torch.float64, device="cuda")
s = ConditionedFrequencySystem(geo, medium, torch.tensor(g["alpha"], dtype=torch.float64, device="cuda"),
                               float(g["f"]))
p = torch.from_numpy(g["p"]).cuda()
np.savez(sys.argv[3], proj=projected_operator(s, p).cpu().numpy(), adj=adjoint_operator(s, p).cpu().numpy())
'''


def test_conditioned_fp64_tile_and_streaming_match_reference(tmp_path):
    golden = np.load(COND_GOLDEN)
    base = {k: v for k, v in os.environ.items() if k != "HELM_COND64_TILED"}
    for name, env in (("tiled", dict(base, HELM_COND64_TILED="1")), ("streaming", base)):
        out = tmp_path / f"{name}.npz"
        print(f"child apply: {name} FP64 conditioned kernel", flush=True)
        subprocess.run([sys.executable, "-c", COND_APPLY, str(ROOT), str(COND_GOLDEN), str(out)], env=env,
                       check=True, timeout=240)
        got = np.load(out)
        assert rel(got["proj"], golden["proj"]) < 1e-12, name
        assert rel(got["adj"], golden["adj"]) < 1e-12, name


SPEC_GOLDEN = Path(__file__).parent / "golden" / "spectral_small.npz"

# FP32 forward and adjoint applies of the S operator on the golden medium, and
# the oracle's k0^2 conj(L) / k0^2 L of the same fields
SPEC_APPLY = r'''
import sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
from oracle import spectral as S
from nsrno_amd.wave_operator import WaveResidualOperator

G = np.load(sys.argv[2])
spec = dict(medium_shape=[int(G["h"]), int(G["w"])], dx_m=float(G["dx"]), x0_m=float(G["x0"]), z0_m=float(G["z0"]),
            angles_deg=list(G["angles"]), source_tref_s=list(G["trefs"]))
model = WaveResidualOperator(spec, iterations=4)
data, _, _ = model.setup(torch.tensor(np.stack((G["c"], G["rho"]))[None], dtype=torch.float32), float(G["f"]))
box = S.Box(int(G["h"]), int(G["w"]), float(G["dx"]), float(G["x0"]), float(G["z0"]))
ax, az, kc, k0 = S.setup(box, np.float32(G["c"]), np.float32(G["rho"]), float(G["f"]))
u = np.random.default_rng(5).normal(size=(3, box.H, box.W)) + 1j * np.random.default_rng(6).normal(size=(3, box.H, box.W))
u32 = torch.from_numpy(u.astype(np.complex64)).cuda()
np.savez(sys.argv[3], fwd=data.plan.apply(u32).cpu().numpy(), adj=data.plan.apply(u32, adjoint=True).cpu().numpy(),
         fwd_ref=k0 ** 2 * np.conj(S.apply_stencil(np.conj(u), ax, az, kc, k0, box.dx)),
         adj_ref=k0 ** 2 * S.apply_stencil(u, ax, az, kc, k0, box.dx))
'''


def test_s_operator_fp32_tile_and_streaming_match_oracle(tmp_path):
    base = {k: v for k, v in os.environ.items() if k != "HELM_STAG8_TILED"}
    got = {}
    for name, env in (("tiled", dict(base, HELM_STAG8_TILED="1")), ("streaming", base)):
        out = tmp_path / f"{name}.npz"
        print(f"child apply: {name} FP32 S operator", flush=True)
        subprocess.run([sys.executable, "-c", SPEC_APPLY, str(ROOT), str(SPEC_GOLDEN), str(out)], env=env,
                       check=True, timeout=240)
        got[name] = np.load(out)
        assert rel(got[name]["fwd"], got[name]["fwd_ref"]) < 5e-6, name
        assert rel(got[name]["adj"], got[name]["adj_ref"]) < 5e-6, name
    assert rel(got["tiled"]["fwd"], got["streaming"]["fwd"]) < 5e-6
