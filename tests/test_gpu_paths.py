"""The solver's alternate kernel paths give the reference's fields too.

Each fast path of the library keeps its predecessor behind an environment
switch that is read once per process (HELM_DCGS_FUSED, HELM_LAZY_BASIS,
HELM_RESID_BLOCK, HELM_STENCIL_BLOCK, HELM_MG_TAIL_LDS), so a test process
only ever runs one side.  This test solves the golden problem (nsrno/theranostic/wave.py
Solver.solve, fields from tests/golden/theranostic_small.npz) in two child
processes, one with every switch off and one with the defaults, and checks
both against the reference's direct-solve fields.
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

SWITCHES = {"HELM_DCGS_FUSED": "0", "HELM_LAZY_BASIS": "0", "HELM_STENCIL_BLOCK": "0", "HELM_MG_TAIL_LDS": "0"}
# the per-node restart residual (the lazy basis stays on, so that kernel runs)
SWITCHES_RESID = {"HELM_RESID_BLOCK": "0"}


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_switched_off_paths_match_default_and_reference(tmp_path):
    golden = np.load(GOLDEN_PATH)
    fields = {}
    defaults = {k: v for k, v in os.environ.items() if k not in SWITCHES and k not in SWITCHES_RESID}
    for name, env in (("alternate", dict(defaults, **SWITCHES)), ("per-node residual", dict(defaults, **SWITCHES_RESID)),
                      ("default", defaults)):
        out = tmp_path / f"{name.replace(' ', '_')}.npy"
        print(f"child solve: {name} paths", flush=True)
        subprocess.run([sys.executable, "-c", SOLVE, str(ROOT), str(GOLDEN_PATH), str(out)], env=env, check=True,
                       timeout=240)
        fields[name] = np.load(out)
    for name, f in fields.items():
        assert rel(f, golden["fields"]) < 1e-7, name
    assert rel(fields["alternate"], fields["default"]) < 1e-7
    assert rel(fields["per-node residual"], fields["default"]) < 1e-7
