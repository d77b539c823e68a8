"""Write the golden vectors of the theranostic CPU forward (run here only).

Imports the reference ``nsrno.theranostic.wave`` from /root/reference (it is
never shipped) and stores small input/output vectors for the functions whose
exact arithmetic no closed form pins: the assembled flux matrix
(Solver.matrix, wave.py:174-188, which fixes the edge-stretch ordering of
wave.py:148-156), the finite-line port weights (Solver.source_weights,
wave.py:190-238) and the direct-solve fields and port data (wave.py:240-278).

Usage:  python tests/golden/make_golden.py   ->  tests/golden/theranostic_small.npz
"""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, "/root/reference")
from nsrno.theranostic.wave import Grid, Solver  # noqa: E402


def main():
    out = {}
    # Grid of the reference unit-test fixture at half the resolution
    # (tests/test_theranostic.py:38-43 uses 81 nodes on [-16,16] mm, 6 mm PML).
    axis = np.linspace(-16.0, 16.0, 41)
    solver = Solver(Grid(axis, axis, pml_mm=6.0))
    positions = np.array([[-6.4, -3.2], [-1.7, -3.1], [3.2, -3.2], [6.3, 1.7]])
    normals = np.array([[0.0, 1.0], [0.2, 1.0], [-0.15, 1.0], [-0.7, 0.7]])
    widths = np.array([0.8, 1.2, 0.6, 1.0])
    weights = solver.source_weights(positions, normals, widths)
    xx, zz = np.meshgrid(axis, axis)
    bump = np.maximum(1 - ((xx + 2.0) / 4.4) ** 2 - ((zz - 1.0) / 3.4) ** 2, 0) ** 2
    bump2 = np.maximum(1 - ((xx - 2.5) / 3.0) ** 2 - ((zz - 4.0) / 4.0) ** 2, 0) ** 2
    m_kappa = 0.12 * bump - 0.07 * bump2
    m_eta = -0.09 * bump + 0.11 * bump2
    frequency = 0.15
    matrix = solver.matrix(frequency, m_kappa, m_eta).tocoo()
    fields, diag = solver.solve(frequency, weights, m_kappa, m_eta, return_diagnostics=True)
    data = solver.sample(fields, weights)
    k0, mass, edge_stretch = solver.coefficients(frequency)
    w = weights.tocoo()
    out.update(axis=axis, pml_mm=6.0, frequency_mhz=frequency, c0=solver.c0_m_s,
               pml_strength=solver.pml_strength,
               positions=positions, normals=normals, widths=widths,
               m_kappa=m_kappa, m_eta=m_eta,
               matrix_row=matrix.row, matrix_col=matrix.col, matrix_val=matrix.data,
               weights_row=w.row, weights_col=w.col, weights_val=w.data,
               fields=fields, data=data, residual=diag["relative_linear_residual"],
               k0=k0, cell_mass=mass, edge_stretch=edge_stretch)
    # A zero-width point element grid (the Green's-function test's ports,
    # tests/test_theranostic.py:145-155) at 40 intervals.
    axis40 = np.linspace(-16.0, 16.0, 41)
    s40 = Solver(Grid(axis40, axis40, pml_mm=6.0))
    pos3 = np.array([[-3.2, 0.0], [3.2, 0.0], [0.0, 4.8]])
    w3 = s40.source_weights(pos3, np.tile([0.0, 1.0], (3, 1)), np.zeros(3))
    out.update(point_data=s40.sample(s40.solve(frequency, w3), w3))
    # A non-square pml-free grid with contrast touching the ghost edges.
    ax = np.linspace(0.0, 4.0, 9)
    az = np.linspace(0.0, 3.0, 7)
    s0 = Solver(Grid(ax, az, pml_mm=0.0), c0_m_s=1500.0)
    rng = np.random.default_rng(3)
    mk0 = 0.2 * rng.uniform(-1, 1, (7, 9))
    me0 = 0.2 * rng.uniform(-1, 1, (7, 9))
    m0 = s0.matrix(0.4, mk0, me0).tocoo()
    out.update(nopml_mk=mk0, nopml_me=me0, nopml_row=m0.row, nopml_col=m0.col, nopml_val=m0.data)
    path = Path(__file__).with_name("theranostic_small.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, path.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
