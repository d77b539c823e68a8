"""Host-side logic of the drop-in (CPU only): validation, port weights, layouts."""
from pathlib import Path

import numpy as np
import pytest
from scipy import sparse

from nsrno_amd import theranostic as T
from nsrno_amd.helmholtz import Ports
from oracle import helmholtz as O

GOLDEN = np.load(Path(__file__).parent / "golden" / "theranostic_small.npz")


def test_source_weights_match_reference_golden():
    axis = GOLDEN["axis"]
    solver = T.Solver(T.Grid(axis, axis, 6.0))
    w = solver.source_weights(GOLDEN["positions"], GOLDEN["normals"], GOLDEN["widths"])
    g = sparse.csc_matrix((GOLDEN["weights_val"], (GOLDEN["weights_row"], GOLDEN["weights_col"])), shape=w.shape)
    assert abs(w - g).max() < 1e-15
    # point elements (zero width) as in the Green's-function test
    pos = np.array([[-3.2, 0.0], [3.2, 0.0], [0.0, 4.8]])
    w3 = solver.source_weights(pos, np.tile([0.0, 1.0], (3, 1)), np.zeros(3))
    o3 = O.source_weights(O.Grid(axis, axis, 6.0), pos, np.tile([0.0, 1.0], (3, 1)), np.zeros(3))
    assert abs(w3 - o3).max() < 1e-15


@pytest.mark.parametrize("bad", [
    dict(positions_mm=np.zeros((2, 3)), normals=np.zeros((2, 3)), widths_mm=np.zeros(2)),
    dict(positions_mm=np.zeros((1, 2)), normals=np.zeros((1, 2)), widths_mm=np.ones(1)),
    dict(positions_mm=np.array([[-15.9, 0.0]]), normals=np.array([[0.0, 1.0]]), widths_mm=np.zeros(1)),
])
def test_source_weights_validation(bad):
    axis = GOLDEN["axis"]
    solver = T.Solver(T.Grid(axis, axis, 6.0))
    with pytest.raises(ValueError):
        solver.source_weights(**bad)


def test_grid_and_contrast_validation():
    with pytest.raises(ValueError):
        T.Grid(np.array([0.0, 1.0]), np.array([0.0, 1.0, 2.0]), 0.0)
    with pytest.raises(ValueError):
        T.Grid(np.linspace(0, 1, 5), np.linspace(0, 2, 5), 0.0)
    with pytest.raises(ValueError):
        T.Grid(np.linspace(0, 1, 5), np.linspace(0, 1, 5), 0.5)
    axis = GOLDEN["axis"]
    solver = T.Solver(T.Grid(axis, axis, 6.0))
    bad = np.zeros(solver.grid.shape)
    bad[20, 20] = -1.0
    with pytest.raises(ValueError):
        solver._contrast(bad, "m_kappa")
    bad[:] = 0
    bad[0, 0] = 0.01
    with pytest.raises(ValueError, match="vanish in the PML"):
        solver._contrast(bad, "m_eta")
    np.testing.assert_array_equal(solver.grid.pml_mask, O.Grid(axis, axis, 6.0).pml_mask)
    with pytest.raises(ValueError):
        solver.solve(0.15, 1j * np.ones((solver.grid.size, 1)))


def test_ports_layout_reproduces_weights():
    axis = GOLDEN["axis"]
    solver = T.Solver(T.Grid(axis, axis, 6.0))
    w = solver.source_weights(GOLDEN["positions"], GOLDEN["normals"], GOLDEN["widths"])
    p = Ports(w, device="cpu")
    rng = np.random.default_rng(0)
    amp = rng.normal(size=w.shape[1]) + 1j * rng.normal(size=w.shape[1])
    # node-major CSR gather equals W @ amp
    out = np.zeros(w.shape[0], complex)
    ptr, nodes = p.node_ptr.numpy(), p.node_idx.numpy()
    for q in range(p.n_nodes):
        s = slice(ptr[q], ptr[q + 1])
        out[nodes[q]] = np.sum(p.w_node.numpy()[s] * amp[p.port_of_tap.numpy()[s]])
    np.testing.assert_allclose(out, w @ amp, rtol=1e-15, atol=1e-15)
    # port-major CSC gather equals W^T u
    u = rng.normal(size=w.shape[0]) + 1j * rng.normal(size=w.shape[0])
    pp = p.port_ptr.numpy()
    tr = [np.sum(p.w_port.numpy()[pp[k]:pp[k + 1]] * u[p.node_of_tap.numpy()[pp[k]:pp[k + 1]]])
          for k in range(w.shape[1])]
    np.testing.assert_allclose(tr, w.T @ u, rtol=1e-14)


def test_coefficients_match_reference_golden():
    axis = GOLDEN["axis"]
    solver = T.Solver(T.Grid(axis, axis, 6.0))
    k0, mass, edge = solver.coefficients(float(GOLDEN["frequency_mhz"]))
    assert abs(k0 - GOLDEN["k0"]) / GOLDEN["k0"] < 1e-15
    np.testing.assert_allclose(mass, GOLDEN["cell_mass"], rtol=1e-15)
    np.testing.assert_allclose(edge, GOLDEN["edge_stretch"], rtol=1e-14)


def test_receiver_response_closed_forms():
    """conditioned_data.receiver_response restated: |H_bp|^2 of the bilinear Butterworth band-pass
    (closed form 1/(1 + r^8) in the prewarped frequency: zero-phase filtering squares the magnitude
    of the order-4 design) times the delay-compensated FIR response."""
    import numpy as np
    from scipy.signal import firwin
    from nsrno_amd.conditioned import receiver_response
    dt = 2.5e-9
    fs = 1 / dt
    taps = firwin(201, 0.1, window=("kaiser", 5.0))
    for f in (3e6, 4.5e6, 6e6, 7.4e6, 9e6):
        pre = lambda x: 2 * fs * np.tan(np.pi * x / fs)  # noqa: E731
        lo, hi, om = pre(4e6), pre(7.5e6), pre(f)
        ratio = (om ** 2 - lo * hi) / (om * (hi - lo))
        band2 = 1.0 / (1.0 + ratio ** 8)
        w = 2 * np.pi * f * dt
        fir = float(np.sum(taps * np.cos(w * (np.arange(201) - 100))))
        assert abs(receiver_response(f, dt) - band2 * fir) < 1e-9


def test_conditioned_geometry_and_validation_on_cpu():
    """Host logic of the dense-path mirror (no GPU needed): the padded box,
    mask, face row and receiver columns of FrequencyGeometry
    (conditioned_wave.py:12-31), its padding guard, and the medium-shape
    ValueError of ConditionedFrequencySystem (:54), raised before any plan."""
    import json
    import torch
    from nsrno_amd.conditioned import ConditionedFrequencySystem, FrequencyGeometry
    from oracle import conditioned as oc
    gold = np.load(Path(__file__).parent / "golden" / "conditioned_small.npz")
    spec = json.loads(str(gold["spec_json"]))
    g = FrequencyGeometry(spec, device="cpu")
    o = oc.Geometry(spec)
    assert (g.H, g.W, g.pz, g.px, g.face) == (o.H, o.W, o.pz, o.px, o.face)
    np.testing.assert_array_equal(g.mask.numpy(), gold["mask"])
    np.testing.assert_array_equal(g.rx_cols.numpy(), gold["rx_cols"])
    assert g.pad == (o.px, o.W - o.w - o.px, o.pz, o.H - o.h - o.pz)
    with pytest.raises(ValueError, match="FFT padding"):
        FrequencyGeometry({**spec, "fft_shape": [70, 128]}, device="cpu")
    with pytest.raises(ValueError, match="Expected native"):
        ConditionedFrequencySystem(g, torch.zeros(2, 10, 10, dtype=torch.float64), torch.zeros(10, 10), 5e6)
    # crop + replicate of the coefficients (conditioned_wave.py:45-49)
    box = g.coefficients(torch.tensor(np.stack((gold["c"], gold["rho"]))))
    np.testing.assert_array_equal(box[0].numpy(), gold["sc"])
    np.testing.assert_array_equal(box[1].numpy(), gold["srho"])
