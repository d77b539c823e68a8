"""GPU: the kernels around the solve against the CPU oracle.

* helm_capture_accumulate (nsrno/native_capture.cu:4-11) through
  FrequencyCapture over 24010 time steps (2401 captures, 32 per launch, a
  last partial batch): the fields equal the oracle's running DFT and numpy's
  FFT of the stored 40 MHz samples / 2401 at the bin frequencies.
* helm_band_rf (nsrno/wave_operator.py:169-180) against the oracle's
  irfft(norm='forward'), incl. a repeated bin, and predict_band_rf's shape.
* helm_born_jacobian (nsrno/theranostic/wave.py:331-389) on 24 ports and a
  bilinear basis of 36 functions against the oracle's sparse-product form.
"""
import numpy as np
import pytest
import torch

from oracle import helmholtz as O
from oracle import spectra as S

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def test_capture_matches_running_dft_and_fft():
    from nsrno_amd.spectra import FrequencyCapture
    rng = np.random.default_rng(3)
    shape, nt = (12, 40), 24010
    bins = np.array([1, 200, 360, 777, 1200])
    freqs = bins * 40e6 / 2401
    # a smooth band-limited signal plus noise, float32 like the native stepper's pressure
    t = np.arange(nt)[:, None, None] * 2.5e-9
    amp = rng.normal(size=(1,) + shape)
    steps = (amp * np.sin(2 * np.pi * 6e6 * t) + 0.1 * rng.normal(size=(nt,) + shape)).astype(np.float32)
    cap = FrequencyCapture(shape, freqs, nt)
    dev = torch.from_numpy(steps).cuda()
    for k in range(nt):
        cap.observe(dev[k])
    fields = cap.flush().cpu().numpy()
    assert cap.first == 2401
    ref = S.capture_fields(steps, freqs)
    assert rel(fields, ref) < 1e-12
    fft = np.fft.fft(steps[::10].astype(np.float64), axis=0)[bins] / 2401
    assert rel(fields, fft) < 1e-12


def test_band_rf_matches_irfft():
    from nsrno_amd.spectra import band_rf
    rng = np.random.default_rng(4)
    bins = np.array([3, 99, 360, 360, 1200, 640])
    x = (rng.normal(size=(192, 6)) + 1j * rng.normal(size=(192, 6))).astype(np.complex64)
    rf = band_rf(torch.from_numpy(x).cuda(), bins).cpu().numpy()
    ref = S.band_rf(x.astype(np.complex128), bins)
    assert rf.shape == (192, 2401) and rf.dtype == np.float32
    assert rel(rf, ref) < 1e-6


def test_predict_band_rf_shape_and_consistency():
    from pathlib import Path
    from nsrno_amd.spectra import band_rf
    from nsrno_amd.wave_operator import WaveResidualOperator
    G = np.load(Path(__file__).parent / "golden" / "spectral_small.npz")
    spec = dict(medium_shape=[int(G["h"]), int(G["w"])], dx_m=float(G["dx"]), x0_m=float(G["x0"]),
                z0_m=float(G["z0"]), angles_deg=list(G["angles"]), source_tref_s=list(G["trefs"]))
    model = WaveResidualOperator(spec, iterations=2)
    medium = torch.tensor(np.stack((G["c"], G["rho"]))[None], dtype=torch.float32)
    bins = [60, 75]
    rf = model.predict_band_rf(medium, 1, bins)
    assert tuple(rf.shape) == (192, 2401)
    preds = torch.stack([model(medium, b * 40e6 / 2401, 1)[0] for b in bins], dim=1)
    assert rel(rf.cpu().numpy(), band_rf(preds, bins).cpu().numpy()) < 1e-6
    with pytest.raises(ValueError):
        model.predict_band_rf(medium, 1, [0])


def test_born_jacobian_kernel_larger_case():
    from scipy import sparse
    from nsrno_amd import theranostic as T
    axis = np.linspace(-12.0, 12.0, 81)
    grid = T.Grid(axis, axis, pml_mm=3.0)
    solver = T.Solver(grid)
    ogrid = O.Grid(axis, axis, 3.0)
    xs = np.linspace(-8.0, 8.0, 24)
    pos = np.stack((xs, np.full(24, -8.0)), axis=1)
    w = solver.source_weights(pos, np.tile([0.0, 1.0], (24, 1)), np.full(24, 0.6))
    f = 0.5
    u0 = solver.solve(f, w)
    # bilinear hat functions on a 6 x 6 lattice inside the interior
    cx = np.linspace(-6.0, 6.0, 6)
    xx, zz = np.meshgrid(axis, axis)
    cols = []
    for z0 in cx:
        for x0 in cx:
            cols.append((np.maximum(1 - np.abs(xx - x0) / 2.4, 0) * np.maximum(1 - np.abs(zz - z0) / 2.4, 0)).ravel())
    basis = sparse.csr_matrix(np.column_stack(cols))
    tx, rx = np.arange(0, 24, 2), np.arange(24)
    jac = solver.born_jacobian(f, u0, tx, rx, basis)
    ref = O.born_jacobian(ogrid, f, u0, tx, rx, basis)
    assert jac.shape == ref.shape == (12 * 24, 72)
    assert rel(jac, ref) < 1e-12
