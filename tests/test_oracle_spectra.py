"""CPU pins of oracle/spectra.py (capture running DFT, band RF synthesis)."""
import numpy as np

from oracle import spectra as S


def test_capture_equals_fft_of_stored_samples():
    """Over 2401 captures (24010 steps of 2.5 ns, every 10th captured) the
    running DFT at the bin frequencies k * 40 MHz / 2401 is numpy's FFT of the
    stored 40 MHz samples divided by 2401 (native_capture.py:37-40)."""
    rng = np.random.default_rng(0)
    nt, stride = 24010, 10
    steps = rng.normal(size=(nt, 3, 5)).astype(np.float32)
    bins = np.array([1, 17, 360, 600, 1200])
    fields = S.capture_fields(steps, bins * 40e6 / 2401, stride=stride)
    ref = np.fft.fft(steps[::stride].astype(np.float64), axis=0)[bins] / 2401
    assert np.abs(fields - ref).max() < 1e-12 * np.abs(ref).max()


def test_capture_skips_uncaptured_steps_and_is_linear():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(95, 4))
    b = rng.normal(size=(95, 4))
    f = [3.1e6, 5.7e6]
    fa, fb = S.capture_fields(a, f), S.capture_fields(b, f)
    assert np.allclose(S.capture_fields(2 * a - b, f), 2 * fa - fb, atol=1e-14)
    c = a.copy()
    c[np.arange(95) % 10 != 0] = 1e9  # only every 10th step enters
    assert np.allclose(S.capture_fields(c, f), fa, atol=1e-14)


def test_band_rf_is_the_hermitian_sum_and_inverts():
    rng = np.random.default_rng(2)
    bins = np.array([5, 80, 1200, 333])
    x = rng.normal(size=(3, 4)) + 1j * rng.normal(size=(3, 4))
    rf = S.band_rf(x, bins)
    t = np.arange(2401)
    direct = sum(2 * np.real(x[:, k, None] * np.exp(2j * np.pi * b * t / 2401)) for k, b in enumerate(bins))
    assert np.abs(rf - direct).max() < 1e-10
    back = np.fft.rfft(rf, axis=-1, norm="forward")
    assert np.allclose(back[:, bins], x, atol=1e-12)
    mask = np.ones(1201, bool)
    mask[bins] = False
    assert np.abs(back[:, mask]).max() < 1e-12
    # a repeated bin keeps its last spectrum, as the reference's assignment does
    rf2 = S.band_rf(np.concatenate([x[:, :1] * 7, x], axis=1), np.concatenate([[5], bins]))
    assert np.allclose(rf2, rf)
