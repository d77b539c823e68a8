"""CPU oracle for the time -> frequency capture and the band RF synthesis.

TEST INFRASTRUCTURE ONLY: imported by tests (and nothing in nsrno_amd).
Plain NumPy in float64/complex128.

* ``capture_phase``  nsrno/native_capture.py:37-40 (phase of the 40 MHz
                     output clock, / 2401)
* ``capture_fields`` nsrno/native_capture.cu:4-11 and :26-31 (accumulate
                     every stride-th pressure snapshot times its phase)
* ``band_rf``        nsrno/wave_operator.py:169-180 (spectrum zero except at
                     the given bins, later duplicates overwriting earlier
                     ones, irfft with n = 2401 and norm 'forward')

Pinned by: at the bin frequencies k * 40 MHz / 2401 over 2401 captures the
captured fields equal numpy's FFT of the stored time samples divided by 2401
(``tests/test_oracle_spectra.py``); the RF synthesis equals the direct
Hermitian sum of its bins, and numpy's rfft of it recovers the spectrum.
"""
import numpy as np


def capture_phase(frequencies_hz, steps, stride=10, dt=2.5e-9, samples=2401):
    times = np.arange(0, steps, stride, dtype=np.float64) * dt
    return np.exp(-2j * np.pi * np.asarray(frequencies_hz, dtype=np.float64)[:, None] * times[None, :]) / samples


def capture_fields(pressure_steps, frequencies_hz, stride=10, dt=2.5e-9, samples=2401):
    """pressure_steps [nt, ...] (one field per time step) -> fields [F, ...]."""
    p = np.asarray(pressure_steps, dtype=np.float64)
    nt = p.shape[0]
    phase = capture_phase(frequencies_hz, nt, stride, dt, samples)
    fields = np.zeros((phase.shape[0],) + p.shape[1:], dtype=np.complex128)
    for t in range(nt):
        if t % stride == 0:
            c = t // stride
            fields += phase[:, c].reshape((-1,) + (1,) * (p.ndim - 1)) * p[t][None]
    return fields


def band_rf(spectra, bins, samples=2401):
    """spectra [R, B] at positive bins [B] -> RF [R, samples] (float64)."""
    spectra = np.asarray(spectra)
    full = np.zeros((spectra.shape[0], samples // 2 + 1), dtype=np.complex128)
    for k, b in enumerate(np.asarray(bins)):
        full[:, int(b)] = spectra[:, k]
    return np.fft.irfft(full, n=samples, axis=-1, norm="forward")
