"""Frequency capture from time snapshots, band-limited RF synthesis and 2-D
FFTs on the GPU.

* ``fft2`` / ``ifft2`` are the library's mixed-radix FFTs (helm_fft2), used
  where the reference calls torch.fft on the path.

* ``FrequencyCapture`` replaces the running DFT of nsrno/native_capture.cu
  (accumulate_pressure, :4-11) and its driver nsrno/native_capture.py:29-59:
  while a caller's time stepper runs, every ``stride``-th pressure snapshot
  is added to complex128 frequency fields with the phase
  exp(-2 pi i f t_c) / 2401 of the 40 MHz output clock (native_capture.py:37-40).
  Snapshots are buffered on the device and accumulated 32 at a time by
  helm_capture_accumulate, so the fields are read and written once per 32
  captures.  The time stepper itself (native_time.cu) is not part of this
  package; any float32 [nz, nx] pressure the caller produces can be added.
* ``band_rf`` replaces the inverse real FFT of
  WaveResidualOperator.predict_band_rf (nsrno/wave_operator.py:169-180):
  receiver spectra at a set of positive bins -> RF traces of 2401 samples,
  irfft(..., norm='forward') of a spectrum that is zero elsewhere
  (helm_band_rf, FP64 sums of an exact integer-phase twiddle table).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib

SAMPLES = 2401            # RF samples of one acquisition (the reference's V5 clock)
CLOCK_HZ = 40e6           # output clock
STEP_S = 2.5e-9           # native time step
STRIDE = 10               # time steps per capture (40 MHz / 2.5 ns)


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def fft2(x, inverse=False):
    """2-D DFT over the last two axes of a complex64/complex128 GPU tensor
    (helm_fft2: mixed-radix Stockham passes; inverse unscaled, as e^{+i})."""
    if x.device.type != "cuda" or x.dtype not in (torch.complex64, torch.complex128) or x.dim() < 2:
        raise ValueError("fft2 takes a complex64 or complex128 GPU tensor [..., H, W]")
    H, W = x.shape[-2:]
    src = x.contiguous()
    out = torch.empty_like(src)
    work = torch.empty_like(src)
    batch = src.numel() // (H * W)
    check(lib().helm_fft2(int(H), int(W), int(batch), int(bool(inverse)), 32 if x.dtype == torch.complex64 else 64,
                          _p(src), _p(out), _p(work), _s()), "helm_fft2")
    return out


def ifft2(x):
    """Inverse 2-D DFT with the 1/(H W) of numpy/torch's ifft2."""
    return fft2(x, inverse=True) / (x.shape[-1] * x.shape[-2])


def capture_phase(frequencies_hz, steps, stride=STRIDE, dt=STEP_S, samples=SAMPLES, device="cuda"):
    """exp(-2 pi i f t_c) / samples at the capture times t_c = c stride dt
    (native_capture.py:37-40) -> complex128 [F, captures]."""
    times = np.arange(0, steps, stride, dtype=np.float64) * dt
    ph = np.exp(-2j * np.pi * np.asarray(frequencies_hz, dtype=np.float64)[:, None] * times[None, :]) / samples
    return torch.from_numpy(ph).to(device)


class FrequencyCapture:
    """Running DFT of pressure snapshots into frequency fields [F, nz, nx] c128."""

    BATCH = 32

    def __init__(self, shape, frequencies_hz, steps, stride=STRIDE, dt=STEP_S, samples=SAMPLES, device="cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("FrequencyCapture needs a GPU (no CPU fallback)")
        self.shape = tuple(int(v) for v in shape)
        self.n = int(np.prod(self.shape))
        self.stride, self.steps = int(stride), int(steps)
        self.phase = capture_phase(frequencies_hz, steps, stride, dt, samples, device).contiguous()
        self.F, self.captures = self.phase.shape
        self.fields = torch.zeros((self.F,) + self.shape, dtype=torch.complex128, device=device)
        self.buffer = torch.empty((self.BATCH,) + self.shape, dtype=torch.float32, device=device)
        self.held = 0        # snapshots in the buffer
        self.first = 0       # capture index of buffer[0]
        self.step = 0        # time steps seen

    def observe(self, pressure):
        """Call once per time step with that step's pressure (float32 [nz, nx]);
        every stride-th step is captured."""
        if self.step % self.stride == 0:
            if self.first + self.held >= self.captures:
                raise ValueError("more time steps than the capture was set up for")
            p = torch.as_tensor(pressure, device=self.buffer.device)
            if tuple(p.shape) != self.shape or p.dtype != torch.float32:
                raise ValueError(f"pressure must be float32 {self.shape}")
            self.buffer[self.held].copy_(p)
            self.held += 1
            if self.held == self.BATCH:
                self.flush()
        self.step += 1

    def add(self, snapshots, first):
        """Accumulate captured snapshots float32 [K, nz, nx] (K <= 32) whose first
        is capture number ``first``."""
        k = snapshots.shape[0]
        if snapshots.dtype != torch.float32 or tuple(snapshots.shape[1:]) != self.shape or not 1 <= k <= self.BATCH:
            raise ValueError(f"snapshots must be float32 [1..{self.BATCH}, {self.shape}]")
        if first < 0 or first + k > self.captures:
            raise ValueError("capture index out of range")
        snap = snapshots.contiguous()
        check(lib().helm_capture_accumulate(_p(snap), k, _p(self.phase), int(first), self.captures, self.F, self.n,
                                            _p(self.fields), _s()), "helm_capture_accumulate")

    def flush(self):
        if self.held:
            self.add(self.buffer[:self.held], self.first)
            self.first += self.held
            self.held = 0
        return self.fields


def band_rf(spectra, bins, samples=SAMPLES):
    """RF traces float32 [R, samples] from spectra [R, B] (complex) at the
    positive bins ``bins`` [B]; as the reference, a bin given twice keeps its
    last spectrum (wave_operator.py:175-178)."""
    bins = np.asarray(bins, dtype=np.int64)
    spectra = torch.as_tensor(spectra)
    if spectra.dim() != 2 or spectra.shape[1] != bins.size or bins.ndim != 1:
        raise ValueError("spectra must be [R, len(bins)]")
    if np.any(bins < 0) or np.any(2 * bins > samples) or samples > 8192:
        raise ValueError("bins must lie in [0, samples/2] (samples <= 8192)")
    last = {int(b): k for k, b in enumerate(bins)}  # later entries overwrite earlier ones
    keep = np.array(sorted(last.values()))
    dev = spectra.device if spectra.device.type == "cuda" else torch.device("cuda")
    x = spectra[:, torch.from_numpy(keep).to(spectra.device)].to(dev, torch.complex64).contiguous()
    b = torch.from_numpy(bins[keep].astype(np.int32)).to(dev)
    rf = torch.empty((x.shape[0], samples), dtype=torch.float32, device=dev)
    check(lib().helm_band_rf(_p(x), _p(b), int(b.numel()), int(x.shape[0]), int(samples), _p(rf), _s()),
          "helm_band_rf")
    return rf
