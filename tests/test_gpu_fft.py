"""GPU: the library's mixed-radix FFTs and the spectral (Fourier-multiplier)
preconditioner, and the reference's warm-start-only forward through them.

* helm_fft2 against numpy's FFT on boxes whose sides have radix-8/4/2 parts
  and odd prime factors (3, 11, 19, 37: register and generic kernels),
  forward and inverse, batched; 1216-wide rows go through the kernel that
  fuses both row passes, 1056-tall columns through the (3 x 11)(8 x 4) plan.
* HELM_PC_SPECTRAL: gain * ifft2(M fft2(v)) (adjoint) and its conjugate map
  against numpy, on boxes with radix 2-5 and radix 11 / 19 sides (the
  multiply, gain and conjugates folded into the first and last passes).
* WaveResidualOperator.forward (nsrno/wave_operator.py:137-154) on the
  reference's golden case: the default model and a model loaded with
  perturbed context net, spectral modulation and shift; gain maps against
  the reference's, predictions, losses and traces against the reference's
  (complex64) and the oracle's (complex128) forward.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu
G = np.load(Path(__file__).parent / "golden" / "spectral_small.npz")


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("H,W", [(32, 96), (1056 // 8, 1216 // 16), (64, 2 * 37), (88, 1216), (1056, 48), (24, 44),
                                 (2 * 31, 2 * 29), (2 * 23, 2 * 17), (4 * 41, 8)])
def test_fft2_matches_numpy(H, W, precision):
    from nsrno_amd import _lib
    rng = np.random.default_rng(H * W)
    x = rng.normal(size=(3, H, W)) + 1j * rng.normal(size=(3, H, W))
    x = x.astype(np.complex64 if precision == 32 else np.complex128)
    src = torch.from_numpy(x).cuda()
    out = torch.empty_like(src)
    work = torch.empty_like(src)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for inverse, ref in ((0, np.fft.fft2(x.astype(np.complex128))),
                         (1, np.fft.ifft2(x.astype(np.complex128)) * H * W)):
        assert _lib.lib().helm_fft2(H, W, 3, inverse, precision, ctypes.c_void_p(src.data_ptr()),
                                    ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(work.data_ptr()), s) == 0
        torch.cuda.synchronize()
        assert rel(out.cpu().numpy(), ref) < (3e-6 if precision == 32 else 1e-14), (H, W, inverse)


@pytest.mark.parametrize("H,W", [(96, 160), (132, 152)])
def test_spectral_preconditioner_apply(H, W):
    from nsrno_amd import _lib
    from nsrno_amd.helmholtz import HelmholtzPlan
    plan = HelmholtzPlan(H, W, 0.05, 3.0, 0.0, nrhs=2, restart=4)
    plan.set_medium(np.ones(H * W), np.ones(H * W, dtype=complex))
    rng = np.random.default_rng(2)
    m = rng.normal(size=(H, W)) + 1j * rng.normal(size=(H, W))
    g = rng.normal(size=(H, W)) + 1j * rng.normal(size=(H, W))
    plan.set_spectral_preconditioner(m, g)
    v = (rng.normal(size=(2, H, W)) + 1j * rng.normal(size=(2, H, W))).astype(np.complex64)
    vt = torch.from_numpy(v).cuda()
    # a flux-form plan: forward solves take the map as given, adjoint solves its conjugate map
    z = plan.precondition(vt, kind=_lib.PC_SPECTRAL, adjoint=False).cpu().numpy()
    ref = g * np.fft.ifft2(m * np.fft.fft2(v.astype(np.complex128)))
    assert rel(z, ref) < 1e-5
    zc = plan.precondition(vt, kind=_lib.PC_SPECTRAL, adjoint=True).cpu().numpy()
    assert rel(zc, np.conj(g * np.fft.ifft2(m * np.fft.fft2(np.conj(v.astype(np.complex128)))))) < 1e-5
    plan.close()


def _model(**state):
    from nsrno_amd.wave_operator import WaveResidualOperator
    spec = dict(medium_shape=[int(G["h"]), int(G["w"])], dx_m=float(G["dx"]), x0_m=float(G["x0"]),
                z0_m=float(G["z0"]), angles_deg=list(G["angles"]), source_tref_s=list(G["trefs"]))
    model = WaveResidualOperator(spec, iterations=4)
    if state:
        model.load_state_dict(state)
    return model


def test_forward_matches_reference_default_model():
    model = _model()
    medium = torch.tensor(np.stack((G["c"], G["rho"]))[None], dtype=torch.float32)
    f = float(G["f"])
    pred, loss, trace = model.forward(medium, f, 1, return_trace=True)
    box = S.Box(int(G["h"]), int(G["w"]), float(G["dx"]), float(G["x0"]), float(G["z0"]))
    opred, oloss, otrace = S.forward(box, np.float32(G["c"]), np.float32(G["rho"]), f,
                                     float(G["angles"][1]) * np.pi / 180, float(G["trefs"][1]), iterations=4)
    print(f"trace {trace} reference {list(G['fwd_trace'])} oracle {otrace}")
    assert np.allclose(trace, otrace, rtol=1e-3)
    assert np.allclose(trace, G["fwd_trace"], rtol=1e-3)
    assert abs(float(loss) - oloss) < 2e-3 * oloss
    # the receiver row sits ~1e-4 below the field maximum: complex64 rounding is a few % there
    assert rel(pred.cpu().numpy(), opred) < 5e-2
    assert rel(pred.cpu().numpy(), G["fwd_pred"]) < 5e-2


def test_forward_with_learned_parts_matches_reference():
    state = {"context.0.weight": G["learned_context_0_weight"], "context.0.bias": G["learned_context_0_bias"],
             "context.2.weight": G["learned_context_2_weight"], "context.2.bias": G["learned_context_2_bias"],
             "spectral": G["learned_spectral"], "log_shift": G["learned_log_shift"]}
    model = _model(**{k: torch.from_numpy(np.asarray(v)) for k, v in state.items()})
    medium = torch.tensor(np.stack((G["c"], G["rho"]))[None], dtype=torch.float32)
    f = float(G["f"])
    gain, mult = model.preconditioner_maps(np.float32(G["c"]), np.float32(G["rho"]), f)
    assert rel(gain, G["learned_gain"]) < 1e-5
    assert rel(mult, G["learned_mult"]) < 1e-5
    pred, loss, trace = model.forward(medium, f, 1, return_trace=True)
    print(f"trace {trace} reference {list(G['learned_trace'])}")
    assert np.allclose(trace, G["learned_trace"], rtol=1e-3)
    assert abs(float(loss) - float(G["learned_loss"])) < 2e-3 * float(G["learned_loss"])
    assert rel(pred.cpu().numpy(), G["learned_pred"]) < 5e-2
