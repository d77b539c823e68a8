"""GPU: helm_spectral_incident_noalloc (the incident field with a static
device work buffer, for callers that cannot allocate) against
helm_spectral_incident, through the C ABI: bit-identical on two calls in a
row, the second on a smaller box, and boxes wider than the 4096-node
buffer rejected."""
import ctypes
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def plan_for(h, w, dx=2e-4, f=2.5e6):
    from nsrno_amd.wave_operator import WaveResidualOperator
    spec = dict(medium_shape=[h, w], dx_m=dx, x0_m=-0.0191 - 2 * dx, z0_m=-3 * dx, angles_deg=[-8.0, 0.0, 8.0],
                source_tref_s=[0.0, 0.0, 0.0])
    model = WaveResidualOperator(spec, iterations=1)
    rng = np.random.default_rng(h * 1000 + w)
    c = 1540.0 + 20.0 * rng.uniform(-1, 1, (h, w))
    rho = 1000.0 + 30.0 * rng.uniform(-1, 1, (h, w))
    data, _, _ = model.setup(torch.tensor(np.stack((c, rho))[None], dtype=torch.float32), f)
    return model, data


def call(name, plan, sin_theta, out):
    from nsrno_amd import _lib
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return getattr(_lib.lib(), name)(plan._h, sin_theta, 0.0191, ctypes.c_void_p(out.data_ptr()), stream)


def test_noalloc_incident_matches_and_rejects_oversize():
    from nsrno_amd import _lib
    cases = [(40, 200, math.sin(math.radians(8.0))), (24, 96, math.sin(math.radians(-5.0)))]
    built = [plan_for(h, w) for h, w, _ in cases]
    refs = []
    for (h, w, s), (model, data) in zip(cases, built):
        ref = torch.empty((model.H, model.W), dtype=torch.complex128, device="cuda")
        assert call("helm_spectral_incident", data.plan, s, ref) == 0
        refs.append(ref)
    outs = []
    for (h, w, s), (model, data) in zip(cases, built):  # two calls in a row, the second smaller
        out = torch.full((model.H, model.W), complex("nan"), dtype=torch.complex128, device="cuda")
        assert call("helm_spectral_incident_noalloc", data.plan, s, out) == 0
        outs.append(out)
    torch.cuda.synchronize()
    assert built[1][0].W < built[0][0].W
    for ref, out in zip(refs, outs):
        assert torch.isfinite(out).all()
        assert torch.equal(out, ref)
    # a box wider than the static buffer is rejected before any launch
    model, data = plan_for(4, 4100, dx=1e-5)
    assert model.W > 4096
    out = torch.empty((model.H, model.W), dtype=torch.complex128, device="cuda")
    assert call("helm_spectral_incident_noalloc", data.plan, 0.0, out) == 1  # HELM_ERR_ARG
    assert b"4096" in _lib.lib().helm_last_error()
