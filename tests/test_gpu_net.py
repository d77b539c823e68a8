"""GPU: BF16 MFMA U-Net forward/backward against the float64 oracle."""
import numpy as np
import pytest
import torch

from oracle import unet as U

pytestmark = pytest.mark.gpu


def bf16_round(a):
    return torch.as_tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def setup(B=2, H=16, W=64, seed=0):
    from nsrno_amd.learned import LearnedNet
    rng = np.random.default_rng(seed)
    params = U.init_params(seed=seed + 1, out_scale=1.0) * 0.7
    params += 0.02 * rng.normal(size=params.size)
    params = bf16_round(params)  # the device computes with bf16 weights (biases stay fp32)
    x = bf16_round(rng.normal(size=(B, H, W, 32)))
    net = LearnedNet(H, W, B, training=True, params=params.astype(np.float32))
    return net, params, x


def test_net_forward_matches_oracle():
    net, params, x = setup()
    out = net.forward(torch.from_numpy(x).to(torch.bfloat16).cuda()).to(torch.float64).cpu().numpy()
    ref, _ = U.forward(params, x)
    assert rel(out, ref) < 3e-2
    # with the device's BF16 storage emulated the remaining difference is FP32 accumulation
    assert rel(out, U.forward(params, x, U.bf16)[0]) < 1e-2
    # odd-sized tiles: width not a multiple of 32, height not of 16
    net2, params2, x2 = setup(B=1, H=24, W=40, seed=3)
    out2 = net2.forward(torch.from_numpy(x2).to(torch.bfloat16).cuda()).to(torch.float64).cpu().numpy()
    assert rel(out2, U.forward(params2, x2)[0]) < 3e-2


def test_net_backward_matches_oracle():
    net, params, x = setup()
    net.forward(torch.from_numpy(x).to(torch.bfloat16).cuda())
    rng = np.random.default_rng(9)
    gout = bf16_round(rng.normal(size=x.shape))
    grad = net.backward(torch.from_numpy(gout).to(torch.bfloat16).cuda()).to(torch.float64).cpu().numpy()
    # BF16 storage of activations and activation gradients emulated in the oracle
    _, cache = U.forward(params, x, U.bf16)
    ref = U.backward(params, cache, gout, U.bf16)
    g = grad.reshape(U.NL, U.PSZ)
    r = ref.reshape(U.NL, U.PSZ)
    for layer in range(U.NL):
        cos = np.dot(g[layer], r[layer]) / (np.linalg.norm(g[layer]) * np.linalg.norm(r[layer]))
        assert cos > 0.998, (layer, cos)
        assert rel(g[layer], r[layer]) < 0.05, (layer, rel(g[layer], r[layer]))
    # and without emulation the FP64 gradient is still close in direction
    _, cache64 = U.forward(params, x)
    r64 = U.backward(params, cache64, gout).reshape(U.NL, U.PSZ)
    for layer in range(U.NL):
        cos = np.dot(g[layer], r64[layer]) / (np.linalg.norm(g[layer]) * np.linalg.norm(r64[layer]))
        assert cos > 0.985, (layer, cos)


def test_adam_step_matches_formula():
    net, params, x = setup(B=1, H=8, W=32)
    net.grad.copy_(torch.linspace(-1, 1, net.grad.numel(), device="cuda"))
    p0 = net.params.clone()
    net.adam(lr=1e-2)
    g = torch.linspace(-1, 1, net.grad.numel(), device="cuda")
    m = 0.1 * g
    v = 0.001 * g * g
    expect = p0 - 1e-2 * (m / 0.1) / (torch.sqrt(v / 0.001) + 1e-8)
    assert torch.allclose(net.params, expect, rtol=1e-5, atol=1e-6)
