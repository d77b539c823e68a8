"""Pins the U-Net oracle's backward pass by central finite differences."""
import numpy as np

from oracle import unet as U


def test_unet_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    params = U.init_params(seed=1, out_scale=1.0) * 0.5
    params += 0.05 * rng.normal(size=params.size)  # nonzero biases
    x = rng.normal(size=(2, 8, 16, 32))
    gout = rng.normal(size=(2, 8, 16, 32))
    out, cache = U.forward(params, x)
    grad = U.backward(params, cache, gout)

    def loss(p):
        return float(np.sum(U.forward(p, x)[0] * gout))

    for layer in (U.C0A, U.C1B, U.C3A, U.C2C, U.C0C, U.COUT):
        for k in (5, 4000, U.WSZ + 3):  # two weights and one bias of the layer
            idx = layer * U.PSZ + k
            d = np.zeros_like(params)
            d[idx] = 1e-6
            fd = (loss(params + d) - loss(params - d)) / 2e-6
            assert abs(fd - grad[idx]) <= 1e-5 * max(1.0, abs(fd)), (layer, k, fd, grad[idx])


def test_unet_shapes_and_linearity_of_output_layer():
    params = U.init_params(seed=2)
    x = np.random.default_rng(3).normal(size=(1, 16, 8, 32))
    out, cache = U.forward(params, x)
    assert out.shape == (1, 16, 8, 32)
    assert cache["e"][3].shape == (1, 2, 1, 32)


def test_bf16_rounding_matches_torch():
    import torch
    x = np.random.default_rng(4).normal(size=1000) * 10.0 ** np.random.default_rng(5).integers(-5, 5, 1000)
    t = torch.as_tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(U.bf16(x), t)
