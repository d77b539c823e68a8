"""Learned preconditioner network on the GPU (helm_net of libhelm.so).

``LearnedNet`` wraps one ``helm_net``: a 4-level, 32-channel 3x3-conv U-Net
run by BF16 MFMA kernels with FP32 accumulation.  Parameters are an FP32
torch tensor [12 * (9*32*32 + 32)] on the GPU owned by this object; Adam
moments live beside them.  Attached to a ``HelmholtzPlan`` it turns the
flexible GMRES direction into z = MG(v) + N(v) (``precond="learned"``).
The reference's learned preconditioner (nsrno/conditioned_wave.py:94-117)
and its Adam training loop (nsrno/train_preconditioner.py:63-78) are the
precedent; see train_preconditioner.py for the harness.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib

NL = 12
WSZ = 9 * 32 * 32
PSZ = WSZ + 32


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def canvas(n):
    """Smallest multiple of 8 >= n."""
    return (int(n) + 7) // 8 * 8


def init_params(seed=0, out_scale=0.1):
    rng = np.random.default_rng(seed)
    p = np.zeros((NL, PSZ))
    for layer in range(NL):
        w = rng.normal(size=(9, 32, 32)) * np.sqrt(2.0 / (9 * 32))
        if layer == NL - 1:
            w *= out_scale
        p[layer, :WSZ] = w.ravel()
    return p.ravel()


class LearnedNet:
    def __init__(self, H, W, batch, training=False, params=None, device="cuda", linear=False):
        if not torch.cuda.is_available():
            raise RuntimeError("LearnedNet needs a GPU")
        self.H, self.W, self.batch = int(H), int(W), int(batch)
        self.device = torch.device(device)
        h = ctypes.c_void_p()
        check(lib().helm_net_create(self.H, self.W, self.batch, int(bool(training)), ctypes.byref(h)),
              "helm_net_create")
        self._h = h
        self.linear = bool(linear)
        check(lib().helm_net_set_linear(h, int(self.linear)), "helm_net_set_linear")
        assert lib().helm_net_param_count() == NL * PSZ
        p = init_params() if params is None else params
        self.params = torch.as_tensor(np.asarray(p, dtype=np.float32), device=self.device).contiguous()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.step_count = 0
        self.sync_params()

    def close(self):
        if getattr(self, "_h", None):
            lib().helm_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync_params(self):
        check(lib().helm_net_set_params(self._h, _p(self.params), _s()), "helm_net_set_params")

    def forward(self, features):
        """features: bf16 [B, H, W, 32] GPU tensor -> bf16 [B, H, W, 32]."""
        b = features.shape[0]
        out = torch.empty_like(features)
        check(lib().helm_net_forward(self._h, _p(features.contiguous()), _p(out), b, _s()), "helm_net_forward")
        return out

    def backward(self, dout):
        """Parameter gradient (fp32, in self.grad) of the last forward for dL/dout (bf16)."""
        check(lib().helm_net_backward(self._h, _p(dout.contiguous()), _p(self.grad), dout.shape[0], _s()),
              "helm_net_backward")
        return self.grad

    def adam(self, lr=2e-3, beta1=0.9, beta2=0.999, eps=1e-8, clip=None):
        """One Adam step on self.grad (global-norm clipping as in train_preconditioner.py:75)."""
        scale = 1.0
        if self.linear:  # biases stay zero so that N is linear
            self.grad.view(NL, PSZ)[:, WSZ:] = 0
        if clip is not None:
            gn = float(torch.linalg.vector_norm(self.grad))
            if gn > clip:
                scale = clip / gn
        self.step_count += 1
        check(lib().helm_adam_step(self.params.numel(), _p(self.params), _p(self.grad), _p(self.m), _p(self.v),
                                   self.step_count, lr, beta1, beta2, eps, scale, _s()), "helm_adam_step")
        self.sync_params()

    def save(self, path):
        np.savez_compressed(path, params=self.params.cpu().numpy(), H=self.H, W=self.W, linear=self.linear)

    @staticmethod
    def load_params(path):
        with np.load(path, allow_pickle=False) as z:
            return z["params"].astype(np.float32)

    @staticmethod
    def load_linear(path):
        with np.load(path, allow_pickle=False) as z:
            return bool(z["linear"]) if "linear" in z.files else False


def attach(plan, net: LearnedNet | None):
    """Use `net` for precond="learned" solves of `plan` (None detaches)."""
    check(lib().helm_plan_set_net(plan._h, None if net is None else net._h), "helm_plan_set_net")
    plan._net = net


def features(plan, v):
    """Net input features of residuals v (complex64 [R, nz, nx]) -> bf16 [R, H, W, 32]."""
    net = plan._net
    r = v.shape[0]
    out = torch.empty((r, net.H, net.W, 32), dtype=torch.bfloat16, device=v.device)
    check(lib().helm_net_features(plan._h, _p(v), _p(out), r, _s()), "helm_net_features")
    return out


def correction(plan, v):
    """N(v): the net's field correction for residuals v (complex64 [R, nz, nx])."""
    z = torch.empty_like(v)
    check(lib().helm_net_correction(plan._h, _p(v), _p(z), v.shape[0], _s()), "helm_net_correction")
    return z
