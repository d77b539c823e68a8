"""CPU oracle of the learned preconditioner's U-Net (forward and backward).

TEST INFRASTRUCTURE ONLY: the checker of the BF16 MFMA kernels in
nsrno_amd/csrc/helm_net.hip.  Float64 NumPy, NHWC, 32 channels, 3x3 convs
with weights [tap][out][in], tap = (dy+1)*3 + (dx+1), zero padding:

    y[b,i,j,o] = bias[o] + sum_t sum_c W[t,o,c] x[b, i+dy, j+dx, c]

Graph (the same as the device net):
    a0 = relu(C0a(in)); e0 = relu(C0b(a0)); for l = 1..3:
    a_l = relu(C_la(pool e_{l-1})); e_l = relu(C_lb(a_l));
    d2 = relu(C2c(up e3 + e2)); d1 = relu(C1c(up d2 + e1)); d0 = relu(C0c(up d1 + e0));
    out = Cout(d0)
Layer order in the parameter vector: C0a C0b C1a C1b C2a C2b C3a C3b C2c C1c C0c Cout,
each [9*32*32 weights | 32 biases].

Pinned by: central finite differences of a scalar loss against ``backward``
(float64, tiny image) in tests/test_oracle_unet.py.
"""
from __future__ import annotations

import numpy as np

NL = 12
WSZ = 9 * 32 * 32
PSZ = WSZ + 32
C0A, C0B, C1A, C1B, C2A, C2B, C3A, C3B, C2C, C1C, C0C, COUT = range(NL)


def bf16(a):
    """Round float64 values to the nearest bfloat16 (ties to even), as float64.

    Emulates the device's BF16 storage of every activation and activation
    gradient (accumulation stays in FP32 there, FP64 here)."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _id(a):
    return a


def split(params):
    p = np.asarray(params, dtype=np.float64).reshape(NL, PSZ)
    return [p[l, :WSZ].reshape(9, 32, 32) for l in range(NL)], [p[l, WSZ:] for l in range(NL)]


def conv(x, w, b):
    B, H, W, _ = x.shape
    xp = np.pad(x, [(0, 0), (1, 1), (1, 1), (0, 0)])
    y = np.zeros((B, H, W, 32)) + (0 if b is None else b)
    for t in range(9):
        dy, dx = t // 3, t % 3
        y += np.einsum("bijc,oc->bijo", xp[:, dy:dy + H, dx:dx + W], w[t])
    return y


def conv_bwd(x, w, gy):
    """(dx, dW, db) of y = conv(x, w, b) given dL/dy."""
    B, H, W, _ = x.shape
    xp = np.pad(x, [(0, 0), (1, 1), (1, 1), (0, 0)])
    dW = np.zeros_like(w)
    dxp = np.zeros_like(xp)
    for t in range(9):
        dy, dx = t // 3, t % 3
        dW[t] = np.einsum("bijo,bijc->oc", gy, xp[:, dy:dy + H, dx:dx + W])
        dxp[:, dy:dy + H, dx:dx + W] += np.einsum("bijo,oc->bijc", gy, w[t])
    return dxp[:, 1:-1, 1:-1], dW, gy.sum(axis=(0, 1, 2))


def pool(x):
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 1::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 1::2])


def pool_bwd(g):
    return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(g):
    return g[:, 0::2, 0::2] + g[:, 1::2, 0::2] + g[:, 0::2, 1::2] + g[:, 1::2, 1::2]


def relu(x):
    return np.maximum(x, 0)


def forward(params, x, rnd=_id):
    """rnd rounds every stored tensor (``bf16`` emulates the device)."""
    W, b = split(params)
    a, e, p, s, d = {}, {}, {}, {}, {}
    a[0] = rnd(relu(conv(x, W[C0A], b[C0A])))
    e[0] = rnd(relu(conv(a[0], W[C0B], b[C0B])))
    for l in range(1, 4):
        p[l] = rnd(pool(e[l - 1]))
        a[l] = rnd(relu(conv(p[l], W[C0A + 2 * l], b[C0A + 2 * l])))
        e[l] = rnd(relu(conv(a[l], W[C0B + 2 * l], b[C0B + 2 * l])))
    coarse = e[3]
    for l, layer in ((2, C2C), (1, C1C), (0, C0C)):
        s[l] = rnd(up(coarse) + e[l])
        d[l] = rnd(relu(conv(s[l], W[layer], b[layer])))
        coarse = d[l]
    out = rnd(conv(d[0], W[COUT], b[COUT]))
    return out, dict(x=x, a=a, e=e, p=p, s=s, d=d)


def backward(params, cache, gout, rnd=_id):
    """Gradient of the parameters [NL, PSZ] given dL/dout; rnd rounds every
    stored activation gradient as the device does."""
    W, _ = split(params)
    grad = np.zeros((NL, PSZ))

    def put(layer, dW, db):
        grad[layer, :WSZ] = dW.ravel()
        grad[layer, WSZ:] = db

    a, e, p, s, d = cache["a"], cache["e"], cache["p"], cache["s"], cache["d"]
    gd, dW, db = conv_bwd(d[0], W[COUT], gout)
    put(COUT, dW, db)
    gd = rnd(gd) * (d[0] > 0)
    skip = {}
    for l, layer in ((0, C0C), (1, C1C), (2, C2C)):
        gs, dW, db = conv_bwd(s[l], W[layer], gd)
        put(layer, dW, db)
        skip[l] = rnd(gs)
        act = d[l + 1] if l < 2 else e[3]
        gd = rnd(up_bwd(skip[l])) * (act > 0)
    ge = gd  # level 3
    pool_grad = None
    for l in (3, 2, 1, 0):
        if l < 3:
            ge = rnd(skip[l] + pool_grad) * (e[l] > 0)
        ga, dW, db = conv_bwd(a[l], W[C0B + 2 * l], ge)
        put(C0B + 2 * l, dW, db)
        ga = rnd(ga) * (a[l] > 0)
        xin = cache["x"] if l == 0 else p[l]
        gp, dW, db = conv_bwd(xin, W[C0A + 2 * l], ga)
        put(C0A + 2 * l, dW, db)
        if l > 0:
            pool_grad = rnd(pool_bwd(rnd(gp)))
    return grad.ravel()


def init_params(seed=0, out_scale=0.1):
    """He-style random initialisation; the output layer is scaled down."""
    rng = np.random.default_rng(seed)
    p = np.zeros((NL, PSZ))
    for l in range(NL):
        w = rng.normal(size=(9, 32, 32)) * np.sqrt(2.0 / (9 * 32))
        if l == COUT:
            w *= out_scale
        p[l, :WSZ] = w.ravel()
    return p.ravel()
