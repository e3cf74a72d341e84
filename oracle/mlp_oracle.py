"""FP64 numpy restatement of the MLP of include/shark_mlp.h (TEST INFRASTRUCTURE ONLY).

No reference counterpart (the reference has no learned components).  `bf16=True` rounds every
product operand to BF16 where the GPU kernels do (weights, layer inputs, output gradients; FP32
values first), so the GPU's forward pass and gradients can be compared to round-off of FP32
accumulation; `bf16=False` is the exact FP64 network, whose gradients are pinned against
central finite differences in tests/test_mlp_oracle.py.
"""
import numpy as np


def bf16(a):
    """Round to BF16 (round to nearest even) from the FP32 value, returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _id(a):
    return np.asarray(a, dtype=np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(layers, x, bf16_ops=True, keep=False):
    """layers = [(W (dout x din), b)], x (n x din).  Returns y (and pre-activations/inputs)."""
    r = bf16 if bf16_ops else _id
    a = _id(x)
    zs, ins = [], []
    for l, (W, b) in enumerate(layers):
        ins.append(a)
        z = r(a) @ r(W).T + _id(b)
        if l < len(layers) - 1:
            zs.append(z)
            a = silu(z)
            if bf16_ops:
                a = a.astype(np.float32).astype(np.float64)
        else:
            a = z
    return (a, zs, ins) if keep else a


def loss_grads(layers, x, t, bf16_ops=True, w=None):
    """Loss 1/(n n_out) sum w (y - t)^2 (w = 1 when None) and [(dW, db)] per layer."""
    r = bf16 if bf16_ops else _id
    y, zs, ins = forward(layers, x, bf16_ops, keep=True)
    n, no = y.shape
    d = y - _id(t)
    wt = np.ones_like(d) if w is None else _id(w)
    loss = float((wt * d * d).sum() / (n * no))
    dz = 2.0 * wt * d / (n * no)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (r(dz).T @ r(ins[l]), dz.sum(axis=0))
        if l > 0:
            dz = (r(dz) @ r(W)) * dsilu(zs[l - 1])
    return loss, grads
