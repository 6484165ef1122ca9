"""CPU fp32 reference of the learned-order MLP (forward, masked BCE, backward),
rounding to bfloat16 exactly where the HIP kernels do: weights and inputs are
bf16, each hidden activation is stored as bf16 after bias and ReLU, dL/dz and
the back-propagated gradients are stored as bf16; every product accumulates in
fp32.  The GPU results are compared against this within tolerances stated in
the tests (summation order differs)."""
import numpy as np


def bf16(x):
    """Round fp32 to the nearest bfloat16 (ties to even), returned as fp32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def forward(Ws, bs, x):
    """Ws: fp32 [out][in] (rounded to bf16 here), x: 0/1 [B][in]."""
    h = bf16(x)
    acts = [h]
    L = len(Ws)
    for l in range(L):
        z = h @ bf16(Ws[l]).T + bs[l]
        if l < L - 1:
            h = bf16(np.maximum(z, 0.0))
            acts.append(h)
        else:
            return z.astype(np.float32), acts


def bce(z, y, mask):
    count = max(float(mask.sum()), 1.0)
    sp = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))
    loss = float((mask * (sp - y * z)).sum())
    sg = 1.0 / (1.0 + np.exp(-z))
    dz = bf16(mask * (sg - y) / count)
    return loss, count, dz


def backward(Ws, acts, dz):
    """Returns (dW list, db list) for dz = dL/dlogits (bf16 values)."""
    L = len(Ws)
    dW, db = [None] * L, [None] * L
    g = dz
    for l in reversed(range(L)):
        db[l] = g.sum(0)
        dW[l] = g.T @ acts[l]
        if l > 0:
            g = bf16((g @ bf16(Ws[l])) * (acts[l] > 0))
    return dW, db
