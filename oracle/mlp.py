"""TEST INFRASTRUCTURE: FP64 NumPy restatement of the traveltime surrogate's
network (5-64-64-64-64-2, SiLU), its mean-squared-error gradient and one Adam
step, in the parameter layout of include/pstm_hip.h (per layer W (out x in)
row-major, then b).  Used to check the HIP training kernels.
"""

from __future__ import annotations

import numpy as np

SIZES = [(5, 64), (64, 64), (64, 64), (64, 64), (64, 2)]


def unpack(p):
    layers, off = [], 0
    for fin, fout in SIZES:
        W = p[off:off + fin * fout].reshape(fout, fin)
        off += fin * fout
        b = p[off:off + fout]
        off += fout
        layers.append((W, b))
    assert off == p.size
    return layers


def silu(z):
    return z / (1.0 + np.exp(-z))


def forward(p, x):
    a = np.asarray(x, np.float64)
    ls = unpack(np.asarray(p, np.float64))
    for i, (W, b) in enumerate(ls):
        z = a @ W.T + b
        a = silu(z) if i < len(ls) - 1 else z
    return a


def loss_grad(p, x, t):
    """(mean squared error over all outputs, gradient in the packed layout)."""
    p = np.asarray(p, np.float64)
    x = np.asarray(x, np.float64)
    ls = unpack(p)
    acts, zs = [x], []
    a = x
    for i, (W, b) in enumerate(ls):
        z = a @ W.T + b
        zs.append(z)
        a = silu(z) if i < len(ls) - 1 else z
        acts.append(a)
    y = acts[-1]
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    d = 2.0 * (y - t) / n
    grads = []
    for i in range(len(ls) - 1, -1, -1):
        W, b = ls[i]
        if i < len(ls) - 1:
            s = 1.0 / (1.0 + np.exp(-zs[i]))
            d = d * s * (1.0 + zs[i] * (1.0 - s))
        grads.append((d.T @ acts[i], d.sum(0)))
        d = d @ W
    g = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in reversed(grads)])
    return loss, g


def adam(p, g, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16_round(a):
    """Nearest bfloat16 (ties to even) of float32 values, returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_pair(a):
    """a (rounded to float32) as the sum of two bfloat16 values hi + lo."""
    a32 = np.asarray(a, dtype=np.float32)
    hi = bf16_round(a32)
    lo = bf16_round((a32.astype(np.float64) - hi).astype(np.float32))
    return hi + lo


# the fused inference carries every pre-activation and activation times
# c = -log2(e) (so that its SiLU needs no multiply before the exponential):
# inputs c x, biases c b (f32), c SiLU(z) = z' / (1 + 2^z') with z' = c z,
# outputs divided by c; the trainer's bf16 mode evaluates unscaled (scale 1)
SILU_C = float(np.float32(-1.44269504088896341))
SILU_INV_C = float(np.float32(-0.693147180559945309))


def _silu_scaled(zc, scale):
    return silu(zc) if scale == 1.0 else zc / (1.0 + np.exp2(zc))


def _bias(b, scale):
    return b if scale == 1.0 else (np.float32(scale) * np.asarray(b, np.float32)).astype(np.float64)


def forward_bf16(p, x, scale=1.0):
    """Forward pass with the numerics of the fused inference: weights rounded to
    bfloat16, the input and each hidden activation carried as a bfloat16 pair,
    products and sums exact (f64 here; f32 accumulation on the device),
    biases in f32.  scale: 1 for the trainer's bf16 mode, SILU_C for the
    fused inference (mlp_eval16's scaled evaluation)."""
    ls = unpack(np.asarray(p, np.float32).astype(np.float64))
    a = bf16_pair(scale * np.asarray(x, np.float32) if scale != 1.0 else x)
    for i, (W, b) in enumerate(ls):
        z = a @ bf16_round(W).T + _bias(b, scale)
        a = bf16_pair(_silu_scaled(z, scale)) if i < len(ls) - 1 else z
    return a if scale == 1.0 else a * SILU_INV_C


def forward_bf16_split(p, x, scale=SILU_C):
    """Forward pass with the split-weight numerics of the fused inference
    (mlp_weight_terms = 2): each weight as a high bfloat16 term plus the
    bfloat16 of its remainder, the input and each hidden activation carried
    as a bfloat16 pair, and the product of the two remainders left out (the
    MFMA sequence w_hi x_hi + w_hi x_lo + w_lo x_hi); biases in f32; values
    scaled as in forward_bf16."""
    ls = unpack(np.asarray(p, np.float32).astype(np.float64))

    def split(a):
        a32 = np.asarray(a, dtype=np.float32)
        hi = bf16_round(a32)
        return hi, bf16_round((a32.astype(np.float64) - hi).astype(np.float32))

    xh, xl = split(scale * np.asarray(x, np.float32) if scale != 1.0 else x)
    for i, (W, b) in enumerate(ls):
        wh, wl = split(W)
        z = (xh + xl) @ wh.T + xh @ wl.T + _bias(b, scale)
        if i < len(ls) - 1:
            xh, xl = split(_silu_scaled(z, scale))
        else:
            return z if scale == 1.0 else z * SILU_INV_C
