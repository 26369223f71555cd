"""TEST INFRASTRUCTURE — CPU restatement of the warm-start MLP kernels.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module; the product never does.

The network (8 -> 256 GELU -> 256 GELU -> 3, BF16 weights and activations,
FP32 accumulation on the GPU) has no counterpart in the reference; this file
restates the GPU kernels' arithmetic in float64 with the same rounding points:

  * every matrix operand the GPU feeds to its matrix cores is rounded to BF16
    (round to nearest, ties to even): normalized inputs, h1, h2, the weights,
    and in the backward pass dy, dz2 and dz1 for the weight gradients and
    bf16(dz2) for the input gradient of layer 2;
  * bias gradients and the layer-3 input gradient use the FP32 values;
  * GELU is the exact erf form, gelu(x) = x * Phi(x).

What pins it (tests/test_mlp.py): finite differences of the loss in float64
(no rounding) against backward() with rounding switched off; closed-form GELU
values; BF16 rounding known answers (ties to even, the carry into the
exponent); and the Adam update against its published definition (Kingma and
Ba, 2015, Algorithm 1).
"""

import math

import numpy as np

IN, HIDDEN, OUT = 8, 256, 3


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even); returns float32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (r << 16).astype(np.uint32).view(np.float32)
    return out.reshape(a.shape)


def bf16_bits(x):
    """The uint16 bit patterns of bf16(x)."""
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


_erf = np.vectorize(math.erf, otypes=[np.float64])


def gelu(x):
    x = np.asarray(x, np.float64)
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    x = np.asarray(x, np.float64)
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def _q(x, quantize):
    return bf16(x).astype(np.float64) if quantize else np.asarray(x, np.float64)


def forward(p, x, quantize=True):
    """p: dict of float32 arrays w1 [256][8], b1, w2 [256][256], b2, w3 [3][256],
    b3, in_shift, in_scale, out_scale (weights given as their BF16 values when
    quantize is set). Returns the activations the GPU records and y_hat."""
    x = np.asarray(x, np.float32)
    xn = ((x - p["in_shift"]) * p["in_scale"]).astype(np.float32)
    xq = _q(xn, quantize)
    w1, w2, w3 = (_q(p[k], quantize) for k in ("w1", "w2", "w3"))
    z1 = xq @ w1.T + p["b1"].astype(np.float64)
    h1 = _q(gelu(z1).astype(np.float32), quantize) if quantize else gelu(z1)
    z2 = h1 @ w2.T + p["b2"].astype(np.float64)
    h2 = _q(gelu(z2).astype(np.float32), quantize) if quantize else gelu(z2)
    y = h2 @ w3.T + p["b3"].astype(np.float64)
    return {"xq": xq, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "y_hat": y}


def mse(y_hat, target, out_scale, huber_delta=0.0):
    """Loss and d loss / d y_hat: mean squared error, or the Huber loss for huber_delta > 0."""
    t = np.asarray(target, np.float64) / np.asarray(out_scale, np.float64)
    d = np.asarray(y_hat, np.float64) - t
    if huber_delta > 0.0:
        ad = np.abs(d)
        inner = ad <= huber_delta
        loss = np.where(inner, 0.5 * d * d, huber_delta * (ad - 0.5 * huber_delta))
        return float(np.mean(loss)), np.clip(d, -huber_delta, huber_delta) / d.size
    return float(np.mean(d * d)), 2.0 * d / d.size


def backward(p, tr, dy, quantize=True):
    dy = np.asarray(dy, np.float64)
    w2, w3 = _q(p["w2"], quantize), _q(p["w3"], quantize)
    g = {}
    g["w3"] = _q(dy, quantize).T @ tr["h2"]
    g["b3"] = dy.sum(0)
    dz2 = (dy @ w3) * gelu_grad(tr["z2"])
    g["w2"] = _q(dz2, quantize).T @ tr["h1"]
    g["b2"] = dz2.sum(0)
    dz1 = (_q(dz2, quantize) @ w2) * gelu_grad(tr["z1"])
    g["w1"] = _q(dz1, quantize).T @ tr["xq"]
    g["b1"] = dz1.sum(0)
    return g, dz2, dz1


def adam(w, g, m1, m2, lr, beta1, beta2, eps, step):
    m1 = beta1 * m1 + (1.0 - beta1) * g
    m2 = beta2 * m2 + (1.0 - beta2) * g * g
    mhat = m1 / (1.0 - beta1 ** step)
    vhat = m2 / (1.0 - beta2 ** step)
    return w - lr * mhat / (np.sqrt(vhat) + eps), m1, m2
