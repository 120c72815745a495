"""CPU FP64 reference of the perceptron's forward pass, loss and gradient
(numpy), the checker of the BF16-MFMA kernels.

``quantize=True`` rounds at the points where the GPU path stores bfloat16
(inputs, weights, hidden activations, the back-propagated errors that enter
the products), so that the remaining difference is FP32-vs-FP64 summation
only; ``quantize=False`` is the exact FP64 network on the same parameters.
"""
import math

import numpy as np

SQRT1_2 = 1.0 / math.sqrt(2.0)


def bf16(a):
    """Round to bfloat16 (nearest even) and back, as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


_erf = np.vectorize(math.erf)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x * SQRT1_2))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x * SQRT1_2)) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def forward(params, x, quantize=True):
    """params: [(W, b)] float arrays (padded); x [B, in_pad].  Returns
    (out [B, out_pad], pre-activations, layer inputs)."""
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    a = q(x)
    pres, ins = [], []
    for l, (W, b) in enumerate(params):
        ins.append(a)
        z = a @ q(W).T + np.asarray(b, dtype=np.float64)
        if l < len(params) - 1:
            if quantize:
                z = z.astype(np.float32).astype(np.float64)   # stored in fp32
            pres.append(z)
            a = q(gelu(z))
        else:
            out = z
    return out, pres, ins


def loss_and_grad(params, x, tgt, nout, quantize=True):
    """Mean-squared error over the first nout outputs and its gradient
    [(dW, db)] (padded shapes)."""
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    out, pres, ins = forward(params, x, quantize)
    B = x.shape[0]
    d = np.zeros_like(out)
    d[:, :nout] = out[:, :nout] - tgt
    loss = float((d * d).sum() / (B * nout))
    dy = 2.0 * d / (B * nout)
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        W, _ = params[l]
        dyq = q(dy)
        grads[l] = (dyq.T @ ins[l], dy.sum(axis=0))
        if l == 0:
            break
        dy = (dyq @ q(W)) * gelu_grad(pres[l - 1])
    return loss, grads
