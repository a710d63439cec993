"""Test infrastructure only: FP64/FP32 numpy restatement of the learned
warm-start network (no reference counterpart: the reference has no learned
component), used to check the HIP forward/backward kernels and the feature
kernel.

Network: 32 -> 128 -> 128 -> 128 -> 1, GELU (tanh form,
0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))) after each hidden layer,
weights rounded to BF16, activations rounded to BF16 at each matmul input,
FP32 accumulation.  Loss: mean squared error.
"""
from __future__ import annotations

import numpy as np

SIZES = (32, 128, 128, 128, 1)
C_GELU = np.sqrt(2.0 / np.pi)


def bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float32."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    u = ((u + r) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * x * x)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    W, b = [], []
    for fi, fo in zip(SIZES[:-1], SIZES[1:]):
        W.append((rng.standard_normal((fi, fo)) * np.sqrt(2.0 / fi)).astype(np.float32))
        b.append(np.zeros(fo, np.float32))
    return W, b


def forward(X, W, b, round_bf16=True):
    """Returns output y [N] and the pre-activations / inputs of every layer."""
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float32))
    acts = [q(X)]
    pre = []
    h = acts[0]
    for li in range(len(W)):
        z = h.astype(np.float64) @ q(W[li]).astype(np.float64) + b[li]
        z = z.astype(np.float32)
        pre.append(z)
        if li < len(W) - 1:
            h = q(gelu(z))
            acts.append(h)
    return pre[-1][:, 0], pre, acts


def backward(X, y_target, W, b, round_bf16=True):
    """Gradients of 0.5 * mean((y - t)^2) w.r.t. W, b.  The gradient flowing
    into a hidden layer through W3 / W2 is rounded to bf16 first (it is the
    matrix-core operand), the one through the 128 -> 1 head is not; weight
    gradients use the bf16 layer inputs and FP32/FP64 gradients."""
    y, pre, acts = forward(X, W, b, round_bf16)
    N = len(y)
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float32))
    g = ((y - y_target) / N).astype(np.float64)[:, None]
    gW = [None] * len(W)
    gb = [None] * len(W)
    for li in reversed(range(len(W))):
        gW[li] = acts[li].astype(np.float64).T @ g
        gb[li] = g.sum(0)
        if li > 0:
            gin = g if li == len(W) - 1 else q(g.astype(np.float32)).astype(np.float64)
            gh = gin @ q(W[li]).astype(np.float64).T
            g = (gh * gelu_grad(pre[li - 1].astype(np.float64))).astype(np.float32).astype(np.float64)
    loss = 0.5 * np.mean((y - y_target) ** 2)
    return loss, gW, gb


def pack(W, b):
    """Packed FP32 layout of pgert_mlp_set_params: W1 b1 W2 b2 W3 b3 w4 b4."""
    parts = []
    for Wi, bi in zip(W, b):
        parts += [np.asarray(Wi, np.float32).ravel(), np.asarray(bi, np.float32).ravel()]
    return np.concatenate(parts)


def unpack(p):
    W, b = [], []
    o = 0
    for fi, fo in zip(SIZES[:-1], SIZES[1:]):
        W.append(p[o:o + fi * fo].reshape(fi, fo)); o += fi * fo
        b.append(p[o:o + fo]); o += fo
    return W, b
