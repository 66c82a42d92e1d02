"""CPU oracle of the sign-risk proposer (MLP 12 -> 64 -> 64 -> 64 -> 1, ReLU).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
forward, gradient and Adam kernels.  The proposer has no counterpart in the
reference (the reference's patcher probes blind stratified samples,
cuda_rasterizer.cu:1045-1118); this module restates the network the HIP
kernels implement, in FP64 numpy.

Parameter blob (float32, 9241 values): W1[12][64] b1[64] W2[64][64] b2[64]
W3[64][64] b3[64] w4[64] b4 | feature mean[12] | 1/std[12].
Layer l computes h_l = relu(h_{l-1} W_l + b_l); logit = h3 . w4 + b4; the
input is (x - mean) * istd.

``bf16=True`` rounds, to BF16 (nearest even), exactly the operands the
matrix-core kernels round: the normalised input, W1..W3, h1 and h2 as layer
inputs, w4 and h3 in the last dot product, and in the backward pass the
operands of every matrix product (dZ3, dZ2, dZ1, H2, H1, X, W2^T, W3^T).
Sums run in FP64 here and in FP32 on the GPU.  ``bf16=False`` is the plain
FP64 network; its gradient is pinned by central finite differences.
"""
from __future__ import annotations

import numpy as np

N_IN = 12
N_H = 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_H
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H * N_H
OFF_W4 = OFF_B3 + N_H
OFF_B4 = OFF_W4 + N_H
N_PARAM = OFF_B4 + 1
OFF_MEAN = N_PARAM
OFF_ISTD = OFF_MEAN + N_IN
N_BLOB = OFF_ISTD + N_IN


def bf16_round(x) -> np.ndarray:
    """FP32 -> BF16 (round to nearest, ties to even) -> FP32 (finite inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(blob):
    b = np.asarray(blob, dtype=np.float32)
    assert b.size >= N_PARAM
    return {
        "W1": b[OFF_W1:OFF_B1].reshape(N_IN, N_H).astype(np.float64),
        "b1": b[OFF_B1:OFF_W2].astype(np.float64),
        "W2": b[OFF_W2:OFF_B2].reshape(N_H, N_H).astype(np.float64),
        "b2": b[OFF_B2:OFF_W3].astype(np.float64),
        "W3": b[OFF_W3:OFF_B3].reshape(N_H, N_H).astype(np.float64),
        "b3": b[OFF_B3:OFF_W4].astype(np.float64),
        "w4": b[OFF_W4:OFF_B4].astype(np.float64),
        "b4": float(b[OFF_B4]),
        "mean": b[OFF_MEAN:OFF_ISTD].astype(np.float32) if b.size >= N_BLOB else np.zeros(N_IN, np.float32),
        "istd": b[OFF_ISTD:N_BLOB].astype(np.float32) if b.size >= N_BLOB else np.ones(N_IN, np.float32),
    }


def init_blob(seed: int = 0, mean=None, istd=None) -> np.ndarray:
    """He-uniform initial weights, zero biases (float32 blob)."""
    rng = np.random.default_rng(seed)
    b = np.zeros(N_BLOB, dtype=np.float32)

    def he(fan_in, shape):
        lim = np.sqrt(6.0 / fan_in)
        return rng.uniform(-lim, lim, size=shape).astype(np.float32)

    b[OFF_W1:OFF_B1] = he(N_IN, N_IN * N_H)
    b[OFF_W2:OFF_B2] = he(N_H, N_H * N_H)
    b[OFF_W3:OFF_B3] = he(N_H, N_H * N_H)
    b[OFF_W4:OFF_B4] = he(N_H, N_H)
    b[OFF_MEAN:OFF_ISTD] = 0.0 if mean is None else np.asarray(mean, np.float32)
    b[OFF_ISTD:N_BLOB] = 1.0 if istd is None else np.asarray(istd, np.float32)
    return b


def _normalise(P, X, bf16):
    # the kernels form (x - mean) * istd in FP32
    x = (np.asarray(X, np.float32).reshape(-1, N_IN) - P["mean"]) * P["istd"]
    return bf16_round(x) if bf16 else x.astype(np.float64)


def forward(blob, X, bf16: bool = True):
    """Returns (logits [n] float64, cache for backward)."""
    P = unpack(blob)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    x = _normalise(P, X, bf16)
    h1 = np.maximum(x @ r(P["W1"]) + P["b1"], 0.0)
    h2 = np.maximum(r(h1) @ r(P["W2"]) + P["b2"], 0.0)
    h3 = np.maximum(r(h2) @ r(P["W3"]) + P["b3"], 0.0)
    z = r(h3) @ r(P["w4"]) + P["b4"]
    return z, (P, x, h1, h2, h3)


def loss_and_grad(blob, X, Y, pos_weight: float = 1.0, bf16: bool = True):
    """sum_i c_i [softplus(z_i) - y_i z_i] (c_i = pos_weight if y_i = 1)
    and its gradient w.r.t. the 9217 parameters (float64)."""
    z, (P, x, h1, h2, h3) = forward(blob, X, bf16)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    y = np.asarray(Y, np.float64).reshape(-1)
    cw = np.where(y > 0.5, pos_weight, 1.0)
    loss = float(np.sum(cw * (np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z))) - y * z)))
    dz = cw * (1.0 / (1.0 + np.exp(-z)) - y)
    g = np.zeros(N_PARAM, dtype=np.float64)
    g[OFF_B4] = dz.sum()
    g[OFF_W4:OFF_B4] = r(h3).T @ dz
    d3 = np.where(h3 > 0.0, dz[:, None] * r(P["w4"])[None, :], 0.0)
    g[OFF_B3:OFF_W4] = d3.sum(0)
    g[OFF_W3:OFF_B3] = (r(h2).T @ r(d3)).reshape(-1)
    d2 = np.where(h2 > 0.0, r(d3) @ r(P["W3"]).T, 0.0)
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W2:OFF_B2] = (r(h1).T @ r(d2)).reshape(-1)
    d1 = np.where(h1 > 0.0, r(d2) @ r(P["W2"]).T, 0.0)
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W1:OFF_B1] = (x.T @ r(d1)).reshape(-1)
    return loss, g


def adam(blob, grad, m, v, n_params, lr, step, beta1=0.9, beta2=0.999, eps=1e-8, grad_scale=1.0,
         weight_decay=0.0):
    """One Adam step on the first n_params values (float32 arrays updated in
    place, in FP32 arithmetic like the kernel: g = grad*scale + wd*p)."""
    f = np.float32
    p = blob[:n_params]
    gg = grad[:n_params].astype(f) * f(grad_scale) + f(weight_decay) * p
    bc1 = f(1.0) - f(beta1) ** f(step)
    bc2 = f(1.0) - f(beta2) ** f(step)
    m[:n_params] = f(beta1) * m[:n_params] + (f(1.0) - f(beta1)) * gg
    v[:n_params] = f(beta2) * v[:n_params] + (f(1.0) - f(beta2)) * gg * gg
    p -= f(lr) * (m[:n_params] / bc1) / (np.sqrt(v[:n_params] / bc2) + f(eps))
