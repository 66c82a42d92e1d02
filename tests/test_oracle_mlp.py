"""Pins the CPU oracle of the sign-risk proposer MLP (oracle/mlp.py).

The FP64 gradient is checked against central finite differences of the FP64
loss; BF16 rounding is checked against torch's own float32 -> bfloat16
conversion; the BF16-emulating gradient is checked to stay close to the FP64
one, and a planted error (a dropped ReLU mask) must fail the finite-difference
check.
"""
import numpy as np
import pytest

from oracle import mlp


def _data(n=96, seed=3):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 12)).astype(np.float32)
    Y = (rng.random(n) < 0.3).astype(np.float32)
    return X, Y


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.normal(size=20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-20, 20, 20000).astype(np.float32),
        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, -0.0, 1e-40], np.float32),
    ]).astype(np.float32)
    ours = mlp.bf16_round(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours, ref)
    # ties go to even
    assert mlp.bf16_round(np.float32(1.0 + 2 ** -8))[()] == 1.0
    assert mlp.bf16_round(np.float32(1.0 + 3 * 2 ** -8))[()] == 1.0 + 2 ** -6


def test_blob_layout():
    assert mlp.N_PARAM == 12 * 64 + 64 + 2 * (64 * 64 + 64) + 64 + 1 == 9217
    assert mlp.N_BLOB == 9241


def test_forward_zero_weights_gives_bias():
    b = np.zeros(mlp.N_BLOB, np.float32)
    b[mlp.OFF_ISTD:] = 1.0
    b[mlp.OFF_B4] = 0.75
    X, _ = _data(10)
    z, _ = mlp.forward(b, X, bf16=False)
    assert np.all(z == 0.75)


def test_forward_matches_loop():
    blob = mlp.init_blob(5)
    blob[mlp.OFF_B1:mlp.OFF_W2] = 0.1
    X, _ = _data(7)
    z, _ = mlp.forward(blob, X, bf16=False)
    P = mlp.unpack(blob)
    for i in range(7):
        h = [max(0.0, sum(float(X[i, k]) * P["W1"][k, j] for k in range(12)) + P["b1"][j]) for j in range(64)]
        for W, bb in (("W2", "b2"), ("W3", "b3")):
            h = [max(0.0, sum(h[k] * P[W][k, j] for k in range(64)) + P[bb][j]) for j in range(64)]
        zz = sum(h[k] * P["w4"][k] for k in range(64)) + P["b4"]
        assert abs(zz - z[i]) <= 1e-9 * (1 + abs(zz))


def _fd_check(grad_fn, blob, X, Y, pw, idx, eps=1e-5):
    _, g = grad_fn(blob, X, Y, pw)
    for i in idx:
        bp = blob.astype(np.float64).copy()
        bm = blob.astype(np.float64).copy()
        bp[i] += eps
        bm[i] -= eps
        lp = _loss64(bp, X, Y, pw)[0]
        lm = _loss64(bm, X, Y, pw)[0]
        fd = (lp - lm) / (2 * eps)
        if abs(fd - g[i]) > 1e-6 * (1 + abs(fd)):
            return False
    return True


def _loss64(blob64, X, Y, pw):
    # FP64 loss with FP64 parameters (no float32 cast of the perturbed blob)
    P = {
        "W1": blob64[mlp.OFF_W1:mlp.OFF_B1].reshape(12, 64), "b1": blob64[mlp.OFF_B1:mlp.OFF_W2],
        "W2": blob64[mlp.OFF_W2:mlp.OFF_B2].reshape(64, 64), "b2": blob64[mlp.OFF_B2:mlp.OFF_W3],
        "W3": blob64[mlp.OFF_W3:mlp.OFF_B3].reshape(64, 64), "b3": blob64[mlp.OFF_B3:mlp.OFF_W4],
        "w4": blob64[mlp.OFF_W4:mlp.OFF_B4], "b4": blob64[mlp.OFF_B4],
    }
    x = X.astype(np.float64)
    h = np.maximum(x @ P["W1"] + P["b1"], 0)
    h = np.maximum(h @ P["W2"] + P["b2"], 0)
    h = np.maximum(h @ P["W3"] + P["b3"], 0)
    z = h @ P["w4"] + P["b4"]
    y = Y.astype(np.float64)
    cw = np.where(y > 0.5, pw, 1.0)
    return float(np.sum(cw * (np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z))) - y * z))), None


def _grad64(blob, X, Y, pw):
    return mlp.loss_and_grad(blob, X, Y, pw, bf16=False)


def _sample_idx(seed=1):
    rng = np.random.default_rng(seed)
    parts = [rng.integers(a, b, 6) for a, b in ((mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2),
                                                (mlp.OFF_W2, mlp.OFF_B2), (mlp.OFF_B2, mlp.OFF_W3),
                                                (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.OFF_W4),
                                                (mlp.OFF_W4, mlp.OFF_B4))]
    return np.concatenate(parts + [np.array([mlp.OFF_B4])])


def test_fp64_gradient_matches_finite_differences():
    blob = mlp.init_blob(2)
    blob[mlp.OFF_B1:mlp.OFF_W2] = 0.05
    blob[mlp.OFF_B2:mlp.OFF_W3] = 0.05
    blob[mlp.OFF_B3:mlp.OFF_W4] = 0.05
    X, Y = _data()
    # the blob carries mean 0 / istd 1, so the FP64 loss above sees the same input
    l64, g = _grad64(blob, X, Y, 3.0)
    assert abs(l64 - _loss64(blob.astype(np.float64), X, Y, 3.0)[0]) < 1e-9 * abs(l64)
    assert _fd_check(_grad64, blob, X, Y, 3.0, _sample_idx())


def test_planted_mask_error_fails_finite_differences():
    blob = mlp.init_blob(2)
    X, Y = _data()

    def broken(b, X, Y, pw):
        # the dZ2 ReLU mask dropped: must be caught
        loss, g = mlp.loss_and_grad(b, X, Y, pw, bf16=False)
        P = mlp.unpack(b)
        z, (P, x, h1, h2, h3) = mlp.forward(b, X, bf16=False)
        y = Y.astype(np.float64)
        dz = np.where(y > 0.5, pw, 1.0) * (1 / (1 + np.exp(-z)) - y)
        d3 = np.where(h3 > 0, dz[:, None] * P["w4"][None, :], 0.0)
        d2 = d3 @ P["W3"].T  # no mask
        g = g.copy()
        g[mlp.OFF_W2:mlp.OFF_B2] = (h1.T @ d2).reshape(-1)
        return loss, g

    assert not _fd_check(broken, blob, X, Y, 3.0, _sample_idx())


def test_bf16_gradient_close_to_fp64():
    blob = mlp.init_blob(4)
    X, Y = _data(512)
    _, g64 = mlp.loss_and_grad(blob, X, Y, 2.0, bf16=False)
    _, g16 = mlp.loss_and_grad(blob, X, Y, 2.0, bf16=True)
    rel = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    assert rel < 0.05
    assert np.dot(g16, g64) / (np.linalg.norm(g16) * np.linalg.norm(g64)) > 0.998


def test_adam_matches_textbook():
    rng = np.random.default_rng(0)
    p = rng.normal(size=64).astype(np.float32)
    g = rng.normal(size=64).astype(np.float32)
    m = np.zeros(64, np.float32)
    v = np.zeros(64, np.float32)
    p0 = p.copy()
    mlp.adam(p, g, m, v, 64, lr=0.01, step=1)
    # first Adam step moves every parameter by lr * sign(g) (up to eps)
    assert np.allclose(p0 - p, 0.01 * np.sign(g), atol=1e-6)
