"""Pins the numpy MLP reference used to check the cache kernels: with its BF16 roundings
switched off, its backpropagated gradient equals central differences of the loss (whose
denominator is held at the unperturbed prediction), and its forward pass equals a plain
matrix product written out per layer."""
import numpy as np

import mlp_ref


def test_reference_gradient_matches_central_differences(monkeypatch):
    monkeypatch.setattr(mlp_ref, "bf16", lambda x: np.asarray(x, np.float64))
    monkeypatch.setattr(mlp_ref, "DTYPE", np.float64)
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, 14787)
    enc = rng.normal(size=(40, 32))
    t = rng.uniform(0, 1, size=(40, 3))
    t[:6] *= 1000.0  # beyond the Huber threshold: the linear branch is exercised too
    loss, g = mlp_ref.loss_and_grad(w, enc, t)
    y0, _, _ = mlp_ref.forward(w, enc)
    sd = np.sqrt(y0 * y0 + mlp_ref.EPS)
    H = mlp_ref.HUBER

    def f(wv):
        y, _, _ = mlp_ref.forward(wv, enc)
        r = np.abs((y - t) / sd)
        return np.where(r <= H, r * r, H * (2 * r - H)).sum() / (3 * len(enc))

    assert abs(f(w) - loss) < 1e-12
    r = np.abs((y0 - t) / sd)
    assert (r > H).any() and (r <= H).any()
    idx = np.concatenate([rng.choice(14787, 60, replace=False), [14784, 14785, 14786, 2048, 2049]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(w)
        e[i] = h
        fd = (f(w + e) - f(w - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_reference_l2_gradient_matches_central_differences(monkeypatch):
    """the L2 loss (mean of (y - t)^2): its gradient by central differences, and the property
    that makes it mean-preserving: the output biases' gradient is twice the mean residual"""
    monkeypatch.setattr(mlp_ref, "bf16", lambda x: np.asarray(x, np.float64))
    monkeypatch.setattr(mlp_ref, "DTYPE", np.float64)
    rng = np.random.default_rng(2)
    w = rng.normal(0, 0.3, 14787)
    enc = rng.normal(size=(40, 32))
    t = rng.uniform(0, 3, size=(40, 3))
    loss, g = mlp_ref.loss_and_grad(w, enc, t, l2=True)

    def f(wv):
        y, _, _ = mlp_ref.forward(wv, enc)
        return ((y - t) ** 2).sum() / (3 * len(enc))

    assert abs(f(w) - loss) < 1e-12
    idx = np.concatenate([rng.choice(14787, 60, replace=False), [14784, 14785, 14786]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(w)
        e[i] = h
        fd = (f(w + e) - f(w - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    y, _, _ = mlp_ref.forward(w, enc)
    np.testing.assert_allclose(g[-3:], 2 * (y - t).mean(0) / 3, rtol=1e-10)


def test_reference_forward_is_the_layer_stack():
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.3, 14787).astype(np.float32)
    enc = rng.normal(size=(8, 32)).astype(np.float32)
    y, hs, _ = mlp_ref.forward(w, enc)
    h = mlp_ref.bf16(enc)
    o = 0
    for l, (i, n) in enumerate(mlp_ref.LAYERS):
        W = mlp_ref.bf16(w[o:o + n * i].reshape(n, i))
        b = w[o + n * i:o + n * i + n]
        o += n * i + n
        z = np.einsum("oi,si->so", W, h) + b
        h = mlp_ref.bf16(np.maximum(z, 0)) if l < 4 else z
    np.testing.assert_allclose(y, h, rtol=1e-5, atol=1e-6)


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5], np.float32)
    np.testing.assert_array_equal(mlp_ref.bf16(x), np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0, -2.5], np.float32))
