"""The learned cost model's HIP kernels against a PyTorch fp32 reference
(on the CPU): forward, gradients of the mean squared loss, Adam steps; then
online learning and the trust rule on a synthetic cost surface.

Tolerance: fp32 with different summation orders, |diff| <= 1e-4 * (1 + |ref|)
for outputs and gradients, 1e-5 absolute after Adam steps of lr 1e-3."""
from __future__ import annotations

import numpy as np
import pytest

from mordred_amd import costmodel as CM

pytestmark = pytest.mark.gpu


def torch_mlp(p):
    import torch
    W1, b1, W2, b2, W3, b3 = [torch.tensor(a.copy(), dtype=torch.float32, requires_grad=True)
                              for a in CM.split_params(p)]

    def f(x):
        h1 = torch.relu(x @ W1 + b1)
        h2 = torch.relu(h1 @ W2 + b2)
        return h2 @ W3 + b3
    return f, [W1, b1, W2, b2, W3, b3]


def data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, CM.FEATURES)).astype(np.float32)
    t = (0.5 * x[:, 0] - 0.25 * x[:, 1] ** 2 + 0.1 * x[:, 2] * x[:, 3] + 1.0).astype(np.float32)
    return x, t


def close(a, b, tol=1e-4):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    assert np.all(np.abs(a - b) <= tol * (1 + np.abs(b))), float(np.max(np.abs(a - b) / (1 + np.abs(b))))


def test_forward_matches_torch(cuda):
    import torch
    m = CM.CostModel(0, seed=3)
    x, _ = data(300, 1)
    f, _ = torch_mlp(m.params())
    with torch.no_grad():
        ref = f(torch.tensor(x)).numpy()
    close(m.predict(x), ref)
    m.close()


def test_gradients_match_torch(cuda):
    import torch
    m = CM.CostModel(0, seed=4)
    x, t = data(200, 2)
    g, loss = m.gradients(x, t)
    f, ps = torch_mlp(m.params())
    l = ((f(torch.tensor(x)) - torch.tensor(t)) ** 2).mean()
    l.backward()
    ref = np.concatenate([p.grad.numpy().reshape(-1) for p in ps])
    close(loss, l.item())
    close(g, ref)
    m.close()


def test_adam_matches_torch(cuda):
    import torch
    m = CM.CostModel(0, seed=5)
    f, ps = torch_mlp(m.params())
    opt = torch.optim.Adam(ps, lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for s in range(5):
        x, t = data(128, 10 + s)
        m.step(x, t, 1e-3)
        opt.zero_grad()
        ((f(torch.tensor(x)) - torch.tensor(t)) ** 2).mean().backward()
        opt.step()
    ref = np.concatenate([p.detach().numpy().reshape(-1) for p in ps])
    assert np.max(np.abs(m.params() - ref)) < 1e-5
    m.close()


def test_online_learning_and_trust(cuda):
    """A cost surface the analytic guess gets wrong by a constant factor: the
    model learns it from observations and becomes trusted."""
    m = CM.CostModel(0, seed=6)
    rng = np.random.default_rng(7)
    assert not m.stats()["trusted"]
    for i in range(400):
        x = rng.uniform(0, 1, CM.FEATURES).astype(np.float32)
        ms = float(np.exp(1.0 + 2.0 * x[0] - 1.0 * x[1] + 0.5 * x[2] * x[3]))
        m.observe(x, ms, analytic_ms=ms * 2.5)
        if i % 20 == 19:
            m.train(40, batch=128, lr=3e-3)
    st = m.stats()
    assert st["samples"] == 400
    assert st["ewma_abs_log_error"] < 0.25 < st["ewma_abs_log_error_analytic"]
    assert st["trusted"]
    m.close()


def test_train_equals_successive_steps(cuda):
    """train() stages every minibatch at once and runs its steps back to back
    with one synchronization; with a replay buffer no larger than the batch
    each step takes every sample in order, so k steps of train() must equal
    k calls of step() on those samples."""
    rng = np.random.default_rng(11)
    xs = rng.uniform(0, 1, (50, CM.FEATURES)).astype(np.float32)
    ms = np.exp(rng.uniform(-1, 2, 50)).astype(np.float64)
    a, b = CM.CostModel(0, seed=12), CM.CostModel(0, seed=12)
    for x, t in zip(xs, ms):
        a.observe(x, float(t), analytic_ms=float(t))
    loss = a.train(3, batch=128, lr=2e-3)
    for _ in range(3):
        b.step(xs, np.log(ms).astype(np.float32), 2e-3)
    np.testing.assert_allclose(a.params(), b.params(), rtol=0, atol=1e-6)
    assert np.isfinite(loss)
    a.close()
    b.close()
