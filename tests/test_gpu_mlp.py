"""GPU kernels of the learned proposal's network (BF16 MFMA forward, backward, Adam) against the
FP64 oracle (oracle/oracle_mlp.cpp), and a training run that has to learn a known target."""
import numpy as np
import pytest

from gpuupbp_amd import renderer as gpu
from oracle import oracle_py

pytestmark = pytest.mark.gpu


def _samples(rng, n, train=True):
    x = rng.uniform(0, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if not train:
        return np.hstack([x, d]).astype(np.float32)
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    W = rng.uniform(0.1, 2.0, (n, 1))
    return np.hstack([x, d, w, W]).astype(np.float32)


def test_forward_matches_fp64_oracle():
    m = gpu.Mlp(seed=5)
    p = m.params()
    s = _samples(np.random.default_rng(1), 3000, train=False)
    g = m.forward(s).astype(np.float64)
    o = oracle_py.mlp_forward(p.astype(np.float64), s)
    # BF16 weights and activations, FP32 accumulation: about 1% per layer
    err = np.linalg.norm(g - o, axis=1) / (np.linalg.norm(o, axis=1) + 1e-3)
    assert np.median(err) < 0.02 and np.percentile(err, 99) < 0.06, (np.median(err), np.percentile(err, 99))


def test_gradient_matches_fp64_oracle():
    m = gpu.Mlp(seed=6)
    p = m.params()
    s = _samples(np.random.default_rng(2), 4096)
    lg, gg = m.loss_grad(s)
    lo, go = oracle_py.mlp_loss_grad(p.astype(np.float64), s)
    assert abs(lg - lo) / abs(lo) < 0.02, (lg, lo)
    cos = np.dot(gg, go) / (np.linalg.norm(gg) * np.linalg.norm(go))
    assert cos > 0.99, cos
    assert np.linalg.norm(gg - go) / np.linalg.norm(go) < 0.12
    # per layer too, so that no layer's gradient can hide behind another's
    cuts = [0, 3264, 3328, 7424, 7488, 11584, 11648, 14208, 14248]
    for a, b in zip(cuts[:-1], cuts[1:]):
        c = np.dot(gg[a:b], go[a:b]) / (np.linalg.norm(gg[a:b]) * np.linalg.norm(go[a:b]) + 1e-30)
        assert c > 0.97, (a, b, c)


def test_adam_step_matches_fp64_oracle():
    # the GPU's Adam (FP32 master weights, FP32 moments) against the FP64 restatement fed with
    # the GPU's own gradients: two steps, so that the moments and both bias corrections count
    m = gpu.Mlp(seed=7)
    rng = np.random.default_rng(4)
    p = m.params().astype(np.float64)
    mo, vo = np.zeros_like(p), np.zeros_like(p)
    for step, lr in ((1, 1e-3), (2, 5e-4)):
        s = _samples(rng, 4096)
        _, g = m.loss_grad(s)
        m.train_step(s, lr=lr)
        got = m.params().astype(np.float64)
        want, mo, vo = oracle_py.adam_step(p, mo, vo, g.astype(np.float64), step, lr)
        # the update is of size ~lr; FP32 parameters carry ~6e-8 relative. The weight gradients
        # are summed with float atomics in no fixed order, so a gradient that nearly cancels can
        # differ between the two evaluations: a few parameters may step differently
        err = np.abs(got - want)
        bad = err > 2e-3 * lr + 2e-7 * np.abs(want)
        assert bad.mean() < 1e-3 and err.max() <= 2.01 * lr, (step, bad.mean(), err.max())
        moved = np.abs(got - p)
        assert np.median(moved) > 0.3 * lr  # the step did happen
        p = got


def _target_batch(rng, n):
    # scattered directions drawn from one vMF lobe (kappa 25) whose mean depends on the position
    x = rng.uniform(0, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    mu = np.stack([np.cos(2 * np.pi * x[:, 0]), np.sin(2 * np.pi * x[:, 0]), 0.5 * np.ones(n)], 1)
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    kappa = 25.0
    u0, u1 = rng.uniform(size=n), rng.uniform(size=n)
    c = 1 + np.log(u0 + (1 - u0) * np.exp(-2 * kappa)) / kappa
    s = np.sqrt(np.maximum(0, 1 - c * c))
    t = np.cross(mu, np.where(np.abs(mu[:, :1]) > 0.9, [[0, 1, 0]], [[1, 0, 0]]))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(mu, t)
    w = mu * c[:, None] + t * (s * np.cos(2 * np.pi * u1))[:, None] + b * (s * np.sin(2 * np.pi * u1))[:, None]
    ent = -np.mean(np.log(kappa / (2 * np.pi * -np.expm1(-2 * kappa)) * np.exp(kappa * (c - 1))))
    return np.hstack([x, d, w, np.ones((n, 1))]).astype(np.float32), ent


def test_training_learns_a_position_dependent_lobe():
    rng = np.random.default_rng(3)
    m = gpu.Mlp(seed=3)
    val, ent = _target_batch(rng, 8192)
    start = m.loss_grad(val)[0]
    for _ in range(300):
        batch, _ = _target_batch(rng, 8192)
        m.train_step(batch, lr=3e-3)
    end = m.loss_grad(val)[0]
    # an untrained mixture is near uniform (loss ~ log 4 pi = 2.53); the target's entropy is
    # about -1.6; the trained proposal must get most of the way there
    assert start > 1.5 and end < start - 2.5 and end < ent + 1.0, (start, end, ent)
