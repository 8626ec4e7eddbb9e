"""The learned proposal's network, host side (no GPU): the FP64 oracle's analytic gradient of
the weighted mixture NLL is pinned by central finite differences, and the loss by the closed
form of a single-lobe target."""
import numpy as np

from oracle import oracle_py


def he_params(seed):
    rng = np.random.default_rng(seed)
    shapes = [(64, 51), (64,), (64, 64), (64,), (64, 64), (64,), (40, 64), (40,)]
    out = []
    for s in shapes:
        out.append(rng.normal(scale=np.sqrt(2.0 / s[1]), size=s).ravel() if len(s) == 2 else rng.normal(scale=0.1, size=s))
    p = np.concatenate(out)
    assert p.size == oracle_py.mlp_param_count() == 14248
    return p


def training_samples(rng, n):
    x = rng.uniform(0, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    W = rng.uniform(0.1, 2.0, (n, 1))
    return np.hstack([x, d, w, W]).astype(np.float32)


def test_encoding_is_the_fourier_features():
    x = np.array([0.3, 0.71, 0.05], np.float32)
    d = np.array([0.6, -0.8, 0.0], np.float32)
    f = oracle_py.mlp_encode(x, d)
    for j in range(3):
        for k in range(8):
            a = 2 ** k * np.pi * float(x[j])
            assert abs(f[(j * 8 + k) * 2] - np.sin(a)) < 1e-6 and abs(f[(j * 8 + k) * 2 + 1] - np.cos(a)) < 1e-6
    assert np.allclose(f[48:], d)


def test_analytic_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = he_params(1)
    s = training_samples(rng, 6)
    _, g = oracle_py.mlp_loss_grad(p, s)
    # every layer's weights and biases, a few entries each
    idx = [0, 17, 50, 3263, 3264 + 5, 3328 + 100, 7424 + 7, 7488 + 3000, 11584 + 63, 11648 + 1234, 14208 + 39, 14247]
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle_py.mlp_loss_grad(pp, s, False)[0] - oracle_py.mlp_loss_grad(pm, s, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_loss_of_a_single_lobe():
    # zero weights except the output biases: every input gives lobe 0 (mu = +z, kappa = softplus(b) + 0.01)
    p = np.zeros(14248)
    b4 = p[14208:]
    b4[2] = 1.0   # mu of lobe 0 along +z
    b4[3] = 3.0   # raw kappa
    b4[4] = 30.0  # lobe 0 takes all the weight
    s = np.array([[0.5, 0.5, 0.5, 0, 0, 1, 0, 0, 1, 1.0], [0.5, 0.5, 0.5, 0, 0, 1, 1, 0, 0, 2.0]], np.float32)
    loss, _ = oracle_py.mlp_loss_grad(p, s)
    k = np.log1p(np.exp(3.0)) + 0.01
    c = k / (2 * np.pi * -np.expm1(-2 * k))
    want = (1.0 * -np.log(c) + 2.0 * -(np.log(c) - k)) / 2
    assert abs(loss - want) < 1e-6 * abs(want)


# ---- FP64 Adam (oracle_adam_step) pinned by what Adam's bias correction fixes ---------------
def test_adam_first_step_moves_every_parameter_by_lr():
    # step 1: m_hat = g and v_hat = g^2 exactly, so the update is lr * g / (|g| + eps)
    rng = np.random.default_rng(0)
    g = rng.normal(size=1000) * np.exp(rng.uniform(-8, 2, 1000))
    p0 = rng.normal(size=1000)
    p1, m, v = oracle_py.adam_step(p0, np.zeros(1000), np.zeros(1000), g, 1, 1e-3)
    assert np.allclose(p0 - p1, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-9, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_adam_constant_gradient_keeps_the_step_size():
    # with a constant gradient the bias-corrected moments stay exactly g and g^2 at every step,
    # so each step moves by lr * sign(g); without the correction the first steps are 3.16x larger
    rng = np.random.default_rng(1)
    g = rng.normal(size=200)
    p, m, v = rng.normal(size=200), np.zeros(200), np.zeros(200)
    for t in range(1, 40):
        q, m, v = oracle_py.adam_step(p, m, v, g, t, 2e-3)
        assert np.allclose(p - q, 2e-3 * g / (np.abs(g) + 1e-8), rtol=1e-7)
        p = q


def test_adam_sign_flip_damps_the_step():
    # gradient +g then -g: m_hat(2) = (0.09 g - 0.1 g) / 0.19, v_hat(2) = g^2, a step of lr / 19
    g = np.array([0.5, -2.0, 3.0])
    p, m, v = oracle_py.adam_step(np.zeros(3), np.zeros(3), np.zeros(3), g, 1, 1.0)
    q, m, v = oracle_py.adam_step(p, m, v, -g, 2, 1.0)
    assert np.allclose(q - p, np.sign(g) / 19.0, rtol=1e-6)
