"""Learned scattering proposal, host side (no GPU): the oracle's product of a von Mises-Fisher
mixture with the phase lobe, and its use inside the renderer's MIS, pinned by what the
mathematics fixes.

* The product pdf must be the density its sampler follows: E[1/pdf] = 4 pi and
  E[f/pdf] = the integral of f, under the product's own samples, for isotropic, forward and
  backward phase lobes.
* One lobe has the closed form kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1)); the
  product of two lobes is the lobe of v = kappa_1 mu_1 + kappa_2 mu_2, and the product pdf is
  the mixture times the phase lobes, normalised (quadrature over the sphere).
* The fitted phase lobes follow the Henyey-Greenstein function (chi-square ratio near 1).
* alpha = 0 is the plain renderer bit for bit.
* With alpha > 0 (in media and at surfaces) the renderer stays unbiased: its converged image
  matches the plain one. A proposal pdf placed in the wrong recursion (say the light
  technique's side) breaks the MIS partition of unity and shows up here as a biased image.
"""
import numpy as np
import pytest

from conftest import box_args
from gpuupbp_amd import capi
from oracle import oracle_py


def random_lobes(rng, kappa_hi=200.0):
    mu = rng.normal(size=(8, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    kappa = np.exp(rng.uniform(np.log(0.05), np.log(kappa_hi), (8, 1)))
    w = rng.uniform(0.05, 1.0, (8, 1))
    return np.hstack([mu, kappa, w]).astype(np.float32)


def strat(n, rng):
    k = int(round(n ** (1 / 3)))
    g = (np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3) +
         rng.uniform(size=(k ** 3, 3))) / k
    return g.astype(np.float32)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def fibonacci_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = np.pi * (1 + 5 ** 0.5) * i
    s = np.sqrt(1 - z * z)
    return np.stack([np.cos(phi) * s, np.sin(phi) * s, z], 1)


def vmf(w, mu, k):
    return k / (2 * np.pi * -np.expm1(-2 * k)) * np.exp(k * (w @ mu - 1))


def phase_table():
    """The fitted phase lobes, read from the generated header: rows (k0, k1, k2, w0, w1, w2)."""
    import os
    import re
    path = os.path.join(os.path.dirname(__file__), "..", "gpuupbp_amd", "csrc", "device", "phase_lobes.h")
    rows = [[float(v) for v in re.findall(r"([-0-9.e+]+)f", line.split("//")[0])]
            for line in open(path) if line.strip().startswith("{") and "// g" in line]
    return np.array(rows)


def phase_lobes(g):
    """Concentrations and weights of the phase lobes at mean cosine g (linear between rows)."""
    tab = phase_table()
    t = min(abs(g), 0.99) / 0.01
    i = min(int(t), len(tab) - 2)
    f = t - i
    row = (1 - f) * tab[i] + f * tab[i + 1]
    return row[:3], row[3:]


def hg(mu, g):
    return (1 - g * g) / (4 * np.pi * (1 + g * g - 2 * g * mu) ** 1.5)


def product_pdf(lobes, arrive, g, quad):
    """q(w) h(w) / Z with h the fitted phase lobes about the arrival and Z by quadrature,
    straight from the definition."""
    ks, ws = phase_lobes(g)
    mu_h = arrive if g >= 0 else -arrive
    wts = lobes[:, 4] / lobes[:, 4].sum()
    mus = lobes[:, :3] / np.linalg.norm(lobes[:, :3], axis=1, keepdims=True)
    q = lambda x: sum(wts[i] * vmf(x, mus[i], float(lobes[i, 3])) for i in range(8))
    h = lambda x: sum(ws[j] * vmf(x, mu_h, ks[j]) for j in range(3))
    z = np.mean(q(quad) * h(quad)) * 4 * np.pi
    return lambda x: q(x) * h(x) / z


def test_phase_lobes_follow_the_phase_function():
    # the fitted mixture against the Henyey-Greenstein function itself: chi-square ratio
    # E_HG[HG / h] close to 1 (one lobe of matching mean cosine gives ~1000 at g = 0.9)
    mu = np.sort(1 - np.concatenate([[0.0], np.geomspace(1e-9, 2.0, 20000)]))
    for g, bound in ((0.0, 1.001), (0.3, 1.01), (0.5, 1.01), (0.9, 1.05), (0.95, 1.1), (-0.6, 1.02)):
        ks, ws = phase_lobes(g)
        # about the lobe axis, as functions of cos = mu
        h = sum(w * k / (2 * np.pi * -np.expm1(-2 * k)) * np.exp(k * (mu - 1)) for w, k in zip(ws, ks))
        p = hg(mu, abs(g))
        assert abs(np.trapezoid(h, mu) * 2 * np.pi - 1) < 1e-3
        chi = np.trapezoid(p * p / h, mu) * 2 * np.pi
        assert chi < bound, (g, chi)


@pytest.mark.parametrize("seed,g", [(0, 0.0), (1, 0.9), (2, -0.5), (3, 0.6)])
def test_product_sampler_follows_its_pdf(seed, g):
    rng = np.random.default_rng(seed)
    lobes = random_lobes(rng, kappa_hi=100.0)
    arrive = unit(rng.normal(size=3))
    out = oracle_py.guide_samples(lobes, strat(100 ** 3, rng), arrive, g).astype(np.float64)
    d, pdf = out[:, :3], out[:, 3]
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-5) and (pdf > 0).all()
    # the reported pdf is the normalised product, from its definition
    quad = fibonacci_sphere(400000)
    ref = product_pdf(lobes, arrive, g, quad)
    rel = np.abs(pdf - ref(d)) / ref(d)
    assert np.max(rel) < 3e-3, np.max(rel)
    # the samples follow it: first and second moments against quadrature of the pdf
    pq = ref(quad)
    for phi in (lambda x: x[:, 0], lambda x: x[:, 1], lambda x: x[:, 2], lambda x: x[:, 0] * x[:, 1],
                lambda x: x[:, 2] ** 2, lambda x: (x @ arrive > 0.8).astype(np.float64)):
        assert abs(np.mean(phi(d)) - np.mean(phi(quad) * pq) * 4 * np.pi) < 3e-3
    if g == 0.0:  # E[1/pdf] = 4 pi when the density is bounded away from zero
        assert abs(np.mean(1 / pdf) - 4 * np.pi) / (4 * np.pi) < 0.03


def cosine_kappa():
    import os
    import re
    path = os.path.join(os.path.dirname(__file__), "..", "gpuupbp_amd", "csrc", "device", "phase_lobes.h")
    return float(re.search(r"kCosineKappa = ([-0-9.e+]+)f", open(path).read()).group(1))


def test_cosine_lobe_follows_the_clamped_cosine():
    # the fitted lobe against max(cos, 0) / pi: same mean cosine to a few percent
    k = cosine_kappa()
    assert abs(1 / np.tanh(k) - 1 / k - 2 / 3) < 0.03


@pytest.mark.parametrize("seed", [5, 6])
def test_surface_product_sampler_follows_its_pdf(seed):
    rng = np.random.default_rng(seed)
    lobes = random_lobes(rng, kappa_hi=100.0)
    n = unit(rng.normal(size=3))
    out = oracle_py.guide_samples(lobes, strat(100 ** 3, rng), n, 0.0, surface=True).astype(np.float64)
    d, pdf = out[:, :3], out[:, 3]
    quad = fibonacci_sphere(400000)
    k = cosine_kappa()
    wts = lobes[:, 4] / lobes[:, 4].sum()
    mus = lobes[:, :3] / np.linalg.norm(lobes[:, :3], axis=1, keepdims=True)
    q = lambda x: sum(wts[i] * vmf(x, mus[i], float(lobes[i, 3])) for i in range(8))
    ref = lambda x: q(x) * vmf(x, n, k)
    z = np.mean(ref(quad)) * 4 * np.pi
    rel = np.abs(pdf - ref(d) / z) / (ref(d) / z)
    assert np.max(rel) < 3e-3, np.max(rel)
    pq = ref(quad) / z
    for phi in (lambda x: x[:, 0], lambda x: x[:, 1], lambda x: x[:, 2], lambda x: (x @ n > 0).astype(np.float64)):
        assert abs(np.mean(phi(d)) - np.mean(phi(quad) * pq) * 4 * np.pi) < 3e-3


def test_product_of_one_lobe_is_three_lobes():
    rng = np.random.default_rng(4)
    arrive = unit([0.3, -0.5, 0.8])
    one = np.zeros((8, 5), np.float32)
    one[:, :3] = unit([1.0, 0.2, -0.3])
    one[:, 3] = 7.0
    one[0, 4] = 1.0
    out = oracle_py.guide_samples(one, strat(20 ** 3, rng), arrive, 0.9).astype(np.float64)
    # lobe x phase lobe j is the vMF of v_j = 7 mu + k_j mu_h, with mass w_j C(7) C(k_j) / C(|v_j|),
    # C(k) = k / (4 pi sinh k)
    ks, ws = phase_lobes(0.9)
    C = lambda k: k / (4 * np.pi * np.sinh(k))
    vs = [7.0 * unit([1.0, 0.2, -0.3]) + k * arrive for k in ks]
    mass = np.array([w * C(7.0) * C(k) / C(np.linalg.norm(v)) for w, k, v in zip(ws, ks, vs)])
    ref = sum(m / mass.sum() * vmf(out[:, :3], v / np.linalg.norm(v), np.linalg.norm(v)) for m, v in zip(mass, vs))
    assert np.max(np.abs(out[:, 3] - ref) / ref) < 1e-3


def test_single_lobe_closed_form():
    for kappa in (0.05, 1.0, 30.0, 900.0):
        lobes = np.zeros((8, 5), np.float32)
        lobes[:, 2] = 1.0
        lobes[:, 3] = kappa
        lobes[0, 4] = 1.0  # only lobe 0 carries weight
        u = np.array([[0.0, 1e-7, 0.3], [0.0, 1 - 1e-7, 0.3]], np.float32)  # g = 0: the mixture itself
        out = oracle_py.guide_samples(lobes, u)
        c = kappa / (2 * np.pi * -np.expm1(-2 * kappa))
        # u0 -> 1 gives mu itself, u0 -> 0 the antipode
        assert out[1, 2] == pytest.approx(1.0, abs=1e-4)
        assert out[1, 3] == pytest.approx(c, rel=2e-3)
        if kappa < 5:
            assert out[0, 2] == pytest.approx(-1.0, abs=1e-3)
            assert out[0, 3] == pytest.approx(c * np.exp(-2 * kappa), rel=5e-3)


def guide_for(sh, res, rng, kappa_hi=40.0):
    s = sh.scene
    c = np.array(list(s.scene_center), np.float32)
    r = s.scene_radius
    lobes = np.stack([random_lobes(rng, kappa_hi) for _ in range(res ** 3)])
    return lobes, c - r, c + r


def _oracle(obj, args, iters, first=0, threads=0):
    rc, cfg = capi.parse_args(["-s", "-1", obj] + args)
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    img, st = oracle_py.render(sh.ptr, cfg, first, iters, threads=threads)
    return img, st, sh


def test_alpha_zero_is_the_plain_renderer(boxfog_obj):
    args = box_args("24x24", extra=("-r_initial", "0.03"))
    # one thread: the threaded oracle adds light-tracing splats in a varying order
    plain, _, sh = _oracle(boxfog_obj, args, 1, threads=1)
    lobes, lo, hi = guide_for(sh, 2, np.random.default_rng(1))
    oracle_py.set_guide(lobes, 2, lo, hi, 0.0)
    try:
        same, _, _ = _oracle(boxfog_obj, args, 1, threads=1)
    finally:
        oracle_py.set_guide(None, 0, None, None, 0.0)
    assert np.array_equal(plain, same)


@pytest.mark.parametrize("alg", ["upbp_bpt", "upbp_bpt+surf+pp3d+pb2d+bb1d+bp2d"])
def test_mixture_keeps_the_estimator_unbiased(boxfog_obj, alg):
    # A deliberately poor proposal (random sharp lobes, alpha 0.9 in the medium and at the
    # diffuse and glossy walls) must converge to the plain image. Same iterations, so the same
    # light paths and primary rays: the two estimates differ only where guided camera paths
    # scatter elsewhere, which is zero-mean noise; over 256 iterations the image means agree to
    # ~0.15% (another seed moves the plain mean by ~0.1%). Evaluating the camera technique at
    # surface light vertices with the BSDF instead of the proposal moves the mean by ~1%, at
    # medium light vertices with the phase function by 0.6-0.8%.
    args = box_args("16x16", alg=alg, extra=("-r_initial", "0.05", "-l", "6"))
    plain, _, sh = _oracle(boxfog_obj, args, 256)
    lobes, lo, hi = guide_for(sh, 3, np.random.default_rng(7), kappa_hi=200.0)
    oracle_py.set_guide(lobes, 3, lo, hi, 0.9)
    try:
        guided, _, _ = _oracle(boxfog_obj, args, 256)
    finally:
        oracle_py.set_guide(None, 0, None, None, 0.0)
    assert np.isfinite(guided).all() and (guided >= 0).all()
    assert abs(guided.mean() - plain.mean()) / plain.mean() < 0.004, (guided.mean(), plain.mean())
    gb = guided.reshape(4, 4, 4, 4, 3).mean(axis=(1, 3))
    pb = plain.reshape(4, 4, 4, 4, 3).mean(axis=(1, 3))
    assert np.max(np.abs(gb - pb) / (pb + 1e-2)) < 0.05
