"""Learned scattering proposal, host side (no GPU): the oracle's von Mises-Fisher mixture and its
use inside the renderer's MIS, pinned by what the mathematics fixes.

* The mixture pdf must be the density its sampler follows: E[1/pdf] = 4 pi and
  E[f/pdf] = the integral of f, under the mixture's own samples.
* One lobe has the closed form kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1)).
* alpha = 0 is the plain renderer bit for bit.
* With alpha > 0 the renderer stays unbiased: its converged image matches the plain one. A
  proposal pdf placed in the wrong recursion (say the light technique's side) breaks the MIS
  partition of unity and shows up here as a biased image.
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


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mixture_sampler_follows_its_pdf(seed):
    rng = np.random.default_rng(seed)
    lobes = random_lobes(rng)
    out = oracle_py.guide_samples(lobes, strat(110 ** 3, rng)).astype(np.float64)
    d, pdf = out[:, :3], out[:, 3]
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-5) and (pdf > 0).all()
    # E[1/pdf] = 4 pi only if the samples follow the reported density everywhere it is positive
    assert abs(np.mean(1 / pdf) - 4 * np.pi) / (4 * np.pi) < 0.02
    z = d[:, 2]
    assert abs(np.mean((1 + z) ** 2 / pdf) - 16 * np.pi / 3) / (16 * np.pi / 3) < 0.02


def test_single_lobe_closed_form():
    for kappa in (0.05, 1.0, 30.0, 900.0):
        lobes = np.zeros((8, 5), np.float32)
        lobes[:, 2] = 1.0
        lobes[:, 3] = kappa
        lobes[0, 4] = 1.0  # only lobe 0 carries weight
        u = np.array([[0.0, 1e-7, 0.3], [0.0, 1 - 1e-7, 0.3]], np.float32)
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
    lobes = np.stack([random_lobes(rng, kappa_hi) for _ in range(res ** 3 * 6)])
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
    # A deliberately poor proposal (random sharp lobes, alpha 0.9) must converge to the plain
    # image. Same iterations, so the same light paths and primary rays: the two estimates differ
    # only where guided camera paths scatter elsewhere, which is zero-mean noise, and the image
    # means agree to ~0.1% (the iteration-to-iteration noise is ~4%). Evaluating the camera
    # technique at light vertices with the phase function instead of the mixture moves the mean
    # by 0.6-0.8%.
    args = box_args("16x16", alg=alg, extra=("-r_initial", "0.05", "-l", "6"))
    plain, _, sh = _oracle(boxfog_obj, args, 64)
    lobes, lo, hi = guide_for(sh, 3, np.random.default_rng(7), kappa_hi=200.0)
    oracle_py.set_guide(lobes, 3, lo, hi, 0.9)
    try:
        guided, _, _ = _oracle(boxfog_obj, args, 64)
    finally:
        oracle_py.set_guide(None, 0, None, None, 0.0)
    assert np.isfinite(guided).all() and (guided >= 0).all()
    assert abs(guided.mean() - plain.mean()) / plain.mean() < 0.003, (guided.mean(), plain.mean())
    gb = guided.reshape(4, 4, 4, 4, 3).mean(axis=(1, 3))
    pb = plain.reshape(4, 4, 4, 4, 3).mean(axis=(1, 3))
    assert np.max(np.abs(gb - pb) / (pb + 1e-2)) < 0.05
