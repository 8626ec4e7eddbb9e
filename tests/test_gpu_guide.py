"""GPU parity of the learned scattering proposal: the vMF product sampler against the oracle's,
and renders with a proposal installed, path for path against the oracle."""
import numpy as np
import pytest

from conftest import assert_whole_image, box_args, oracle_independent
from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu
from oracle import oracle_py

pytestmark = pytest.mark.gpu


def _lobes(rng, n, kappa_hi=200.0):
    mu = rng.normal(size=(n * 8, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    kappa = np.exp(rng.uniform(np.log(0.05), np.log(kappa_hi), (n * 8, 1)))
    w = rng.uniform(0.05, 1.0, (n * 8, 1))
    return np.hstack([mu, kappa, w]).astype(np.float32).reshape(n, 8, 5)


@pytest.mark.parametrize("mean_cosine", [0.0, 0.9, -0.5])
def test_product_sampler_matches_oracle(mean_cosine):
    rng = np.random.default_rng(2)
    for lobes in _lobes(rng, 4):
        u = rng.uniform(0, 1, (20000, 3)).astype(np.float32)
        arrive = rng.normal(size=3)
        arrive /= np.linalg.norm(arrive)
        g = gpu.guide_samples(lobes, u, arrive, mean_cosine)
        o = oracle_py.guide_samples(lobes, u, arrive, mean_cosine)
        # same FP32 operations; expf / logf / sinf / cosf may differ in an ulp, which the
        # inversion 1 + log(u + (1 - u) exp(-2 kappa)) / kappa amplifies for small kappa
        dd = np.max(np.abs(g[:, :3] - o[:, :3]), axis=1)
        assert np.mean(dd < 2e-5) > 0.999 and dd.max() < 1e-2, (np.mean(dd < 2e-5), dd.max())
        rp = np.abs(g[:, 3] - o[:, 3]) / o[:, 3]
        assert np.mean(rp < 1e-3) > 0.999, np.mean(rp < 1e-3)


def test_surface_product_sampler_matches_oracle():
    rng = np.random.default_rng(3)
    for lobes in _lobes(rng, 3):
        u = rng.uniform(0, 1, (20000, 3)).astype(np.float32)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        g = gpu.guide_samples(lobes, u, n, surface=True)
        o = oracle_py.guide_samples(lobes, u, n, surface=True)
        dd = np.max(np.abs(g[:, :3] - o[:, :3]), axis=1)
        assert np.mean(dd < 2e-5) > 0.999 and dd.max() < 1e-2, (np.mean(dd < 2e-5), dd.max())
        rp = np.abs(g[:, 3] - o[:, 3]) / o[:, 3]
        assert np.mean(rp < 1e-3) > 0.999, np.mean(rp < 1e-3)


def _render(obj, args, lobes, res, lo, hi, alpha):
    rc, cfg = capi.parse_args(["-s", "-1", obj] + args)
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    with gpu.Renderer(sh.ptr, cfg) as r:
        r.set_guide_lobes(lobes, res, lo, hi, alpha)
        st = r.run_iteration(0)
        g = r.read_accum(1)
    oracle_py.set_guide(lobes, res, lo, hi, alpha)
    try:
        o, ost = oracle_py.render(sh.ptr, cfg, 0, 1)
        o2 = oracle_independent(sh.ptr, cfg, 0, 1)
    finally:
        oracle_py.set_guide(None, 0, None, None, 0.0)
    return g, o, st, ost, o2


@pytest.mark.parametrize("alpha", [0.0, 0.7])
def test_guided_render_path_parity(boxfog_obj, alpha):
    rc, cfg = capi.parse_args(["-s", "-1", boxfog_obj])
    sh = capi.SceneHandle(boxfog_obj, 64, 64)
    c, r = np.array(list(sh.scene.scene_center), np.float32), sh.scene.scene_radius
    lobes = _lobes(np.random.default_rng(11), 3 ** 3, kappa_hi=60.0)
    g, o, st, ost, o2 = _render(boxfog_obj, box_args("48x48", extra=("-r_initial", "0.03")), lobes, 3, c - r, c + r,
                            alpha)
    assert st.path_errors == 0 and st.surface_vertices + st.medium_vertices == ost.light_vertices
    rel = np.abs(g - o) / np.maximum(np.abs(o), 1e-3)
    assert np.mean(rel < 1e-3) > 0.93, np.mean(rel < 1e-3)
    assert abs(g.mean() - o.mean()) / o.mean() < 0.01
    assert_whole_image(g, o, o2)
