"""The oracle's restatement of the learned proposal's training samples (CPU): which scatterings
give samples and where they lie. The weight itself (the luminance a path collects after the
scattering over lum(T f cos) / cp) is this library's own definition, with no counterpart in the
reference; the GPU test of the same samples checks the GPU against it."""
import numpy as np
import pytest

from conftest import box_args
from gpuupbp_amd import capi
from oracle import oracle_py


@pytest.fixture(scope="module")
def fogbox(boxfog_obj):
    rc, cfg = capi.parse_args(["-s", "-1", boxfog_obj] + box_args("32x32", extra=("-r_initial", "0.03")))
    assert rc == 0
    sh = capi.SceneHandle(boxfog_obj, cfg.res_x, cfg.res_y)
    yield sh, cfg
    sh.close()


def _keys(a):
    return set(map(tuple, np.round(a[:, :7], 5)))


def test_samples_come_from_the_scatterings_that_train(fogbox):
    sh, cfg = fogbox
    both = oracle_py.training_samples(sh.ptr, cfg, 0, True, True)
    media = oracle_py.training_samples(sh.ptr, cfg, 0, True, False)
    surfaces = oracle_py.training_samples(sh.ptr, cfg, 0, False, True)
    assert len(oracle_py.training_samples(sh.ptr, cfg, 0, False, False)) == 0
    assert len(media) > 50 and len(surfaces) > 50
    # a scattering is in a medium or at a surface: the two sets partition the samples
    assert len(both) == len(media) + len(surfaces)
    assert _keys(both) == _keys(media) | _keys(surfaces)
    # media samples lie in the fog sphere (radius 0.55 about the origin), surface samples on the
    # walls of the box [-1, 1]^3 (the only scattering surfaces)
    assert np.all(np.linalg.norm(media[:, 1:4], axis=1) < 0.55 + 1e-4)
    assert np.all(np.abs(np.max(np.abs(surfaces[:, 1:4]), axis=1) - 1.0) < 1e-3)
    # unit scattered directions, positive finite weights, one pixel per camera path
    assert np.allclose(np.linalg.norm(both[:, 4:7], axis=1), 1.0, atol=1e-5)
    assert np.all(both[:, 7] > 0.0) and np.isfinite(both).all()
    assert np.all((both[:, 0] >= 0) & (both[:, 0] < cfg.res_x * cfg.res_y))


def test_samples_are_deterministic(fogbox):
    sh, cfg = fogbox
    a = oracle_py.training_samples(sh.ptr, cfg, 3, True, True, threads=4)
    b = oracle_py.training_samples(sh.ptr, cfg, 3, True, True, threads=1)
    assert np.array_equal(a, b)
