"""B-B1D bundles (GPU): walking a camera wave's j-th query pieces as one bundle (one beam-tree walk
per bundle, every ray testing the sub-beams it reaches) accepts the same beam pairs and renders the
same image as the per-piece walk -- whether every bundle is walked whole, none is (all handed to
the per-piece walk) or some are, on the first bounce or on all of them."""
import numpy as np
import pytest

from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu
from gpuupbp_amd import workloads as wl

pytestmark = pytest.mark.gpu

OFF = dict(bundle_bounces=0)
SETTINGS = [dict(bundle_bounces=1), dict(bundle_bounces=64, bundle_factor=1e6), dict(bundle_bounces=64, bundle_factor=4.0),
            dict(bundle_bounces=2, bundle_factor=0.0), dict(bundle_bounces=64, bundle_factor=16.0, graph=1),
            dict(bundle_bounces=1, bundle_piece_factor=128.0)]


def _render(obj, args, tuning, iters=2):
    rc, cfg = capi.parse_args(["-s", "-1", obj] + args)
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    with gpu.Renderer(sh.ptr, cfg) as r:
        r.set_tuning(**tuning)
        st = [r.run_iteration(i) for i in range(iters)]
        img = r.read_accum(iters)
    sh.close()
    return img, st


def _counts(st):
    return [(s.pb2d_accepts, s.bb1d_accepts, s.merge_accepts, s.path_errors) for s in st]


@pytest.fixture(scope="module")
def blobs_obj(tmp_path_factory):
    return wl.scene(str(tmp_path_factory.mktemp("bundle")), "blobs", 1)


@pytest.mark.parametrize("tuning", SETTINGS)
def test_bundles_accept_the_per_piece_pairs(blobs_obj, tuning):
    args = wl.scaled(wl.BLOBS_ARGS, 8)  # 160x90, 14,062 light paths
    ref, rst = _render(blobs_obj, args, OFF)
    img, st = _render(blobs_obj, args, tuning)
    print(tuning, "bundles", [s.bundles for s in st], "left", [s.bundle_pieces_left for s in st], "accepts",
          [s.bb1d_accepts for s in st])
    assert _counts(st) == _counts(rst) and all(s.path_errors == 0 for s in st)
    assert all(s.bundles == 0 and s.bundle_pieces_left == 0 for s in rst)
    if tuning.get("bundle_factor", 32.0) > 0.0:
        assert all(s.bundles > 0 and s.bundle_tests > 0 for s in st)
    else:  # only bundles of one piece (or of one line) are narrow enough: the rest go per piece
        assert all(s.bundle_pieces_left > 0 for s in st)
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-3)
    assert np.isfinite(img).all()
    assert np.mean(rel < 1e-4) > 0.99 and abs(img.mean() - ref.mean()) / ref.mean() < 1e-5


def test_bundles_in_a_dense_fog(boxfog_obj):
    # every bounce bundled, in the fog box at a large radius: many sub-beams per bundle
    args = ["-r", "96x64", "-l", "8", "-r_initial", "0.03", "-a", "upbp_all"]
    ref, rst = _render(boxfog_obj, args, OFF, iters=3)
    img, st = _render(boxfog_obj, args, dict(bundle_bounces=64, bundle_factor=1e6), iters=3)
    assert _counts(st) == _counts(rst)
    assert all(s.bundles > 0 for s in st)
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-3)
    assert np.mean(rel < 1e-4) > 0.99 and abs(img.mean() - ref.mean()) / ref.mean() < 1e-5
