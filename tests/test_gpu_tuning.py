"""Tuning knobs change speed only (GPU): every setting of sub-beam length, query piece length,
merge blocks per CU, piece order, packet traversal, leaf data, bounce-kernel block size and piece
bundles accepts the same photon and beam pairs and renders the same image as the defaults."""
import numpy as np
import pytest

from conftest import box_args
from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu

pytestmark = pytest.mark.gpu

SETTINGS = [dict(subbeam_factor=8.0, query_piece_factor=32.0), dict(subbeam_factor=24.0, query_piece_factor=128.0),
            dict(subbeam_factor=0.0, query_piece_factor=0.0), dict(merge_blocks_per_cu=1), dict(piece_order=1),
            dict(piece_order=2), dict(packet_bounces=0), dict(packet_bounces=64),
            dict(packet_bounces=64, piece_order=2), dict(packet_bounces=64, subbeam_factor=0.0, query_piece_factor=0.0),
            dict(piece_order=3), dict(piece_order=4), dict(piece_order=3, graph=1), dict(beam_leaf=1),
            dict(beam_leaf=1, piece_order=3, packet_bounces=64), dict(block_size=256), dict(block_size=128),
            dict(bundle_bounces=0), dict(bundle_bounces=64, bundle_factor=1e6), dict(bundle_bounces=3, bundle_factor=2.0)]


def _render(obj, tuning):
    rc, cfg = capi.parse_args(["-s", "-1", obj] + box_args("64x64", extra=("-r_initial", "0.03")))
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    with gpu.Renderer(sh.ptr, cfg) as r:
        if tuning:
            r.set_tuning(**tuning)
        st = [r.run_iteration(i) for i in range(2)]
        img = r.read_accum(2)
    return img, [(s.pb2d_accepts, s.bb1d_accepts, s.merge_accepts, s.path_errors) for s in st]


@pytest.mark.parametrize("tuning", SETTINGS)
def test_tuning_changes_speed_only(boxfog_obj, tuning):
    ref, ref_counts = _render(boxfog_obj, None)
    img, counts = _render(boxfog_obj, tuning)
    assert counts == ref_counts and all(c[3] == 0 for c in counts)
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-3)
    assert np.mean(rel < 1e-4) > 0.99 and abs(img.mean() - ref.mean()) / ref.mean() < 1e-5
