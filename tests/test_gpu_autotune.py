"""The autotuner (GPU): it returns a tuning from its grid, leaves the accumulated image as it
was, and the renderer continues to the same image as without it (the knobs change speed only)."""
import numpy as np
import pytest

from conftest import box_args
from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu

pytestmark = pytest.mark.gpu


def test_autotune_keeps_the_image(boxfog_obj):
    rc, cfg = capi.parse_args(["-s", "-1", boxfog_obj] + box_args("64x64", extra=("-r_initial", "0.03")))
    assert rc == 0
    sh = capi.SceneHandle(boxfog_obj, cfg.res_x, cfg.res_y)
    with gpu.Renderer(sh.ptr, cfg) as r:
        for i in range(3):
            r.run_iteration(i, want_stats=False)
        plain = r.read_accum(3)
    with gpu.Renderer(sh.ptr, cfg) as r:
        r.run_iteration(0, want_stats=False)
        one = r.read_accum(1)
        tuning, ms = r.autotune(iteration=1, repeats=1)
        assert np.array_equal(r.read_accum(1), one)  # trial iterations leave no trace
        assert tuning["subbeam_factor"] in (8.0, 12.0, 16.0, 24.0) and tuning["query_piece_factor"] in (32.0, 64.0, 128.0)
        assert tuning["merge_blocks_per_cu"] in (0, 1, 2) and tuning["piece_order"] in (0, 2) and ms > 0
        assert tuning["packet_bounces"] in (0, 1, 2, 64) and tuning["bundle_bounces"] in (0, 1, 2)
        for i in (1, 2):
            r.run_iteration(i, want_stats=False)
        tuned = r.read_accum(3)
    # same accepted pairs; only the floating-point summation order may differ
    rel = np.abs(tuned - plain) / np.maximum(np.abs(plain), 1e-3)
    assert np.mean(rel < 1e-4) > 0.99 and abs(tuned.mean() - plain.mean()) / plain.mean() < 1e-5, rel.max()
