"""Batched jobs (GPU): one iteration of each of several independent renderers of uneven size,
driven side by side, gives each renderer the image it renders on its own."""
import numpy as np
import pytest

from conftest import box_args
from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu

pytestmark = pytest.mark.gpu

JOBS = [("box", "48x48", "5"), ("boxfog", "64x40", "8"), ("boxfog", "32x32", "4"), ("box", "80x64", "12"),
        ("boxfog", "96x72", "6"), ("box", "40x56", "9")]


def _renderers(scene_dir):
    out = []
    for kind, res, length in JOBS:
        obj = str(scene_dir / kind / (kind + ".obj"))
        rc, cfg = capi.parse_args(["-s", "-1", obj] + box_args(res, extra=("-r_initial", "0.03", "-l", length)))
        assert rc == 0
        sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
        out.append((sh, gpu.Renderer(sh.ptr, cfg)))
    return out


def test_batch_matches_each_job_alone(scene_dir):
    alone = []
    for sh, r in _renderers(scene_dir):
        for i in range(3):
            r.run_iteration(i, want_stats=False)
        alone.append(r.read_accum(3))
        r.close()
    jobs = _renderers(scene_dir)
    rs = [r for _, r in jobs]
    for i in range(3):
        gpu.run_batch(rs, [i] * len(rs), threads=4)
    for (sh, r), ref in zip(jobs, alone):
        img = r.read_accum(3)
        r.close()
        # the same paths; only the order of floating-point atomic sums may differ
        rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-3)
        assert np.mean(rel < 1e-4) > 0.99 and abs(img.mean() - ref.mean()) / ref.mean() < 1e-5
