"""The gpuupbp CLI (the drop-in for the reference's main) on the GPU: iteration loop,
continuous output names and the final EXR, compared with the library's own mean image."""
import os
import subprocess

import numpy as np
import pytest

from conftest import box_args
from gpuupbp_amd import capi
from gpuupbp_amd.renderer import Renderer
from test_host_cpu import _read_exr

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cli_renders_and_writes_exr(box_obj, tmp_path):
    exe = os.path.join(ROOT, "gpuupbp_amd", "gpuupbp")
    out = str(tmp_path / "img.exr")
    args = ["-s", "-1", box_obj] + box_args("40x32") + ["-i", "4", "-continuous_output", "2", "-o", out]
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "done in" in r.stdout and "100%" in r.stdout
    assert os.path.exists(str(tmp_path / "img-2.exr")) and os.path.exists(str(tmp_path / "img-4.exr"))
    rc, cfg = capi.parse_args(args)
    sh = capi.SceneHandle(box_obj, cfg.res_x, cfg.res_y)
    with Renderer(sh.ptr, cfg) as rr:
        for i in range(4):
            rr.run_iteration(i, want_stats=False)
        ref = rr.read_accum(4)
    _, img = _read_exr(out)
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-2)
    assert np.mean(rel < 2e-3) > 0.97  # HALF storage plus atomic summation order


def test_cli_learned_proposal(boxfog_obj, tmp_path):
    # -guide / -guide_media switch the learned proposal on; the image stays an unbiased
    # estimate of the same scene (compared by its mean with the plain CLI render)
    exe = os.path.join(ROOT, "gpuupbp_amd", "gpuupbp")
    base = ["-s", "-1", boxfog_obj] + box_args("48x48", extra=("-r_initial", "0.03", "-l", "6")) + ["-i", "24"]
    plain, guided = str(tmp_path / "plain.exr"), str(tmp_path / "guided.exr")
    r = subprocess.run([exe] + base + ["-o", plain], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe] + base + ["-guide", "0.65", "-guide_media", "0.5", "-o", guided], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _, a = _read_exr(plain)
    _, b = _read_exr(guided)
    assert np.isfinite(b).all() and abs(b.mean() - a.mean()) / a.mean() < 0.03, (a.mean(), b.mean())
