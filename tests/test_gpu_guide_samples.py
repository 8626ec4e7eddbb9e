"""Training samples of the learned proposal (GPU) against the oracle's restatement
(oracle_training_samples). The first trained iteration renders without the proposal, so the GPU's
camera paths are the oracle's path for path: every scattering that trains (in a medium, at a
guidable surface, as the proposal's weights switch each on) gives the same sample -- position,
scattered direction and weight, the luminance the path collects after the scattering over
lum(T f cos) / cp -- up to FP32 summation order and the few paths whose FP32 arithmetic diverges."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from conftest import box_args
from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu
from oracle import oracle_py

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alpha,surface_alpha", [(0.5, 0.65), (0.5, 0.0), (0.0, 0.65)])
def test_training_samples_match_the_oracle(boxfog_obj, alpha, surface_alpha):
    rc, cfg = capi.parse_args(["-s", "-1", boxfog_obj] + box_args("64x64", extra=("-r_initial", "0.03")))
    assert rc == 0
    sh = capi.SceneHandle(boxfog_obj, cfg.res_x, cfg.res_y)
    with gpu.Renderer(sh.ptr, cfg) as r:
        r.set_guide(alpha=alpha, surface_alpha=surface_alpha)
        r.run_iteration(0, want_stats=False)
        g, lo, edge = r.read_guide_samples()
    o = oracle_py.training_samples(sh.ptr, cfg, 0, alpha > 0.0, surface_alpha > 0.0)
    sh.close()
    print("samples: GPU %d, oracle %d; weight sums %.4f / %.4f" % (len(g), len(o), g[:, 9].sum(), o[:, 7].sum()))
    assert len(o) > 200
    assert abs(len(g) - len(o)) <= 0.02 * len(o) + 5
    assert np.all(g[:, 3:6] == 0.0) and np.all(g[:, 9] > 0.0) and np.isfinite(g).all()
    # the oracle's world positions in the record grid's unit cube
    ok = np.concatenate([(o[:, 1:4] - lo) / edge, o[:, 4:7]], axis=1)
    dist, idx = cKDTree(ok).query(np.concatenate([g[:, 0:3], g[:, 6:9]], axis=1))
    matched = dist < 1e-4
    print("matched %.4f" % matched.mean())
    assert matched.mean() > 0.95
    wg, wo = g[matched, 9], o[idx[matched], 7]
    rel = np.abs(wg - wo) / np.maximum(np.abs(wo), 1e-6)
    print("weights within 1e-3: %.4f" % np.mean(rel < 1e-3))
    assert np.mean(rel < 1e-3) > 0.98
    assert abs(g[:, 9].sum() - o[:, 7].sum()) <= 0.02 * o[:, 7].sum()
