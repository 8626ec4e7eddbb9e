"""Batched jobs (GPU): one iteration of each of several independent renderers of uneven size and
path length -- driven side by side from host threads (upbp_run_batch), or as rounds of one launch
per stage over all of them, eager or replayed from a captured HIP graph (upbp_batch_*) -- gives
each renderer the image it renders on its own."""
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


def _same(img, ref):
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-3)
    return np.isfinite(img).all() and np.mean(rel < 1e-4) > 0.99 and abs(img.mean() - ref.mean()) <= 1e-5 * abs(
        ref.mean()) + 1e-12


@pytest.mark.parametrize("use_graph", [1, 0])
def test_batch_in_one_launch_per_stage_matches_each_job_alone(scene_dir, use_graph):
    # Rounds: the first sizes every job's photon lists (jobs alone), the second runs batched and
    # eagerly, later ones replay the captured round (use_graph) or run batched eagerly again.
    # Every job's image after five rounds equals its own five iterations.
    n_it = 5
    alone = []
    for sh, r in _renderers(scene_dir):
        for i in range(n_it):
            r.run_iteration(i, want_stats=False)
        alone.append(r.read_accum(n_it))
        r.close()
        sh.close()
    jobs = _renderers(scene_dir)
    rs = [r for _, r in jobs]
    with gpu.Batch(rs, use_graph=use_graph) as bt:
        for i in range(n_it):
            bt.run([i] * len(rs))
            # every round's images against the jobs' own, with the program's own work between
            if i >= 2:
                for r in rs:
                    assert np.isfinite(r.read_accum(i + 1)).all()
        counts = bt.counts()
    print("rounds, graph replays, eager batched, alone:", counts)
    if use_graph:
        assert counts == (n_it, n_it - 2, 1, 1), counts
    else:
        assert counts == (n_it, 0, n_it - 1, 1), counts
    for (sh, r), ref, job in zip(jobs, alone, JOBS):
        img = r.read_accum(n_it)
        r.close()
        sh.close()
        assert _same(img, ref), job


def test_batch_rejects_bad_arguments(scene_dir):
    jobs = _renderers(scene_dir)[:2]
    rs = [r for _, r in jobs]
    with pytest.raises(RuntimeError):
        gpu.Batch([rs[0], rs[0]])
    with pytest.raises(RuntimeError):
        gpu.Batch([])
    for sh, r in jobs:
        r.close()
        sh.close()


def test_batch_after_a_photon_list_overflow(scene_dir):
    # B-B1D radius shrinking per iteration without B-P2D: the sub-beam list outgrows its capacity in
    # batched rounds; that job's iteration is then rendered again alone, its lists resized, and the
    # next round runs batched eagerly (the captured graph is out of date) before it is captured
    # again. The image stays that of the job's own iterations.
    obj = str(scene_dir / "boxfog" / "boxfog.obj")
    args = box_args("48x48", alg="upbp_bpt+surf+pp3d+pb2d+bb1d",
                    extra=("-r_initial", "0.03", "-r_alpha_bb1d", "0.5"))
    rc, cfg = capi.parse_args(["-s", "-1", obj] + args)
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    n_it = 8
    ref = gpu.Renderer(sh.ptr, cfg)
    for i in range(n_it):
        ref.run_iteration(i, want_stats=False)
    want = ref.read_accum(n_it)
    ref.close()
    jobs = _renderers(scene_dir)[:3]
    r = gpu.Renderer(sh.ptr, cfg)
    rs = [r] + [x for _, x in jobs]
    with gpu.Batch(rs) as bt:
        for i in range(n_it):
            bt.run([i] * len(rs))
        counts = bt.counts()
    print("rounds, graph replays, eager batched, alone:", counts)
    assert counts[0] == n_it and counts[3] == 1 and counts[2] >= 2, counts
    assert _same(r.read_accum(n_it), want)
    r.close()
    for s2, x in jobs:
        x.close()
        s2.close()
    sh.close()
