"""Per-record MIS weights of the GPU renderer against the FP64 full-path MIS oracle (GPU).

The renderer's MIS weight log (upbp_set_mis_log) records, for sampled camera paths and the first
light paths of one iteration, the constant-time FP32 weight of every MIS-weighted contribution the
production kernels compute: light hits, next-event estimation, vertex connections, light tracing,
SURF / P-P3D / B-P2D point merges in the camera bounce kernel, and P-B2D / B-B1D merges in the
persistent segment merge kernels. The oracle renders the same iteration on the same Philox streams
and logs the same contributions with the primitive densities of their full paths, from which
oracle_mis.cpp forms the balance-heuristic weight p_T / sum p_T' over every technique that samples
the path, in FP64 (with the reference's two conventions that depart from it, see
test_mis_oracle.py).

Records are paired by (technique, k, s, camera path, light path). GPU and oracle trace the same
paths, up to FP32 transcendental differences that let a few percent of the paths take another
turn; a pair counts as the same contribution when its key group has the same size on both sides
and the GPU weight equals the oracle's own FP32 constant-time weight within 1e-3. Every technique
must pair nearly all of its oracle records that way (a missing or doubled merge pair in the
production kernels, or a wrong term in a GPU weight, breaks the pairing of its whole technique),
and the paired GPU weights must agree with the FP64 balance heuristic within 1e-4 for 99.9% of
the records and within 1e-2 for all.

Sum over techniques: the FP64 weights of one path's techniques are p_T / sum p_T' and sum to one
by construction, so weights within 1e-4 of them sum to one within 1e-4 per technique (the
reference's next-event convention excepted, which test_mis_oracle.py measures).
"""
import os
from collections import defaultdict

import numpy as np
import pytest

from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu
from oracle import oracle_py

pytestmark = pytest.mark.gpu

ALL = "upbp_bpt+surf+pp3d+pb2d+bb1d+bp2d"
STRIDE, N_LT = 3, 2000


def _guide_lobes(seed=5, res=3):
    rng = np.random.default_rng(seed)
    n = res ** 3 * 8
    mu = rng.normal(size=(n, 3))
    return np.hstack([mu / np.linalg.norm(mu, axis=1, keepdims=True), np.exp(rng.uniform(-2, 4, (n, 1))),
                      rng.uniform(0.05, 1, (n, 1))]).astype(np.float32)


def _records(scene_dir, kind, beams=(), guide_alpha=0.0, res=64):
    obj = capi.generate_scene(kind, 0, os.path.join(str(scene_dir), kind + "_mis.obj"))
    args = ["-s", "-1", obj, "-r", "%dx%d" % (res, res), "-l", "7", "-a", ALL, "-r_initial", "0.03",
            "-r_alpha", "1"] + list(beams)
    rc, cfg = capi.parse_args(args)
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    c, rad = np.array(list(sh.scene.scene_center)), sh.scene.scene_radius
    lobes = _guide_lobes()
    print("mis weights: %s %s guide %.1f" % (kind, " ".join(beams), guide_alpha), flush=True)
    try:
        with gpu.Renderer(sh.ptr, cfg) as r:
            if guide_alpha > 0:
                r.set_guide_lobes(lobes, 3, c - rad, c + rad, guide_alpha)
            r.set_mis_log(1 << 23, STRIDE, N_LT)
            st = r.run_iteration(0)
            g = r.read_mis_log(1 << 23)
        if guide_alpha > 0:
            oracle_py.set_guide(lobes, 3, c - rad, c + rad, guide_alpha)
        o = oracle_py.mis_check(sh.ptr, cfg, 0, np.arange(0, res * res, STRIDE), N_LT, max_records=4000000)
    finally:
        oracle_py.set_guide(None, 0, None, None, 0.0)
        sh.close()
    assert st.path_errors == 0
    return g, o


def _pair(g, o):
    """Pairs GPU and oracle records per key; returns, per technique, (oracle records, paired
    records, relative GPU-vs-FP64 errors of the paired ones)."""
    gk = defaultdict(list)
    for t, k, s, i, lp, w in zip(g["tech"], g["k"], g["s"], g["id"], g["lpath"], g["w"]):
        gk[(int(t), int(k), int(s), int(i), int(lp))].append(float(w))
    ok = defaultdict(list)
    for row in o:
        ok[(int(row[0]), int(row[2]), int(row[1]), int(row[6]), int(row[7]))].append((row[3], row[5]))
    out = defaultdict(lambda: [0, 0, []])
    for key, orecs in ok.items():
        t = key[0]
        out[t][0] += len(orecs)
        grecs = gk.get(key, [])
        if len(grecs) != len(orecs):
            continue
        orecs = sorted(orecs)
        grecs = sorted(grecs)
        if any(abs(gw - ow) > 1e-3 * max(ow, 1e-12) for gw, (ow, _) in zip(grecs, orecs)):
            continue
        out[t][1] += len(orecs)
        out[t][2].extend(abs(gw - ref) / max(ref, 1e-12) for gw, (_, ref) in zip(grecs, orecs))
    return out


@pytest.mark.parametrize("kind,beams,guide", [("box", (), 0.0), ("boxfog", (), 0.0),
                                              ("boxfog", ("-qbt", "S", "-pbt", "S"), 0.0),
                                              ("boxfog", ("-qbt", "S", "-pbt", "L"), 0.0), ("boxfog", (), 0.7)])
def test_gpu_weights_match_fp64_balance_heuristic(scene_dir, kind, beams, guide):
    g, o = _records(scene_dir, kind, beams, guide)
    assert len(g) > 1000 and len(o) > 1000
    assert np.all((g["w"] > 0) & (g["w"] <= 1.0 + 1e-6))
    res = _pair(g, o)
    names = oracle_py.MIS_TECH
    summary = {names[t]: (v[0], v[1]) for t, v in res.items()}
    print(summary)
    need = ["nee", "bpt", "light_tracing", "surf"] + (["pp3d", "pb2d", "bp2d", "bb1d"] if kind == "boxfog" else [])
    for n in need:
        t = [k for k, v in names.items() if v == n][0]
        assert res[t][0] > 50, (n, summary)
    pooled = []
    for t, (n_o, n_p, rel) in res.items():
        # box: every path pairs up to FP32 transcendental differences; the fog's distance and
        # scattering decisions diverge on a few percent of the paths
        assert n_p >= (0.97 if kind == "box" else 0.9) * n_o, (names[t], n_o, n_p)
        rel = np.asarray(rel)
        pooled.append(rel)
        if len(rel):
            # per technique: the 99.9th percentile within 1e-4 (a handful of records of a small
            # technique may sit just beyond it: paths whose geometry differs from the oracle's by
            # FP32 transcendental rounding, where the weight is sensitive to it)
            assert np.percentile(rel, 99.9) < 1e-4, (names[t], np.percentile(rel, 99.9))
            assert np.mean(rel > 1e-4) < 5e-3 and rel.max() < 1e-2, (names[t], rel.max())
    # over all records: 99.9% within 1e-4 of the FP64 balance heuristic, none beyond 1e-2
    rel = np.concatenate(pooled)
    print("records %d, beyond 1e-4: %d, max %.3g" % (len(rel), int(np.sum(rel > 1e-4)), rel.max()))
    assert np.mean(rel > 1e-4) < 1e-3 and rel.max() < 1e-2


def test_mis_log_switches_off(box_obj):
    rc, cfg = capi.parse_args(["-s", "-1", box_obj, "-r", "32x32", "-l", "6", "-a", ALL])
    sh = capi.SceneHandle(box_obj, cfg.res_x, cfg.res_y)
    with gpu.Renderer(sh.ptr, cfg) as r:
        r.set_mis_log(1 << 20, 1, 100)
        r.run_iteration(0)
        n0 = len(r.read_mis_log())
        r.set_mis_log(0)
        r.run_iteration(1)
        with pytest.raises(RuntimeError):
            r.read_mis_log()
        r.set_mis_log(16, 1, 100)  # far too small: the overflow is reported
        r.run_iteration(2)
        with pytest.raises(RuntimeError):
            r.read_mis_log()
        img = r.read_accum(3)
    sh.close()
    assert n0 > 100 and np.isfinite(img).all()
