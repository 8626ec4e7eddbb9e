"""GPU parity tests: the HIP path (through the C ABI of libupbp.so) against the CPU oracle.

Range queries and closest hits must match brute force bit for bit (index sets and FP32 hit
distances). Renders must match the oracle path for path: both use the same counter-based
Philox streams, so one iteration of the box workload agrees pixel by pixel up to FP32
transcendental-function differences (tolerances below), and converged images agree in relMSE.
"""
import numpy as np
import pytest

from conftest import assert_whole_image, box_args, oracle_independent
from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu
from oracle import oracle_py

pytestmark = pytest.mark.gpu


def _sets(offsets, idx):
    return [np.sort(idx[offsets[i]:offsets[i + 1]]) for i in range(len(offsets) - 1)]


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _cases(kind, rng, n_prims, n_q, spread=1.0):
    if kind in (0, 1):
        prims = rng.uniform(-spread, spread, (n_prims, 3)).astype(np.float32)
        prims[: n_prims // 10] = prims[0]  # a cluster of identical points (Morton ties)
    else:
        o = rng.uniform(-spread, spread, (n_prims, 3))
        d = _unit(rng.normal(size=(n_prims, 3)))
        ln = rng.uniform(0.05, 1.5, (n_prims, 1))
        prims = np.hstack([o, d, ln]).astype(np.float32)
    if kind in (0, 2):
        q = rng.uniform(-spread, spread, (n_q, 3)).astype(np.float32)
        if kind == 0:
            k = min(20, n_prims)
            q[:k] = prims[:k]  # queries on photons
    else:
        o = rng.uniform(-spread, spread, (n_q, 3))
        d = _unit(rng.normal(size=(n_q, 3)))
        t0 = rng.uniform(0.0, 0.3, (n_q, 1))
        t1 = t0 + rng.uniform(0.1, 2.0, (n_q, 1))
        q = np.hstack([o, d, t0, t1]).astype(np.float32)
    return prims, q


@pytest.mark.parametrize("kind,radius", [(0, 0.05), (1, 0.02), (2, 0.03), (3, 0.02)])
def test_range_query_matches_brute_force(kind, radius):
    rng = np.random.default_rng(100 + kind)
    n_prims = 20000 if kind in (0, 1) else 6000
    prims, q = _cases(kind, rng, n_prims, 1500)
    go, gi = gpu.range_query(kind, prims, q, radius)
    oo, oi = oracle_py.range_query(kind, prims, q, radius)
    assert len(gi) == len(oi) and len(gi) > 100
    for a, b in zip(_sets(go, gi), _sets(oo, oi)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kind,radius", [(1, 0.02), (3, 0.02)])
def test_range_query_wide_tree_matches_brute_force(kind, radius):
    # the 4-wide quantised collapse walked by the segment merge kernels loses no leaf
    rng = np.random.default_rng(300 + kind)
    prims, q = _cases(kind, rng, 20000 if kind == 1 else 6000, 1500, spread=3.0)
    go, gi = gpu.range_query(kind, prims, q, radius, subbeam_len=0.05 if kind == 3 else 0.0, wide=True)
    oo, oi = oracle_py.range_query(kind, prims, q, radius)
    assert len(gi) == len(oi) and len(gi) > 100
    for a, b in zip(_sets(go, gi), _sets(oo, oi)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kind,sub", [(2, 0.0), (2, 0.07), (3, 0.0), (3, 0.03), (3, 0.31)])
def test_range_query_subbeams_report_each_beam_once(kind, sub):
    # beams cut into sub-beams (as the renderer's beam map does) give the brute-force index set
    rng = np.random.default_rng(200 + kind)
    prims, q = _cases(kind, rng, 5000, 1200)
    go, gi = gpu.range_query(kind, prims, q, 0.02, subbeam_len=sub)
    oo, oi = oracle_py.range_query(kind, prims, q, 0.02)
    assert len(gi) == len(oi) and len(gi) > 50
    for a, b in zip(_sets(go, gi), _sets(oo, oi)):
        assert np.array_equal(a, b)


def test_range_query_near_parallel_beams():
    # beams within a few degrees of the query direction, around the parallel cut-off
    # |cos^2 - 1| < 1e-5 of test_intersection_beam_beam (intersection.cuh:153), far from the origin
    rng = np.random.default_rng(11)
    n = 4000
    qd = _unit(rng.normal(size=(300, 3)))
    qo = rng.uniform(2.0, 6.0, (300, 3))
    q = np.hstack([qo, qd, np.zeros((300, 1)), np.full((300, 1), 4.0)]).astype(np.float32)
    k = rng.integers(0, 300, n)
    ang = 10.0 ** rng.uniform(-3.3, -1.5, n)
    perp = _unit(np.cross(qd[k], rng.normal(size=(n, 3))))
    d = _unit(qd[k] + perp * ang[:, None])
    t = rng.uniform(0.0, 4.0, (n, 1))
    off = _unit(rng.normal(size=(n, 3))) * rng.uniform(0.0, 0.03, (n, 1))
    o = qo[k] + qd[k] * t + off - d * rng.uniform(0.1, 1.0, (n, 1))
    prims = np.hstack([o, d, rng.uniform(0.2, 2.0, (n, 1))]).astype(np.float32)
    for sub in (0.0, 0.05):
        go, gi = gpu.range_query(3, prims, q, 0.02, subbeam_len=sub)
        oo, oi = oracle_py.range_query(3, prims, q, 0.02)
        assert len(gi) == len(oi) and len(gi) > 100
        for a, b in zip(_sets(go, gi), _sets(oo, oi)):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_range_query_edge_sizes(kind):
    rng = np.random.default_rng(7)
    for n_prims in (0, 1, 2, 3, 5):
        prims, q = _cases(kind, rng, max(n_prims, 1), 64, spread=0.05)
        prims = prims[:n_prims]
        oo, oi = oracle_py.range_query(kind, prims, q, 0.2)
        for wide in (False, True):
            go, gi = gpu.range_query(kind, prims, q, 0.2, wide=wide)
            assert np.array_equal(go, oo)
            for a, b in zip(_sets(go, gi), _sets(oo, oi)):
                assert np.array_equal(a, b)


def test_range_query_closed_boundary():
    # |q - p|^2 == r^2 exactly must be accepted (point_point uses <=, intersection.cuh:71-75)
    prims = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.25, 0.0]], np.float32)
    q = np.array([[0.25, 0.0, 0.0], [0.0, 0.0, 0.0]], np.float32)
    go, gi = gpu.range_query(0, prims, q, 0.25)
    sets = _sets(go, gi)
    assert list(sets[0]) == [0, 1] and list(sets[1]) == [0, 2]


def test_closest_hit_matches_brute_force():
    rng = np.random.default_rng(5)
    n = 4000
    c = rng.uniform(-2, 2, (n, 1, 3))
    tris = (c + rng.normal(scale=0.15, size=(n, 3, 3))).reshape(n, 9).astype(np.float32)
    o = rng.uniform(-3, 3, (3000, 3))
    d = _unit(rng.normal(size=(3000, 3)))
    rays = np.hstack([o, d, np.full((3000, 1), 1e-3), np.full((3000, 1), 1e30)]).astype(np.float32)
    gt, gp = gpu.closest_hit(tris, rays)
    ot, op = oracle_py.closest_hit(tris, rays)
    assert np.array_equal(np.isinf(gt), np.isinf(ot))
    hit = ~np.isinf(ot)
    assert hit.sum() > 300
    assert np.array_equal(gt[hit], ot[hit])
    assert np.mean(gp[hit] == op[hit]) > 0.999


def _render_pair(obj, args, iters=1, first=0, indep=False):
    rc, cfg = capi.parse_args(["-s", "-1", obj] + args)
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    with gpu.Renderer(sh.ptr, cfg) as r:
        stats = [r.run_iteration(first + i) for i in range(iters)]
        g = r.read_accum(iters)
    o, st = oracle_py.render(sh.ptr, cfg, first, iters)
    if indep:
        return g, o, stats, st, oracle_independent(sh.ptr, cfg, first, iters)
    return g, o, stats, st


def test_render_path_parity_one_iteration(box_obj):
    g, o, stats, st, o2 = _render_pair(box_obj, box_args("64x64"), iters=1, indep=True)
    assert np.isfinite(g).all()
    assert stats[0].path_errors == 0 and stats[0].beam_drops == 0
    assert stats[0].surface_vertices + stats[0].medium_vertices == st.light_vertices
    assert stats[0].beams == st.beams
    rel = np.abs(g - o) / np.maximum(np.abs(o), 1e-3)
    # same random streams: the large majority of pixels agree to FP32 rounding
    assert np.mean(rel < 1e-3) > 0.97, np.mean(rel < 1e-3)
    assert abs(g.mean() - o.mean()) / o.mean() < 0.01
    assert_whole_image(g, o, o2)


@pytest.mark.parametrize("beams", [(), ("-qbt", "S", "-pbt", "S")])
def test_render_path_parity_dense_medium(boxfog_obj, beams):
    # denser anisotropic fog and a glossy wall: every merge estimator contributes; the GPU's
    # MIS weights must follow the oracle's, which the FP64 MIS oracle pins
    g, o, stats, _, o2 = _render_pair(boxfog_obj, box_args("48x48", extra=("-r_initial", "0.03") + beams), iters=1,
                                      indep=True)
    assert stats[0].path_errors == 0 and stats[0].beam_drops == 0
    assert stats[0].pb2d_accepts > 0 and stats[0].bb1d_accepts > 0
    rel = np.abs(g - o) / np.maximum(np.abs(o), 1e-3)
    assert np.mean(rel < 1e-3) > 0.95, np.mean(rel < 1e-3)
    assert abs(g.mean() - o.mean()) / o.mean() < 0.01
    assert_whole_image(g, o, o2)


def test_render_converged_relmse(box_obj):
    g, o, stats, _ = _render_pair(box_obj, box_args("48x48"), iters=8)
    relmse = np.mean((g - o) ** 2 / (o ** 2 + 1e-2))
    assert relmse < 1e-3, relmse
    assert sum(s.path_errors for s in stats) == 0


@pytest.mark.parametrize("alg", ["upbp_bpt", "upbp_surf+pp3d", "upbp_pb2d", "upbp_bb1d", "upbp_bp2d+bpt"])
def test_technique_subsets_match_oracle(box_obj, alg):
    g, o, _, _ = _render_pair(box_obj, box_args("32x32", alg=alg), iters=2)
    relmse = np.mean((g - o) ** 2 / (o ** 2 + 1e-2))
    assert relmse < 2e-3, (alg, relmse)


def test_iteration_offset_and_radius_schedule(box_obj):
    # -r_alpha 0.75 shrinks radii per iteration; iteration 3 must match the oracle's iteration 3
    args = ["-r", "32x32", "-l", "6", "-a", "upbp_bpt+surf+pp3d+pb2d+bb1d", "-r_initial", "0.02", "-r_alpha", "0.75",
            "-ioff", "5"]
    g, o, _, _ = _render_pair(box_obj, args, iters=1, first=3)
    relmse = np.mean((g - o) ** 2 / (o ** 2 + 1e-2))
    assert relmse < 1e-3, relmse
