"""Benchmark: UPBP renderer iteration on one MI355X.

Workload (the plain-UPBP-at-scale configuration): seeded grid of 64 noise-displaced
icosphere blobs (1,310,720 triangles) on a floor, three overlapping homogeneous media
spheres with priorities 1-3, two quad lights and a point light, 1280x720, max path
length 12, 900,000 light paths per iteration, radius 0.005, upbp_all
(BPT + surface PM + P-P3D + P-B2D + B-B1D), FP32 transport, FP64 accumulation.

One step = one 1-spp iteration (light pass, photon LBVH builds, camera pass, FP64
accumulate) with the scene resident in HBM. The default run then measures the learned
proposal against the plain renderer at equal wall-clock time (60 s each) on the slab room,
against the committed plain reference render (gpuupbp_amd/data/), and the 64 batched jobs.
Prints one JSON line; it also carries the committed stress measurement (profiles/stress_4k.json).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-learned] [--no-batch]
                       [--autotune]

Other measurements (each prints its own JSON line):
  --learned SECONDS                    only the learned-vs-plain comparison (guide settings:
                                       --alpha, --surface-alpha, --guide-res, --guide-steps,
                                       --guide-lr, --guide-history; comma lists compare several)
  --batch JOBS                         batched independent jobs: one after another, from host
                                       threads (upbp_run_batch), one launch per stage (upbp_batch_*)
  --stress                             3840x2160, 2.05M triangles, 8 media, 8M light paths
  --reference-chunk FIRST:LAST         render part of the slab-room reference (GPU), then
  --reference-merge CHUNK...           join the parts into gpuupbp_amd/data/ (no GPU)
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "samples/s per 1-spp iter + equal-time relMSE, learned vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0
RES = (1280, 720)
ARGS = ["-r", "%dx%d" % RES, "-l", "12", "-pcpi", "900000", "-r_initial", "0.005", "-a", "upbp_all"]


def make_scene(workdir):
    from gpuupbp_amd import capi
    obj = os.path.join(workdir, "blobs.obj")
    if not os.path.exists(obj + ".aux"):
        capi.generate_scene("blobs", 1, obj)
    return obj


def cpu_baseline(scene_ptr, cfg):
    """Oracle (scalar FP32 port of the reference estimators) on the host cores: the full light pass
    and photon-map build, plus an evenly spread sample of camera paths whose time is scaled to
    the full frame."""
    from oracle import oracle_py
    threads = min(16, os.cpu_count() or 1)
    n_cam = 65536
    t = oracle_py.render_sample_times(scene_ptr, cfg, n_cam, int(cfg.path_count_per_iter), threads)
    npx = cfg.res_x * cfg.res_y
    t_iter = t[0] + t[1] + t[2] * (npx / float(n_cam))
    return {
        "value": npx / t_iter,
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": "one iteration: all %d light paths + photon maps (%.1f s + %.1f s), camera pass on %d of %d "
                  "pixels (%.1f s) scaled to the frame" % (int(cfg.path_count_per_iter), t[0], t[1], n_cam, npx, t[2]),
    }


def measured_traffic(kernel):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        t = json.load(open(path))
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel or t.get("workload") != "blobs":
        return None
    return t["traffic_bytes_per_launch"]


def batch_jobs(workdir, n_jobs=64, iters=4, threads=8):
    """Batched jobs: the n_jobs independent renderers of the seed-2 mix (gpuupbp_amd/workloads.py:
    box, fog box, blob grids and slab rooms of 10 to 195k triangles and 1 to 4 media, resolution
    in {64, 128, 256, 512}^2, max length in [4, 16]). Each renders `iters` iterations per mode:
    one job after another, all at once from host threads (upbp_run_batch), and in rounds of one
    launch per stage over all jobs (upbp_batch_run), eagerly and replayed from the captured HIP
    graph. Returns the measurement."""
    from gpuupbp_amd import capi
    from gpuupbp_amd import workloads as wl
    from gpuupbp_amd.renderer import Batch, Renderer, run_batch
    jobs = []
    for obj, argv, px, kind, length in wl.batch_mix(workdir, n_jobs, 2):
        rc, cfg = capi.parse_args(argv)
        assert rc == 0
        sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
        jobs.append((sh, Renderer(sh.ptr, cfg), px, kind, length))
    rs = [r for _, r, _, _, _ in jobs]
    pixels = sum(px for _, _, px, _, _ in jobs)
    run_batch(rs, [0] * len(rs), threads)  # warm-up: buffers grow to their sizes
    for r in rs:
        r.sync()
    it = 1
    t0 = time.perf_counter()
    for _ in range(iters):
        for r in rs:
            r.run_iteration(it, want_stats=False)
            r.sync()
        it += 1
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(iters):
        run_batch(rs, [it] * len(rs), threads)
        it += 1
    t_thr = time.perf_counter() - t0
    res = {}
    for label, graph in (("eager", 0), ("graph", 1)):
        with Batch(rs, use_graph=graph) as bt:
            for _ in range(2):  # warm-up rounds: an eager one, then the capture
                bt.run([it] * len(rs))
                it += 1
            t0 = time.perf_counter()
            for _ in range(iters):
                bt.run([it] * len(rs))
                it += 1
            res[label] = (time.perf_counter() - t0, bt.counts())
    for sh, r, _, _, _ in jobs:
        r.close()
        sh.close()
    mix = {k: sum(1 for _, _, _, kk, _ in jobs if kk == k) for k in wl.BATCH_KINDS}
    t_eager, c_eager = res["eager"]
    t_graph, c_graph = res["graph"]
    return {"jobs": n_jobs, "iterations_per_job": iters, "threads": threads, "job_mix": mix,
            "pixels_per_round": pixels, "seconds_sequential": round(t_seq, 3), "seconds_threads": round(t_thr, 3),
            "seconds_stage_launches": round(t_eager, 3), "seconds_stage_launches_graph": round(t_graph, 3),
            "samples_per_s_sequential": pixels * iters / t_seq, "samples_per_s_threads": pixels * iters / t_thr,
            "samples_per_s_stage_launches": pixels * iters / t_eager,
            "samples_per_s_stage_launches_graph": pixels * iters / t_graph,
            "speedup_threads": t_seq / t_thr, "speedup_stage_launches_graph": t_seq / t_graph,
            "rounds_graph_replays_eager_alone": {"eager": list(c_eager), "graph": list(c_graph)}}


def stress(workdir, steps=3, warmup=1):
    """Combined stress workload: 3840x2160, 2.05M-triangle blob grid with eight overlapping media
    (g from -0.3 to 0.95), l = 16, 8M light paths per iteration, all estimators; the plain
    renderer, then the learned proposal (library defaults, trained online) on the same scene."""
    import ctypes as C
    from gpuupbp_amd import capi
    from gpuupbp_amd.renderer import Renderer
    obj = os.path.join(workdir, "stress.obj")
    if not os.path.exists(obj + ".aux"):
        capi.generate_scene("stress", 1, obj)
    rc, cfg = capi.parse_args(["-s", "-1", obj, "-r", "3840x2160", "-l", "16", "-pcpi", "8000000",
                               "-r_initial", "0.005", "-a", "upbp_all"])
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    out = {"workload": "stress: 10x10 noise-displaced icospheres (%d tris), 8 media spheres (g -0.3..0.95), "
                       "3840x2160, l=16, 8M light paths/iter, r=0.005, upbp_all" % sh.scene.real.count}
    for label, guide in (("plain", None), ("learned", {})):
        r = Renderer(sh.ptr, cfg)
        if guide is not None:
            r.set_guide(**guide)
        for i in range(warmup):
            r.run_iteration(i, want_stats=False)
        r.sync()
        t0 = time.perf_counter()
        for i in range(steps):
            r.run_iteration(warmup + i, want_stats=False)
        r.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        st = r.run_iteration(warmup + steps, want_stats=True)
        free, total = C.c_uint64(), C.c_uint64()
        capi.lib().upbp_memory_info(C.byref(free), C.byref(total))
        r.close()
        out[label] = {"ms_per_iteration": round(ms, 1), "samples_per_s": 3840 * 2160 / (ms * 1e-3),
                      "surface_points": st.surface_vertices, "medium_points": st.medium_vertices,
                      "beams": st.beams, "subbeams": st.subbeams, "bb1d_accepts": st.bb1d_accepts,
                      "path_errors": st.path_errors, "vertex_drops": st.vertex_drops, "beam_drops": st.beam_drops,
                      "device_memory_used_gib": round((total.value - free.value) / 2 ** 30, 1)}
        print("bench: stress %s %.1f ms per iteration" % (label, ms), file=sys.stderr, flush=True)
    sh.close()
    return out


def committed(name):
    """A committed measurement under profiles/ (JSON written from a bench.py --batch / --stress
    run), or None."""
    try:
        return json.load(open(os.path.join(ROOT, "profiles", name)))
    except (OSError, ValueError):
        return None


REFERENCE = os.path.join(ROOT, "gpuupbp_amd", "data", "slabroom_1280x720_reference.npz")
# the reference render's random streams: shifted by -ioff so that they share nothing with the
# equal-time renders (iterations 0.. on unshifted streams), while every render follows the same
# radius schedule from iteration 0
REFERENCE_IOFF = 1 << 24


def slabroom(workdir, res=(1280, 720), ioff=0):
    from gpuupbp_amd import capi
    from gpuupbp_amd import workloads as wl
    obj = wl.scene(workdir, "slabroom", 3)
    args = list(wl.SLABROOM_ARGS)
    args[args.index("-r") + 1] = "%dx%d" % res
    rc, cfg = capi.parse_args(["-s", "-1", obj] + args + ["-ioff", str(ioff)])
    assert rc == 0
    return obj, cfg


def reference_chunk(workdir, first, last, seconds, out_dir):
    """Renders iterations [first, last) of the plain slab-room reference (or as many as fit in
    `seconds`) and writes their FP64 sum with the range it covers to out_dir."""
    import numpy as np
    from gpuupbp_amd import capi
    from gpuupbp_amd.renderer import Renderer
    obj, cfg = slabroom(workdir, ioff=REFERENCE_IOFF)
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    r = Renderer(sh.ptr, cfg)
    t0 = last_print = time.perf_counter()
    it = first
    while it < last:
        r.run_iteration(it, want_stats=False)
        it += 1
        if it % 16 == 0:
            r.sync()
            now = time.perf_counter()
            if now - last_print > 20:
                print("bench: reference iterations %d..%d, %.0f s" % (first, it, now - t0), file=sys.stderr, flush=True)
                last_print = now
            if now - t0 > seconds:
                break
    img = r.read_accum(1)  # the sum over the iterations rendered
    r.close()
    sh.close()
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, "ref_%06d_%06d.npy" % (first, it))
    np.save(path, img)
    return {"first": first, "end": it, "seconds": round(time.perf_counter() - t0, 1), "file": path}


def reference_merge(paths):
    """Joins reference chunks (contiguous iteration ranges from 0) into the committed reference
    image: the FP64 mean, stored as FP16 (relative precision 5e-4, far below the error measured)."""
    import numpy as np
    chunks = sorted((tuple(int(v) for v in os.path.basename(p)[4:-4].split("_")), p) for p in paths)
    at, total = 0, None
    for (a, b), p in chunks:
        if a != at:
            raise SystemExit("reference chunks are not contiguous at iteration %d" % at)
        img = np.load(p)
        total = img if total is None else total + img
        at = b
    os.makedirs(os.path.dirname(REFERENCE), exist_ok=True)
    np.savez_compressed(REFERENCE, image=(total / at).astype(np.float16), iterations=np.int64(at),
                        ioff=np.int64(REFERENCE_IOFF))
    return {"iterations": at, "file": REFERENCE}


def learned_vs_plain(workdir, seconds=60.0, guides=({},)):
    """The learned proposal against the plain renderer at equal wall-clock time on the slab room
    (194,970 triangles, dense g = 0.9 slab lit through a 5% aperture, 1280x720, l = 16): relMSE of
    each against the committed plain reference, rendered from iteration 0 on other random streams
    for at least 64 times as many iterations as the plain renderer manages here. The learned run's
    time includes its online training. guides: learned-proposal settings to compare (dicts of
    upbp_guide_config fields), each against the same reference and plain run. None if the
    reference is missing."""
    import numpy as np
    from gpuupbp_amd import capi
    from gpuupbp_amd.renderer import Renderer
    if not os.path.exists(REFERENCE):
        return None
    ref_npz = np.load(REFERENCE)
    ref = ref_npz["image"].astype(np.float64)
    ref_iters = int(ref_npz["iterations"])
    obj, cfg = slabroom(workdir)
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)

    def run(g, label):
        r = Renderer(sh.ptr, cfg)
        if g is not None:
            r.set_guide(**g)
        r.sync()
        t0 = last = time.perf_counter()
        n = 0
        while True:
            r.run_iteration(n, want_stats=False)
            r.sync()
            n += 1
            now = time.perf_counter()
            if now - last > 20:
                print("bench: %s %d iterations, %.0f s" % (label, n, now - t0), file=sys.stderr, flush=True)
                last = now
            if now - t0 >= seconds:
                break
        img = r.read_accum(n)
        r.close()
        return img, n, time.perf_counter() - t0

    rel = lambda im: float(np.mean((im - ref) ** 2 / (ref ** 2 + 1e-2)))
    plain, n_plain, t_plain = run(None, "plain")
    rp = rel(plain)
    learned = []
    for g in guides:
        img, n, t = run(g, "learned %s" % g)
        rl = rel(img)
        learned.append({"guide": g, "iterations": n, "wall_s": round(t, 2), "relmse": rl,
                        "ratio_plain_over_learned": rp / rl})
    sh.close()
    first = learned[0]
    return {"workload": "slabroom: 194,970 tris, slab sigma_t 7.23 (U[2,10], seed 3), albedo 0.95, g 0.9, "
                        "5% aperture, 1280x720, l=16, r=0.01, upbp_all",
            "seconds": seconds, "reference_iterations": ref_iters,
            "reference_spp_over_plain": ref_iters / n_plain,
            "iterations_plain": n_plain, "iterations_learned": first["iterations"],
            "wall_plain_s": round(t_plain, 2), "wall_learned_s": first["wall_s"],
            "relmse_plain": rp, "relmse_learned": first["relmse"],
            "ratio_plain_over_learned": first["ratio_plain_over_learned"], "target": 1.5,
            "guide": first["guide"], "learned": learned,
            "measured": "in this run (raw relMSE, nothing subtracted)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--subbeam", type=float, default=None, help="sub-beam length in beam radii")
    ap.add_argument("--piece", type=float, default=None, help="query piece length in merge radii")
    ap.add_argument("--autotune", action="store_true",
                    help="run the library's autotuner (upbp_autotune) before the warm-up and use its choice")
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="size of the plain copy kernel run once for bandwidth and counter calibration (0: skip)")
    ap.add_argument("--stress", action="store_true", help="combined stress workload (3840x2160, 2M tris, 8 media)")
    ap.add_argument("--batch", type=int, default=None, metavar="JOBS",
                    help="batched-jobs measurement over JOBS independent renderers")
    ap.add_argument("--batch-threads", type=int, default=8)
    ap.add_argument("--no-batch", dest="batch_in_default", action="store_false",
                    help="skip the batched-jobs measurement of the default run")
    ap.add_argument("--learned", type=float, default=None,
                    help="only the learned-vs-plain equal-time comparison (slab room), SECONDS per renderer; "
                         "guide settings below (comma lists compare several against one plain run)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-vs-plain comparison of the default run")
    ap.add_argument("--learned-seconds", type=float, default=60.0,
                    help="seconds per renderer of the default run's learned-vs-plain comparison")
    ap.add_argument("--reference-chunk", default=None, metavar="FIRST:LAST",
                    help="render iterations [FIRST, LAST) of the plain slab-room reference (at most "
                         "--reference-seconds) into gpurun_out/reference/")
    ap.add_argument("--reference-seconds", type=float, default=900.0)
    ap.add_argument("--reference-merge", nargs="+", default=None, metavar="CHUNK",
                    help="join reference chunks into the committed reference image (no GPU)")
    ap.add_argument("--alpha", default=None,
                    help="learned-proposal mixture weight in media (library default); a comma list compares several")
    ap.add_argument("--surface-alpha", default=None,
                    help="mixture weight at surfaces (library default); a comma list pairs with --alpha")
    ap.add_argument("--guide-res", default=None, help="record grid cells per axis (comma list allowed)")
    ap.add_argument("--guide-steps", default=None, help="Adam steps per iteration (comma list allowed)")
    ap.add_argument("--guide-lr", default=None, help="Adam step size (comma list allowed)")
    ap.add_argument("--guide-history", default=None, help="training-sample ring size (comma list allowed)")
    a = ap.parse_args()
    if a.reference_merge:
        print(json.dumps(reference_merge(a.reference_merge)))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: the iteration-sharded multi-GPU path is not built", file=sys.stderr)
        sys.exit(2)

    from gpuupbp_amd import capi
    from gpuupbp_amd.renderer import Renderer

    workdir = a.workdir or tempfile.mkdtemp(prefix="upbp_bench_")
    os.makedirs(workdir, exist_ok=True)
    if a.stress:
        print(json.dumps({"metric": "combined stress, samples/s, 1 MI355X", **stress(workdir)}))
        return
    if a.batch is not None:
        res = batch_jobs(workdir, n_jobs=a.batch, threads=a.batch_threads)
        print(json.dumps({"metric": "batched jobs, samples/s, 1 MI355X", **res}))
        return
    if a.reference_chunk:
        first, last = (int(v) for v in a.reference_chunk.split(":"))
        print(json.dumps(reference_chunk(workdir, first, last, a.reference_seconds,
                                         os.path.join(ROOT, "gpurun_out", "reference"))))
        return
    if a.learned is not None:
        lists = {k: ([conv(v) for v in str(val).split(",")] if val is not None else [None])
                 for k, val, conv in (("alpha", a.alpha, float), ("surface_alpha", a.surface_alpha, float),
                                      ("grid_res", a.guide_res, int), ("steps_per_iteration", a.guide_steps, int),
                                      ("learning_rate", a.guide_lr, float), ("history", a.guide_history, int))}
        n = max(len(v) for v in lists.values())
        guides = [{k: v[i % len(v)] for k, v in lists.items() if v[i % len(v)] is not None} for i in range(n)]
        res = learned_vs_plain(workdir, a.learned, guides)
        if res is None:
            raise SystemExit("the slab-room reference image is missing: " + REFERENCE)
        print(json.dumps({"metric": "equal-time relMSE, learned vs plain, 1 MI355X", **res}))
        return
    obj = make_scene(workdir)
    rc, cfg = capi.parse_args(["-s", "-1", obj] + ARGS)
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)

    copy = None
    if a.copy_mib > 0:
        import ctypes as C
        cms = C.c_double()
        nbytes = a.copy_mib << 20
        if capi.lib().upbp_test_copy(nbytes, 3, C.byref(cms)) == 0:
            copy = {"bytes": nbytes, "ms": cms.value, "GB_per_s": 2.0 * nbytes / (cms.value * 1e-3) / 1e9}
    r = Renderer(sh.ptr, cfg)
    tune = r.set_tuning(subbeam_factor=a.subbeam, query_piece_factor=a.piece)
    tuning = {f: getattr(tune, f) for f, _ in type(tune)._fields_}
    autotune = None
    if a.autotune:
        ta = time.perf_counter()
        tuning, tms = r.autotune(iteration=0, repeats=1)
        autotune = {"seconds": round(time.perf_counter() - ta, 2), "ms_per_iteration": round(tms, 2)}
        print("bench: autotuned %s (%.1f ms per iteration)" % (tuning, tms), file=sys.stderr, flush=True)
    print("bench: scene loaded, %d warm-up + %d timed iterations" % (a.warmup, a.steps), file=sys.stderr, flush=True)
    for i in range(a.warmup):
        r.run_iteration(i, want_stats=False)
    r.sync()
    t0 = time.perf_counter()
    for i in range(a.steps):
        r.run_iteration(a.warmup + i, want_stats=False)
    r.sync()
    t1 = time.perf_counter()
    # one more iteration with device counters and per-kernel HIP events (untimed)
    st = r.run_iteration(a.warmup + a.steps, want_stats=True)
    r.close()

    ms = (t1 - t0) * 1e3 / a.steps
    npx = cfg.res_x * cfg.res_y
    launches = max(1, st.dominant_launches)
    dom_ms = st.ms_dominant / launches
    achieved = (st.dominant_bytes / launches) / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": npx / (ms * 1e-3),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {
            "workload": "blobs: 64 noise-displaced icospheres (1,310,720 tris), 3 media spheres, 2 quad + 1 point "
                        "light, 1280x720, l=12, 900k light paths/iter, r=0.005, upbp_all",
            "resolution": "%dx%d" % RES,
            "max_path_length": 12,
            "light_paths": int(cfg.path_count_per_iter),
            "parallelism": "single GPU",
            "tuning": tuning,
            "autotune": autotune,
        },
        "paths_per_s": (npx + cfg.path_count_per_iter) / (ms * 1e-3),
        "queries_per_s": (st.point_queries + st.segment_queries) / (ms * 1e-3),
        "iteration_breakdown_ms": {"light": st.ms_light, "build": st.ms_build, "camera": st.ms_camera,
                                   "camera_bounce": st.ms_bounce, "merge_pb2d": st.ms_merge_pb2d,
                                   "merge_bb1d": st.ms_merge_bb1d},
        "counters": {"light_steps": st.light_extend_steps, "camera_steps": st.camera_extend_steps,
                     "surface_points": st.surface_vertices, "medium_points": st.medium_vertices, "beams": st.beams,
                     "subbeams": st.subbeams, "bpt_connections": st.bpt_connections,
                     "point_queries": st.point_queries, "segment_queries": st.segment_queries,
                     "merge_accepts": st.merge_accepts, "pb2d_accepts": st.pb2d_accepts,
                     "bb1d_accepts": st.bb1d_accepts, "seg_nodes": st.seg_nodes,
                     "seg_candidates": st.seg_candidates, "seg_pieces": st.seg_pieces,
                     "bundles": st.bundles, "bundle_pieces_left": st.bundle_pieces_left,
                     "bundle_nodes": st.bundle_nodes, "bundle_leaves": st.bundle_leaves,
                     "bundle_tests": st.bundle_tests, "path_errors": st.path_errors,
                     "beam_drops": st.beam_drops},
        "roofline": {
            "kernel": st.dominant_kernel.decode(),
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(st.dominant_kernel.decode()),
            "traffic_source": "profiles/seg_merge_blobs_hbm_traffic.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; bytes)",
            "avg_launch_ms": dom_ms,
            "algorithmic_bytes_per_launch": st.dominant_bytes / launches,
            "measured_copy": copy,
        },
    }
    # the iteration as a whole: compulsory bytes of every pass (path state read and written per
    # extend step, vertex / beam records written and read, LBVH builds, framebuffers) over the
    # iteration's wall time
    t_iter = st.ms_total * 1e-3
    out["iteration_roofline"] = {"algorithmic_bytes": st.algorithmic_bytes,
                                 "achieved": st.algorithmic_bytes / t_iter / 1e9 if t_iter > 0 else 0.0,
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": st.algorithmic_bytes / t_iter / 1e9 / HBM_PEAK_GBS if t_iter > 0 else 0.0}
    # segment queries with their counted traversal: 44 bytes per query (record and result), 128
    # per tree node fetched (one 8-wide node), 48 per sub-beam or point record tested by the
    # per-piece walk, 48 + 32 per sub-beam a bundle stages (record, and its box from its node),
    # the record of every accepted photon (112) or beam (128), over the two merge stages' time
    t_q = (st.ms_merge_pb2d + st.ms_merge_bb1d) * 1e-3
    b_q = (44.0 * st.segment_queries + 128.0 * (st.seg_nodes + st.bundle_nodes) + 48.0 * st.seg_candidates +
           80.0 * st.bundle_leaves + 112.0 * st.pb2d_accepts + 128.0 * st.bb1d_accepts)
    out["query_roofline"] = {"bytes": b_q, "queries": st.segment_queries,
                             "achieved": b_q / t_q / 1e9 if t_q > 0 else 0.0, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": b_q / t_q / 1e9 / HBM_PEAK_GBS if t_q > 0 else 0.0}
    out["paths_per_s_frac_of_roofline"] = out["iteration_roofline"]["frac"]
    print("bench: GPU done, %.1f ms per iteration" % ms, file=sys.stderr, flush=True)
    if not a.no_learned:
        out["learned_vs_plain"] = learned_vs_plain(workdir, a.learned_seconds)
    if a.batch_in_default:
        print("bench: batched jobs", file=sys.stderr, flush=True)
        out["batched_jobs"] = batch_jobs(workdir, n_jobs=64, iters=3, threads=8)
        out["batched_jobs"]["measured"] = "in this run"
    stress_run = committed("stress_4k.json")
    if stress_run:
        out["stress"] = {"workload": stress_run["workload"], "ms_per_iteration": stress_run["plain"]["ms_per_iteration"],
                         "samples_per_s": stress_run["plain"]["samples_per_s"],
                         "source": "profiles/stress_4k.md (python bench.py --stress)"}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sh.ptr, cfg)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    sh.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
