"""Benchmark: UPBP renderer iteration on one MI355X.

Workload (the plain-UPBP-at-scale configuration): seeded grid of 64 noise-displaced
icosphere blobs (1,310,720 triangles) on a floor, three overlapping homogeneous media
spheres with priorities 1-3, two quad lights and a point light, 1280x720, max path
length 12, 900,000 light paths per iteration, radius 0.005, upbp_all
(BPT + surface PM + P-P3D + P-B2D + B-B1D), FP32 transport, FP64 accumulation.

One step = one 1-spp iteration (light pass, photon LBVH builds, camera pass, FP64
accumulate) with the scene resident in HBM. Prints one JSON line; it also carries the
committed measurements of the longer runs below (profiles/*.json).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--autotune]

Other measurements (each prints its own JSON line):
  --learned SECONDS [--ref-seconds S]  learned proposal vs plain renderer at equal time on the
                                       slab room (guide settings: --alpha, --surface-alpha,
                                       --guide-res, --guide-steps, --guide-lr, --guide-history;
                                       comma lists compare several against one reference)
  --batch JOBS                         batched independent jobs (upbp_run_batch)
  --stress                             3840x2160, 2.05M triangles, 8 media, 8M light paths
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


LEARNED_ARGS = ["-l", "16", "-r_initial", "0.01", "-a", "upbp_all"]


def equal_time(seconds, ref_seconds, workdir, res, guides):
    """Learned proposal vs plain renderer at equal wall-clock time on the slab-room workload
    (196k triangles, dense g=0.9 slab lit through a 5% aperture): relMSE of each against a
    plain reference render of ref_seconds (two halves on other random streams). Each learned
    run's time includes its online training. The reference's own error, estimated from its
    halves as relMSE(half_a, half_b) / 4, is reported and also taken off every figure.
    guides: learned-proposal settings to compare (dicts of upbp_guide_config fields), each
    against the same reference and plain run. Returns the result dictionary (the first setting
    at the top level, all of them under "learned")."""
    import numpy as np
    from gpuupbp_amd import capi
    from gpuupbp_amd.renderer import Renderer
    obj = os.path.join(workdir, "slabroom.obj")
    if not os.path.exists(obj + ".aux"):
        capi.generate_scene("slabroom", 3, obj)
    rc, cfg = capi.parse_args(["-s", "-1", obj, "-r", "%dx%d" % res] + LEARNED_ARGS)
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)

    def run(sec, g, first, label):
        r = Renderer(sh.ptr, cfg)
        if g is not None:
            r.set_guide(**g)
        r.sync()
        t0 = time.perf_counter()
        n, last = 0, t0
        while True:
            r.run_iteration(first + n, want_stats=False)
            r.sync()
            n += 1
            now = time.perf_counter()
            if now - last > 20:
                print("bench: %s %d iterations, %.0f s" % (label, n, now - t0), file=sys.stderr, flush=True)
                last = now
            if now - t0 >= sec:
                break
        img = r.read_accum(n)
        r.close()
        return img, n, time.perf_counter() - t0

    ref_a, n_a, _ = run(ref_seconds / 2, None, 1 << 20, "reference a")
    ref_b, n_b, _ = run(ref_seconds / 2, None, 1 << 21, "reference b")
    ref = (ref_a * n_a + ref_b * n_b) / (n_a + n_b)
    rel = lambda im, r: float(np.mean((im - r) ** 2 / (r ** 2 + 1e-2)))
    ref_err = rel(ref_a, ref_b) / 4
    plain, n_plain, t_plain = run(seconds, None, 0, "plain")
    rp = rel(plain, ref)
    learned = []
    for g in guides:
        img, n, t = run(seconds, g, 0, "learned %s" % g)
        rl = rel(img, ref)
        learned.append({"guide": g, "iterations": n, "wall_s": round(t, 2), "relmse": rl,
                        "relmse_less_reference": rl - ref_err, "ratio_plain_over_learned": rp / rl,
                        "ratio_less_reference": (rp - ref_err) / (rl - ref_err)})
    first = learned[0]
    return {"workload": "slabroom: 194,970 tris, slab sigma_t 7.23 (U[2,10], seed 3), albedo 0.95, g 0.9, "
                        "5%% aperture, %dx%d, l=16, r=0.01, upbp_all" % res,
            "seconds": seconds, "reference_seconds": ref_seconds, "reference_iterations": n_a + n_b,
            "iterations_plain": n_plain, "iterations_learned": first["iterations"],
            "wall_plain_s": round(t_plain, 2), "wall_learned_s": first["wall_s"],
            "relmse_plain": rp, "relmse_learned": first["relmse"],
            "relmse_ratio_plain_over_learned": first["ratio_plain_over_learned"],
            "reference_relmse_estimate": ref_err,
            "relmse_plain_less_reference": rp - ref_err, "relmse_learned_less_reference": first["relmse_less_reference"],
            "relmse_ratio_less_reference": first["ratio_less_reference"],
            "guide": first["guide"], "learned": learned}


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
    """Batched jobs: n_jobs independent renderers drawn with seed 2 -- scene box / boxfog /
    slab room (small and 195k-triangle), resolution in {64, 128, 256, 512}^2, max length in
    [4, 16], radius 0.01 -- each rendering `iters` iterations, one after another and then all at
    once through upbp_run_batch. Returns the measurement."""
    import numpy as np
    from gpuupbp_amd import capi
    from gpuupbp_amd.renderer import Renderer, run_batch
    rng = np.random.default_rng(2)
    kinds = ["box", "boxfog", "slabroom_small", "slabroom"]
    objs = {}
    jobs = []
    for j in range(n_jobs):
        kind = kinds[int(rng.integers(len(kinds)))]
        seed = int(rng.integers(4)) if kind != "slabroom" else 3
        key = (kind, seed)
        if key not in objs:
            d = os.path.join(workdir, "batch_%s_%d" % key)
            os.makedirs(d, exist_ok=True)
            objs[key] = capi.generate_scene(kind, seed, os.path.join(d, kind + ".obj"))
        res = int(rng.choice([64, 128, 256, 512]))
        length = int(rng.integers(4, 17))
        rc, cfg = capi.parse_args(["-s", "-1", objs[key], "-r", "%dx%d" % (res, res), "-l", str(length),
                                   "-r_initial", "0.01", "-a", "upbp_all"])
        assert rc == 0
        sh = capi.SceneHandle(objs[key], cfg.res_x, cfg.res_y)
        jobs.append((sh, Renderer(sh.ptr, cfg), res * res, kind, length))
    rs = [r for _, r, _, _, _ in jobs]
    pixels = sum(px for _, _, px, _, _ in jobs)
    run_batch(rs, [0] * len(rs), threads)  # warm-up: buffers grow to their sizes
    for r in rs:
        r.sync()
    t0 = time.perf_counter()
    for i in range(1, iters + 1):
        for r in rs:
            r.run_iteration(i, want_stats=False)
            r.sync()
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    for i in range(iters + 1, 2 * iters + 1):
        run_batch(rs, [i] * len(rs), threads)
    t_batch = time.perf_counter() - t0
    for sh, r, _, _, _ in jobs:
        r.close()
        sh.close()
    mix = {k: sum(1 for _, _, _, kk, _ in jobs if kk == k) for k in kinds}
    return {"jobs": n_jobs, "iterations_per_job": iters, "threads": threads, "job_mix": mix,
            "pixels_per_round": pixels, "seconds_sequential": round(t_seq, 3), "seconds_batched": round(t_batch, 3),
            "samples_per_s_sequential": pixels * iters / t_seq, "samples_per_s_batched": pixels * iters / t_batch,
            "speedup": t_seq / t_batch}


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


def learned_vs_plain():
    """The committed equal-time comparison of the learned proposal with the plain renderer on the
    slab room (bench.py --learned; a ten-minute run, so the default run reports the committed
    measurement rather than repeating it), or None."""
    path = os.path.join(ROOT, "profiles", "learned_vs_plain_slabroom.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    # the run with the library's default guide settings (lr 8e-3 is the default)
    best = next((x for x in d["learned"] if x["guide"] in ({}, {"learning_rate": 0.008})), d["learned"][0])
    return {"workload": d["workload"], "seconds": d["seconds"], "reference_seconds": d["reference_seconds"],
            "relmse_plain": d["relmse_plain"], "relmse_learned": best["relmse"], "guide": best["guide"],
            "ratio_plain_over_learned": best["ratio_plain_over_learned"],
            "ratio_less_reference_error": best["ratio_less_reference"],
            "source": "profiles/learned_vs_plain_slabroom.md (python bench.py --learned 60 --ref-seconds 480)"}


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
                    help="batched-jobs measurement over JOBS independent renderers (upbp_run_batch)")
    ap.add_argument("--batch-threads", type=int, default=8)
    ap.add_argument("--learned", type=float, default=None,
                    help="seconds per renderer of the learned-vs-plain equal-time comparison (slab room)")
    ap.add_argument("--ref-seconds", type=float, default=None, help="reference render time (default 5 x --learned)")
    ap.add_argument("--learned-res", default="1280x720")
    ap.add_argument("--alpha", default=None,
                    help="learned-proposal mixture weight in media (library default); a comma list compares several")
    ap.add_argument("--surface-alpha", default=None,
                    help="mixture weight at surfaces (library default); a comma list pairs with --alpha")
    ap.add_argument("--guide-res", default=None, help="record grid cells per axis (comma list allowed)")
    ap.add_argument("--guide-steps", default=None, help="Adam steps per iteration (comma list allowed)")
    ap.add_argument("--guide-lr", default=None, help="Adam step size (comma list allowed)")
    ap.add_argument("--guide-history", default=None, help="training-sample ring size (comma list allowed)")
    a = ap.parse_args()
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
    if a.learned is not None:
        w, h = (int(v) for v in a.learned_res.split("x"))
        lists = {k: ([conv(v) for v in str(val).split(",")] if val is not None else [None])
                 for k, val, conv in (("alpha", a.alpha, float), ("surface_alpha", a.surface_alpha, float),
                                      ("grid_res", a.guide_res, int), ("steps_per_iteration", a.guide_steps, int),
                                      ("learning_rate", a.guide_lr, float), ("history", a.guide_history, int))}
        n = max(len(v) for v in lists.values())
        guides = [{k: v[i % len(v)] for k, v in lists.items() if v[i % len(v)] is not None} for i in range(n)]
        res = equal_time(a.learned, a.ref_seconds or 5 * a.learned, workdir, (w, h), guides)
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
                     "path_errors": st.path_errors,
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
    out["learned_vs_plain"] = learned_vs_plain()
    batch = committed("batched_jobs.json")
    if batch:
        best = max(batch["runs"], key=lambda x: x["speedup"])
        out["batched_jobs"] = {k: best[k] for k in ("jobs", "threads", "samples_per_s_sequential",
                                                    "samples_per_s_batched", "speedup")}
        out["batched_jobs"]["source"] = "profiles/batched_jobs.md (python bench.py --batch 64)"
    stress_run = committed("stress_4k.json")
    if stress_run:
        out["stress"] = {"workload": stress_run["workload"], "ms_per_iteration": stress_run["plain"]["ms_per_iteration"],
                         "samples_per_s": stress_run["plain"]["samples_per_s"],
                         "source": "profiles/stress_4k.md (python bench.py --stress)"}
    print("bench: GPU done, %.1f ms per iteration" % ms, file=sys.stderr, flush=True)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sh.ptr, cfg)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    sh.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
