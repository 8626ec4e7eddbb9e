"""Seeded workloads shared by the benchmark and the GPU tests.

Each workload is a procedural scene written as OBJ + MTL + AUX (upbp_generate_scene) and the
reference-compatible command line that renders it. Sizes are those of the benchmark
configurations; tests run them at full size or scaled down as they state.
"""
import os

import numpy as np

from . import capi

# the plain-UPBP-at-scale blob scene: 64 noise-displaced icospheres (1,310,720 triangles), three
# overlapping media spheres with priorities 1-3, two quad lights and a point light
BLOBS_ARGS = ["-r", "1280x720", "-l", "12", "-pcpi", "900000", "-r_initial", "0.005", "-a", "upbp_all"]
# the learned-proposal slab room: 194,970 triangles, a dense g = 0.9 slab lit through a 5% aperture
SLABROOM_ARGS = ["-r", "1280x720", "-l", "16", "-r_initial", "0.01", "-a", "upbp_all"]
# the combined stress scene: 2,048,006 triangles, eight media (g -0.3 .. 0.95)
STRESS_ARGS = ["-r", "3840x2160", "-l", "16", "-pcpi", "8000000", "-r_initial", "0.005", "-a", "upbp_all"]
# the CPU-runnable box: 5 walls, a quad light, an isotropic medium sphere r = 0.3
BOX_ARGS = ["-r", "256x256", "-l", "8", "-a", "upbp_bpt+surf+pp3d+pb2d+bb1d+bp2d", "-r_initial", "0.01",
            "-r_alpha", "1"]


def scene(workdir, kind, seed):
    """Writes (once) and returns the OBJ path of a generated scene."""
    d = os.path.join(workdir, "%s_%d" % (kind, seed))
    os.makedirs(d, exist_ok=True)
    obj = os.path.join(d, kind + ".obj")
    if not os.path.exists(obj + ".aux"):
        capi.generate_scene(kind, seed, obj)
    return obj


def scaled(args, factor):
    """The command line with the resolution divided by `factor` per axis and the light path count
    (when given) by factor^2: the same scene and path statistics on fewer pixels."""
    out = list(args)
    i = out.index("-r")
    w, h = (int(v) for v in out[i + 1].split("x"))
    out[i + 1] = "%dx%d" % (max(1, w // factor), max(1, h // factor))
    if "-pcpi" in out:
        j = out.index("-pcpi")
        out[j + 1] = str(max(1, int(float(out[j + 1])) // (factor * factor)))
    return out


BATCH_KINDS = ["box", "boxfog", "blobs_small", "slabroom_small", "slabroom"]


def batch_mix(workdir, n_jobs=64, seed=2):
    """The batched-jobs mix: n_jobs independent render jobs drawn with `seed` -- scene box / boxfog
    (10 and 1,290 triangles, one medium), a blob grid (1 to 9 blobs, 20k to 184k triangles, 1 to 4
    media) or the slab room (small, or the 194,970-triangle one); resolution in {64, 128, 256,
    512}^2; max path length in [4, 16]; radius 0.01; all estimators. Returns a list of
    (obj, argv, pixels, kind, max length)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n_jobs):
        kind = BATCH_KINDS[int(rng.integers(len(BATCH_KINDS)))]
        sseed = int(rng.integers(12)) if kind == "blobs_small" else (int(rng.integers(4)) if kind != "slabroom" else 3)
        obj = scene(workdir, kind, sseed)
        res = int(rng.choice([64, 128, 256, 512]))
        length = int(rng.integers(4, 17))
        argv = ["-s", "-1", obj, "-r", "%dx%d" % (res, res), "-l", str(length), "-r_initial", "0.01", "-a", "upbp_all"]
        jobs.append((obj, argv, res * res, kind, length))
    return jobs
