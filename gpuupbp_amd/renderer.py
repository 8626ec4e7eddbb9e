"""Python host-side mirror of the reference's renderer interface.

Mirrors Renderer::init / run_iteration / cleanup (gpuupbp/renderer/renderer.hpp:46-55)
and the main loop of gpuupbp/gpuupbp.cpp:57-148 over the C ABI of libupbp.so.
Raises if the HIP library is not built or no GPU is visible: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import capi


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Renderer:
    def __init__(self, scene_ptr, cfg):
        self.cfg = cfg
        self.h = C.c_void_p()
        rc = capi.lib().upbp_create(scene_ptr, C.byref(cfg), C.byref(self.h))
        if rc != 0:
            raise RuntimeError("upbp_create failed (%d): %s" % (rc, capi.lib().upbp_last_error(None).decode()))
        self.iterations = 0

    def run_iteration(self, iteration=None, want_stats=True):
        it = self.iterations if iteration is None else iteration
        st = capi.IterStats() if want_stats else None
        rc = capi.lib().upbp_run_iteration(self.h, it, C.byref(st) if want_stats else None)
        if rc != 0:
            raise RuntimeError("upbp_run_iteration failed (%d): %s" % (rc, capi.lib().upbp_last_error(self.h).decode()))
        self.iterations += 1
        return st

    def set_guide(self, **kw):
        """Learned proposal trained online (fields of upbp_guide_config; unnamed ones default)."""
        g = capi.GuideConfig()
        capi.lib().upbp_guide_defaults(C.byref(g))
        for k, v in kw.items():
            setattr(g, k, v)
        rc = capi.lib().upbp_set_guide(self.h, C.byref(g))
        if rc != 0:
            raise RuntimeError("upbp_set_guide failed: %s" % capi.lib().upbp_last_error(self.h).decode())

    def set_guide_lobes(self, lobes, res, lo, hi, alpha):
        """Install explicit learned-proposal lobes ((res^3, 8, 5): mu.xyz, kappa, weight)."""
        lobes = np.ascontiguousarray(lobes, np.float32)
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        rc = capi.lib().upbp_set_guide_lobes(self.h, _p(lobes), res, _p(lo), _p(hi), alpha)
        if rc != 0:
            raise RuntimeError("upbp_set_guide_lobes failed: %s" % capi.lib().upbp_last_error(self.h).decode())

    def autotune(self, iteration=0, repeats=1):
        """Runs upbp_autotune: returns the chosen tuning (dict) and its ms per iteration."""
        t = capi.Tuning()
        ms = C.c_float()
        rc = capi.lib().upbp_autotune(self.h, iteration, repeats, C.byref(t), C.byref(ms))
        if rc != 0:
            raise RuntimeError("upbp_autotune failed: %s" % capi.lib().upbp_last_error(self.h).decode())
        return {f: getattr(t, f) for f, _ in capi.Tuning._fields_}, ms.value

    def set_tuning(self, **kw):
        """Speed knobs of the C ABI's upbp_tuning (subbeam_factor, query_piece_factor,
        merge_blocks_per_cu, piece_order, packet_bounces, block_size, beam_leaf, graph,
        bundle_bounces, bundle_factor, bundle_piece_factor); unnamed fields keep their defaults."""
        t = capi.Tuning()
        capi.lib().upbp_tuning_defaults(C.byref(t))
        for k, v in kw.items():
            if v is not None:
                setattr(t, k, v)
        rc = capi.lib().upbp_set_tuning(self.h, C.byref(t))
        if rc != 0:
            raise RuntimeError("upbp_set_tuning failed: %s" % capi.lib().upbp_last_error(self.h).decode())
        return t

    def set_mis_log(self, capacity, pixel_stride=1, light_paths=0):
        """upbp_set_mis_log: log every MIS-weighted contribution of the camera paths whose index is
        a multiple of pixel_stride and of light tracing from the first light_paths light paths
        (capacity 0: off)."""
        rc = capi.lib().upbp_set_mis_log(self.h, capacity, pixel_stride, light_paths)
        if rc != 0:
            raise RuntimeError("upbp_set_mis_log failed: %d %s" % (rc, capi.lib().upbp_last_error(self.h).decode()))

    def read_guide_samples(self, max_samples=1 << 22):
        """The training ring's samples ((n, 10): unit-cube position, zero arrival, scattered
        direction, weight), the record grid's low corner and its edge length."""
        buf = np.zeros((max_samples, 10), np.float32)
        grid = np.zeros(4, np.float32)
        n = C.c_uint64()
        rc = capi.lib().upbp_read_guide_samples(self.h, _p(buf), max_samples, C.byref(n), _p(grid))
        if rc != 0:
            raise RuntimeError("upbp_read_guide_samples failed: %d %s"
                               % (rc, capi.lib().upbp_last_error(self.h).decode()))
        return buf[: min(n.value, max_samples)].copy(), grid[:3].copy(), float(grid[3])

    def read_mis_log(self, max_records=1 << 22):
        """The last iteration's MIS weight log as a structured array with fields tech, k, s, id
        (camera path; light tracing: light path), lpath (light path; light hits and next-event
        estimation: the light) and w (FP32 weight)."""
        buf = np.zeros((max_records, 4), np.uint32)
        n = C.c_uint64()
        rc = capi.lib().upbp_read_mis_log(self.h, _p(buf), max_records, C.byref(n))
        if rc != 0:
            raise RuntimeError("upbp_read_mis_log failed: %d %s" % (rc, capi.lib().upbp_last_error(self.h).decode()))
        r = buf[: min(n.value, max_records)]
        out = np.zeros(len(r), dtype=[("tech", "i4"), ("k", "i4"), ("s", "i4"), ("id", "i8"), ("lpath", "i8"),
                                      ("w", "f8")])
        out["tech"] = r[:, 0] & 255
        out["k"] = (r[:, 0] >> 8) & 255
        out["s"] = r[:, 0] >> 16
        out["id"] = r[:, 1]
        out["lpath"] = r[:, 2]
        out["w"] = r[:, 3].view(np.float32)
        return out

    def sync(self):
        rc = capi.lib().upbp_sync(self.h)
        if rc != 0:
            raise RuntimeError("upbp_sync failed: %s" % capi.lib().upbp_last_error(self.h).decode())

    def read_accum(self, iters_done=None):
        n = self.iterations if iters_done is None else iters_done
        img = np.zeros((self.cfg.res_y, self.cfg.res_x, 3), np.float64)
        rc = capi.lib().upbp_read_accum(self.h, _p(img), n)
        if rc != 0:
            raise RuntimeError("upbp_read_accum failed: %s" % capi.lib().upbp_last_error(self.h).decode())
        return img

    def close(self):
        if self.h:
            capi.lib().upbp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def range_query(kind, prims, queries, radius, max_indices=None, subbeam_len=0.0, wide=False):
    """Exact accepted index sets of an LBVH range query (CSR: offsets, indices). Beams (kinds 2, 3)
    enter the LBVH as sub-beams of length about subbeam_len (0: whole beams); wide=True walks the
    4-wide collapse (segment kinds 1 and 3)."""
    prims = np.ascontiguousarray(prims, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    nq = queries.shape[0]
    cap = max_indices if max_indices is not None else max(1, nq * max(prims.shape[0], 1))
    offsets = np.zeros(nq + 1, np.uint32)
    idx = np.zeros(cap, np.uint32)
    n = C.c_uint64()
    rc = capi.lib().upbp_test_range_query_wide(kind, _p(prims), prims.shape[0], _p(queries), nq, radius,
                                               subbeam_len, int(wide), _p(offsets), _p(idx), cap, C.byref(n))
    if rc != 0:
        raise RuntimeError("upbp_test_range_query failed: %d" % rc)
    return offsets, idx[: n.value]


def sphere_hits(spheres, rays):
    """GPU analytic sphere test of every (ray, sphere) pair: (n_rays, n_spheres, 3) of root count,
    near root, far root (inf if absent)."""
    spheres = np.ascontiguousarray(spheres, np.float32)
    rays = np.ascontiguousarray(rays, np.float32)
    out = np.zeros((rays.shape[0], spheres.shape[0], 3), np.float32)
    rc = capi.lib().upbp_test_sphere_hits(_p(spheres), spheres.shape[0], _p(rays), rays.shape[0], _p(out))
    if rc != 0:
        raise RuntimeError("upbp_test_sphere_hits failed: %d" % rc)
    return out


def env_samples(rgb, u):
    """GPU environment-map sampling and lookup of an (H, W, 3) map at u (n, 2): (n, 11) floats
    (direction, pdf, radiance, lookup pdf, lookup radiance)."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    env = capi.EnvMap(rgb.shape[1], rgb.shape[0], rgb.ctypes.data_as(C.POINTER(C.c_float)))
    out = np.zeros((u.shape[0], 11), np.float32)
    rc = capi.lib().upbp_test_env_samples(C.byref(env), _p(u), u.shape[0], _p(out))
    if rc != 0:
        raise RuntimeError("upbp_test_env_samples failed: %d" % rc)
    return out


class Mlp:
    """The learned proposal's network on the GPU (BF16 MFMA forward / backward / Adam)."""

    def __init__(self, seed=3):
        self.h = C.c_void_p()
        rc = capi.lib().upbp_mlp_create(seed, C.byref(self.h))
        if rc != 0:
            raise RuntimeError("upbp_mlp_create failed: %d" % rc)
        self.n = int(capi.lib().upbp_mlp_param_count())

    def close(self):
        if self.h:
            capi.lib().upbp_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.n, np.float32)
        if capi.lib().upbp_mlp_get_params(self.h, _p(p)) != 0:
            raise RuntimeError("upbp_mlp_get_params failed")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if capi.lib().upbp_mlp_set_params(self.h, _p(p)) != 0:
            raise RuntimeError("upbp_mlp_set_params failed")

    def forward(self, samples):
        samples = np.ascontiguousarray(samples, np.float32)
        out = np.zeros((samples.shape[0], 40), np.float32)
        if capi.lib().upbp_mlp_forward(self.h, _p(samples), samples.shape[0], _p(out)) != 0:
            raise RuntimeError("upbp_mlp_forward failed")
        return out

    def loss_grad(self, samples):
        samples = np.ascontiguousarray(samples, np.float32)
        loss = C.c_float()
        g = np.zeros(self.n, np.float32)
        if capi.lib().upbp_mlp_loss_grad(self.h, _p(samples), samples.shape[0], C.byref(loss), _p(g)) != 0:
            raise RuntimeError("upbp_mlp_loss_grad failed")
        return loss.value, g

    def train_step(self, samples, lr=1e-3):
        samples = np.ascontiguousarray(samples, np.float32)
        loss = C.c_float()
        if capi.lib().upbp_mlp_train_step(self.h, _p(samples), samples.shape[0], lr, C.byref(loss)) != 0:
            raise RuntimeError("upbp_mlp_train_step failed")
        return loss.value


def guide_samples(lobes, u, arrive=(0.0, 0.0, 1.0), g=0.0, surface=False):
    """GPU sampling of one learned-proposal record (8, 5) times the phase lobes of a path
    arriving along `arrive` in a medium of mean cosine g (or, with surface, times the cosine
    lobe about the normal `arrive`), at u (n, 3): (n, 4) direction, pdf."""
    lobes = np.ascontiguousarray(lobes, np.float32)
    arrive = np.ascontiguousarray(arrive, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    out = np.zeros((u.shape[0], 4), np.float32)
    rc = capi.lib().upbp_test_guide_samples(_p(lobes), _p(arrive), g, int(surface), _p(u), u.shape[0], _p(out))
    if rc != 0:
        raise RuntimeError("upbp_test_guide_samples failed: %d" % rc)
    return out


def run_batch(renderers, iterations, threads=8):
    """upbp_run_batch: one iteration of each renderer (iterations[i] for renderers[i]) with up to
    `threads` host threads driving their streams side by side."""
    n = len(renderers)
    hs = (C.c_void_p * n)(*[r.h for r in renderers])
    its = np.ascontiguousarray(iterations, np.int32)
    rc = capi.lib().upbp_run_batch(hs, _p(its), n, threads)
    if rc != 0:
        raise RuntimeError("upbp_run_batch failed: %d" % rc)
    for r, it in zip(renderers, iterations):
        r.iterations = max(r.iterations, int(it) + 1)


class Batch:
    """upbp_batch_*: many renderers rendered in rounds of one launch per stage (see the C ABI).
    run(iterations) renders one iteration of each renderer; counts() returns (rounds, graph
    replays, eager batched rounds, rounds of jobs alone). Close it before its renderers."""

    def __init__(self, renderers, use_graph=True):
        self.renderers = list(renderers)
        self.use_graph = int(bool(use_graph))
        hs = (C.c_void_p * len(self.renderers))(*[r.h for r in self.renderers])
        self.h = C.c_void_p()
        rc = capi.lib().upbp_batch_create(hs, len(self.renderers), C.byref(self.h))
        if rc != 0:
            raise RuntimeError("upbp_batch_create failed: %d %s" % (rc, capi.lib().upbp_batch_last_error(None).decode()))

    def run(self, iterations):
        its = np.ascontiguousarray(iterations, np.int32)
        if len(its) != len(self.renderers):
            raise ValueError("one iteration per renderer")
        rc = capi.lib().upbp_batch_run(self.h, _p(its), self.use_graph)
        if rc != 0:
            raise RuntimeError("upbp_batch_run failed: %d %s" % (rc, capi.lib().upbp_batch_last_error(self.h).decode()))
        for r, it in zip(self.renderers, its):
            r.iterations = max(r.iterations, int(it) + 1)

    def counts(self):
        v = [C.c_uint64() for _ in range(4)]
        capi.lib().upbp_batch_counts(self.h, *[C.byref(x) for x in v])
        return tuple(int(x.value) for x in v)

    def close(self):
        if self.h:
            capi.lib().upbp_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def closest_hit(verts, rays):
    verts = np.ascontiguousarray(verts, np.float32)
    rays = np.ascontiguousarray(rays, np.float32)
    t = np.zeros(rays.shape[0], np.float32)
    prim = np.zeros(rays.shape[0], np.int32)
    rc = capi.lib().upbp_test_closest_hit(_p(verts), verts.shape[0], _p(rays), rays.shape[0], _p(t), _p(prim))
    if rc != 0:
        raise RuntimeError("upbp_test_closest_hit failed: %d" % rc)
    return t, prim


def save_exr(path, rgb):
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = rgb.shape
    rc = capi.lib().upbp_save_exr(path.encode(), _p(rgb), w, h)
    if rc != 0:
        raise RuntimeError("upbp_save_exr failed: %d" % rc)


def save_image(path, rgb):
    """EXR, BMP (gamma 2.2) or HDR by the file extension, like the reference's save()."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = rgb.shape
    rc = capi.lib().upbp_save_image(path.encode(), _p(rgb), w, h)
    if rc != 0:
        raise RuntimeError("upbp_save_image failed: %d" % rc)


def render(obj_path, argv, iterations=None, want_stats=False):
    """Parse argv like the CLI, load the OBJ scene, render, return the mean image."""
    rc, cfg = capi.parse_args(["-s", "-1", obj_path] + list(argv))
    if rc != 0:
        raise RuntimeError("command line rejected")
    sh = capi.SceneHandle(obj_path, cfg.res_x, cfg.res_y)
    n = iterations if iterations is not None else cfg.iteration_count
    stats = []
    with Renderer(sh.ptr, cfg) as r:
        for i in range(n):
            s = r.run_iteration(i, want_stats)
            if want_stats:
                stats.append(s)
        img = r.read_accum(n)
    sh.close()
    return (img, stats) if want_stats else img
