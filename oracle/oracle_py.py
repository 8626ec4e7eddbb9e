"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
_lib = None


class OracleStats(C.Structure):
    _fields_ = [("light_vertices", C.c_uint64), ("surface_vertices", C.c_uint64),
                ("medium_vertices", C.c_uint64), ("beams", C.c_uint64), ("path_errors", C.c_uint64),
                ("seconds", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        h = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        h.oracle_render.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, C.POINTER(OracleStats)]
        h.oracle_render_sample.argtypes = [vp, vp, C.c_int, C.c_uint32, C.c_uint32, C.c_int, vp, vp]
        h.oracle_training_samples.argtypes = [vp, vp, C.c_int, C.c_int32, C.c_int32, C.c_int, vp, C.c_uint64,
                                              C.POINTER(C.c_uint64)]
        h.oracle_range_query.argtypes = [C.c_int, vp, C.c_uint32, vp, C.c_uint32, C.c_float, vp, vp, C.c_uint64,
                                         C.POINTER(C.c_uint64)]
        h.oracle_closest_hit.argtypes = [vp, C.c_uint32, vp, C.c_uint32, vp, vp]
        h.oracle_sphere_hits.argtypes = [vp, C.c_uint32, vp, C.c_uint32, vp]
        h.oracle_env_samples.argtypes = [vp, vp, C.c_uint32, vp]
        h.oracle_set_guide_lobes.argtypes = [vp, C.c_int32, vp, vp, C.c_float]
        h.oracle_guide_samples.argtypes = [vp, vp, C.c_float, C.c_int32, vp, C.c_uint32, vp]
        h.oracle_mlp_param_count.restype = C.c_uint32
        h.oracle_mlp_encode.argtypes = [vp, vp, vp]
        h.oracle_mlp_forward.argtypes = [vp, vp, C.c_uint32, vp]
        h.oracle_mlp_loss_grad.argtypes = [vp, vp, C.c_uint32, vp, vp]
        h.oracle_camera_ray.argtypes = [vp, vp, vp, C.c_uint32, vp]
        h.oracle_world_to_raster.argtypes = [vp, vp, vp, C.c_uint32, vp]
        h.oracle_hg_pdf.argtypes = [C.c_float, C.c_float]
        h.oracle_hg_pdf.restype = C.c_float
        h.oracle_hg_sample.argtypes = [C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        h.oracle_fresnel.argtypes = [C.c_float, C.c_float]
        h.oracle_fresnel.restype = C.c_float
        h.oracle_philox.argtypes = [vp, vp, vp]
        h.oracle_mis_check.argtypes = [vp, vp, C.c_int, vp, C.c_uint32, C.c_uint32, C.c_int, vp, C.c_uint32,
                                       C.POINTER(C.c_uint32)]
        h.oracle_adam_step.argtypes = [vp, vp, vp, vp, C.c_uint32, C.c_int, C.c_double]
        h.oracle_bsdf_eval.argtypes = [vp, vp, C.c_float, C.c_int, vp, C.c_uint32, vp]
        h.oracle_bsdf_sample.argtypes = [vp, vp, C.c_float, C.c_int, vp, C.c_uint32, vp]
        _lib = h
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def render(scene_ptr, cfg, first_iteration=0, n_iterations=1, threads=0, want_fb=False):
    """Mean FP64 image (H, W, 3) over n_iterations; optionally the FP32 framebuffer of the first."""
    out = np.zeros((cfg.res_y, cfg.res_x, 3), np.float64)
    fb = np.zeros((cfg.res_y, cfg.res_x, 3), np.float32) if want_fb else None
    st = OracleStats()
    rc = lib().oracle_render(C.cast(scene_ptr, C.c_void_p), C.cast(C.pointer(cfg), C.c_void_p), first_iteration,
                             n_iterations, threads, _p(out), _p(fb) if want_fb else None, C.byref(st))
    if rc != 0:
        raise RuntimeError("oracle_render failed: %d" % rc)
    return (out, fb, st) if want_fb else (out, st)


def render_sample_times(scene_ptr, cfg, n_camera_paths, n_light_paths, threads=0, iteration=0, want_stats=False):
    """Wall seconds (light pass, photon-map build, camera pass) of a bounded sample iteration
    (and the light pass's counts with want_stats)."""
    t = np.zeros(3, np.float64)
    st = OracleStats()
    rc = lib().oracle_render_sample(C.cast(scene_ptr, C.c_void_p), C.cast(C.pointer(cfg), C.c_void_p), iteration,
                                    n_camera_paths, n_light_paths, threads, _p(t), C.cast(C.pointer(st), C.c_void_p))
    if rc != 0:
        raise RuntimeError("oracle_render_sample failed: %d" % rc)
    return (t, st) if want_stats else t


def training_samples(scene_ptr, cfg, iteration=0, train_media=True, train_surfaces=True, threads=0):
    """The learned proposal's training samples of one iteration rendered without the proposal:
    (n, 8) float32 rows (pixel, world position, scattered direction, weight)."""
    n = C.c_uint64()
    rc = lib().oracle_training_samples(C.cast(scene_ptr, C.c_void_p), C.cast(C.pointer(cfg), C.c_void_p), iteration,
                                       int(train_media), int(train_surfaces), threads, None, 0, C.byref(n))
    if rc != 0:
        raise RuntimeError("oracle_training_samples failed: %d" % rc)
    out = np.zeros((max(1, n.value), 8), np.float32)
    rc = lib().oracle_training_samples(C.cast(scene_ptr, C.c_void_p), C.cast(C.pointer(cfg), C.c_void_p), iteration,
                                       int(train_media), int(train_surfaces), threads, _p(out), n.value, C.byref(n))
    if rc != 0:
        raise RuntimeError("oracle_training_samples failed: %d" % rc)
    return out[: n.value]


def range_query(kind, prims, queries, radius, max_indices=None):
    prims = np.ascontiguousarray(prims, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    nq = queries.shape[0]
    npr = prims.shape[0]
    cap = max_indices if max_indices is not None else max(1, nq * npr)
    offsets = np.zeros(nq + 1, np.uint32)
    idx = np.zeros(cap, np.uint32)
    n = C.c_uint64()
    rc = lib().oracle_range_query(kind, _p(prims), npr, _p(queries), nq, radius, _p(offsets), _p(idx), cap,
                                  C.byref(n))
    if rc != 0:
        raise RuntimeError("oracle_range_query failed: %d" % rc)
    return offsets, idx[: n.value]


def closest_hit(verts, rays):
    verts = np.ascontiguousarray(verts, np.float32)
    rays = np.ascontiguousarray(rays, np.float32)
    t = np.zeros(rays.shape[0], np.float32)
    prim = np.zeros(rays.shape[0], np.int32)
    lib().oracle_closest_hit(_p(verts), verts.shape[0], _p(rays), rays.shape[0], _p(t), _p(prim))
    return t, prim


def sphere_hits(spheres, rays):
    """(n_rays, n_spheres, 3): root count in (tmin, tmax), near root, far root (inf if absent)."""
    spheres = np.ascontiguousarray(spheres, np.float32)
    rays = np.ascontiguousarray(rays, np.float32)
    out = np.zeros((rays.shape[0], spheres.shape[0], 3), np.float32)
    lib().oracle_sphere_hits(_p(spheres), spheres.shape[0], _p(rays), rays.shape[0], _p(out))
    return out


def guide_samples(lobes, u, arrive=(0.0, 0.0, 1.0), g=0.0, surface=False):
    """One learned-proposal record (8, 5) times the phase lobes of a path arriving along `arrive`
    in a medium of mean cosine g (or, with surface, the cosine lobe about the normal `arrive`),
    sampled at u (n, 3): (n, 4) direction and product pdf."""
    lobes = np.ascontiguousarray(lobes, np.float32)
    arrive = np.ascontiguousarray(arrive, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    out = np.zeros((u.shape[0], 4), np.float32)
    lib().oracle_guide_samples(_p(lobes), _p(arrive), g, int(surface), _p(u), u.shape[0], _p(out))
    return out


def mlp_param_count():
    return int(lib().oracle_mlp_param_count())


def mlp_encode(x, d):
    out = np.zeros(51, np.float32)
    lib().oracle_mlp_encode(_p(np.ascontiguousarray(x, np.float32)), _p(np.ascontiguousarray(d, np.float32)), _p(out))
    return out


def mlp_forward(params, samples):
    """FP64 forward: samples (n, 6) -> raw outputs (n, 40)."""
    params = np.ascontiguousarray(params, np.float64)
    samples = np.ascontiguousarray(samples, np.float32)
    out = np.zeros((samples.shape[0], 40), np.float64)
    lib().oracle_mlp_forward(_p(params), _p(samples), samples.shape[0], _p(out))
    return out


def mlp_loss_grad(params, samples, want_grad=True):
    """FP64 weighted mixture NLL of samples (n, 10) and its gradient."""
    params = np.ascontiguousarray(params, np.float64)
    samples = np.ascontiguousarray(samples, np.float32)
    loss = C.c_double()
    g = np.zeros(params.shape[0], np.float64) if want_grad else None
    lib().oracle_mlp_loss_grad(_p(params), _p(samples), samples.shape[0], C.byref(loss), _p(g) if want_grad else None)
    return loss.value, g


def adam_step(params, m, v, grads, step, lr):
    """FP64 Adam step (beta1 0.9, beta2 0.999, eps 1e-8); returns new (params, m, v)."""
    p = np.array(params, np.float64)
    mm = np.array(m, np.float64)
    vv = np.array(v, np.float64)
    g = np.ascontiguousarray(grads, np.float64)
    if lib().oracle_adam_step(_p(p), _p(mm), _p(vv), _p(g), len(p), step, lr) != 0:
        raise RuntimeError("oracle_adam_step failed")
    return p, mm, vv


def set_guide(lobes, res, lo, hi, alpha):
    """Install the learned proposal for later oracle renders (alpha 0: off)."""
    if alpha == 0:
        lib().oracle_set_guide_lobes(None, 0, None, None, 0.0)
        return
    lobes = np.ascontiguousarray(lobes, np.float32)
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    rc = lib().oracle_set_guide_lobes(_p(lobes), res, _p(lo), _p(hi), alpha)
    if rc != 0:
        raise RuntimeError("oracle_set_guide_lobes failed: %d" % rc)


class _EnvMap(C.Structure):  # upbp_env_map
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.POINTER(C.c_float))]


def env_samples(rgb, u):
    """Environment map (H, W, 3) sampled at u (n, 2): (n, 11) as upbp_test_env_samples."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    env = _EnvMap(rgb.shape[1], rgb.shape[0], rgb.ctypes.data_as(C.POINTER(C.c_float)))
    out = np.zeros((u.shape[0], 11), np.float32)
    rc = lib().oracle_env_samples(C.byref(env), _p(u), u.shape[0], _p(out))
    if rc != 0:
        raise RuntimeError("oracle_env_samples failed: %d" % rc)
    return out


def camera_rays(scene_ptr, cfg, raster_xy):
    raster_xy = np.ascontiguousarray(raster_xy, np.float32)
    d = np.zeros((raster_xy.shape[0], 3), np.float32)
    lib().oracle_camera_ray(C.cast(scene_ptr, C.c_void_p), C.cast(C.pointer(cfg), C.c_void_p), _p(raster_xy),
                            raster_xy.shape[0], _p(d))
    return d


def world_to_raster(scene_ptr, cfg, pts):
    pts = np.ascontiguousarray(pts, np.float32)
    r = np.zeros((pts.shape[0], 2), np.float32)
    lib().oracle_world_to_raster(C.cast(scene_ptr, C.c_void_p), C.cast(C.pointer(cfg), C.c_void_p), _p(pts),
                                 pts.shape[0], _p(r))
    return r


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().oracle_philox(_p(c), _p(k), _p(o))
    return o


MIS_TECH = {0: "light_hit", 1: "nee", 2: "bpt", 3: "light_tracing", 4: "surf", 5: "pp3d", 6: "pb2d", 7: "bp2d",
            8: "bb1d"}


def mis_check(scene_ptr, cfg, iteration, pixels, n_light_tracing_paths, threads=0, max_records=2000000):
    """Records (technique, s, k, constant-time weight, FP64 balance-heuristic weight, the same
    with the reference's next-event weight convention, camera path (pixel; light tracing: light
    path), light path (next-event estimation and light hits: the light)) of every contribution
    on the given camera paths and the first light paths of one iteration."""
    pixels = np.ascontiguousarray(pixels, np.uint32)
    out = np.zeros((max_records, 8), np.float64)
    n = C.c_uint32()
    rc = lib().oracle_mis_check(C.cast(scene_ptr, C.c_void_p), C.cast(C.pointer(cfg), C.c_void_p), iteration,
                                _p(pixels), len(pixels), n_light_tracing_paths, threads, _p(out), max_records,
                                C.byref(n))
    if rc != 0:
        raise RuntimeError("oracle_mis_check failed: %d" % rc)
    return out[: min(n.value, max_records)]


def _material(diffuse=(0, 0, 0), phong=(0, 0, 0), exponent=1.0, mirror=(0, 0, 0), ior=-1.0):
    from gpuupbp_amd import capi
    m = capi.Material()
    m.diffuse[:] = diffuse
    m.phong[:] = phong
    m.phong_exponent = exponent
    m.mirror[:] = mirror
    m.ior = ior
    m.priority = -1
    m.real = 1
    m.med_idx = -1
    return m


def bsdf_eval(material, ray_dir, dirs, ior=-1.0, from_light=False):
    """Surface BSDF (normal +z) of a ray travelling along ray_dir, evaluated toward dirs (n x 3):
    n x 7 (f.xyz, cosine at the generated side, forward pdf, reverse pdf, continuation prob.)."""
    m = _material(**material)
    rd = np.ascontiguousarray(ray_dir, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    out = np.zeros((len(d), 7), np.float32)
    if lib().oracle_bsdf_eval(C.byref(m), _p(rd), ior, int(from_light), _p(d), len(d), _p(out)) != 0:
        raise RuntimeError("oracle_bsdf_eval failed")
    return out


def bsdf_sample(material, ray_dir, u, ior=-1.0, from_light=False):
    """Samples of the same BSDF with variates u (n x 3): n x 9 (direction, pdf, cosine, f.xyz,
    event bits: 1 diffuse, 2 Phong, 4 reflect, 8 refract)."""
    m = _material(**material)
    rd = np.ascontiguousarray(ray_dir, np.float32)
    uu = np.ascontiguousarray(u, np.float32)
    out = np.zeros((len(uu), 9), np.float32)
    if lib().oracle_bsdf_sample(C.byref(m), _p(rd), ior, int(from_light), _p(uu), len(uu), _p(out)) != 0:
        raise RuntimeError("oracle_bsdf_sample failed")
    return out
