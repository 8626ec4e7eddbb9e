"""ctypes mirror of include/upbp_capi.h (plumbing for tests, bench and smoke).

The renderer itself lives in libupbp.so (HIP kernels + C++ host); this module only
declares the C structs and function signatures. Loading fails loudly if the
library has not been built.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libupbp.so")

UPBP_BEAM_SHORT = 1 << 0
UPBP_BEAM_LONG = 1 << 1
UPBP_BPT = 1 << 2
UPBP_SURF = 1 << 3
UPBP_PP3D = 1 << 4
UPBP_PB2D = 1 << 5
UPBP_BB1D = 1 << 6
UPBP_BP2D = 1 << 7
UPBP_PREVIOUS = 1 << 9
UPBP_COMPATIBLE = 1 << 10
UPBP_SPECULAR_ONLY = 1 << 12
UPBP_HELP = 1
UPBP_LIGHT_AREA, UPBP_LIGHT_BACKGROUND, UPBP_LIGHT_DIRECTIONAL, UPBP_LIGHT_POINT = 0, 1, 2, 3


class Config(C.Structure):
    _fields_ = [
        ("algorithm_flags", C.c_uint32), ("iteration_count", C.c_int32), ("max_time", C.c_float),
        ("res_x", C.c_uint32), ("res_y", C.c_uint32), ("max_path_length", C.c_uint32),
        ("min_path_length", C.c_uint32), ("path_count_per_iter", C.c_float),
        ("query_beam_type", C.c_uint32), ("photon_beam_type", C.c_uint32),
        ("bb1d_radius_initial", C.c_float), ("bb1d_radius_alpha", C.c_float),
        ("bb1d_used_light_subpath_count", C.c_float),
        ("bp2d_radius_initial", C.c_float), ("bp2d_radius_alpha", C.c_float),
        ("pb2d_radius_initial", C.c_float), ("pb2d_radius_alpha", C.c_float),
        ("pp3d_radius_initial", C.c_float), ("pp3d_radius_alpha", C.c_float),
        ("surf_radius_initial", C.c_float), ("surf_radius_alpha", C.c_float),
        ("continuous_output", C.c_int32), ("ignore_fully_spec_paths", C.c_int32), ("gpu_assert", C.c_int32),
        ("iteration_offset", C.c_uint64), ("do_log", C.c_int32), ("use_shading_normal", C.c_int32),
        ("scene_id", C.c_int32), ("scene_obj_file", C.c_char * 1024), ("output_name", C.c_char * 1024),
        ("additional_args", C.c_char * 512),
        ("guide_surface_alpha", C.c_float), ("guide_media_alpha", C.c_float),
    ]


class Material(C.Structure):
    _fields_ = [("diffuse", C.c_float * 3), ("phong", C.c_float * 3), ("phong_exponent", C.c_float),
                ("mirror", C.c_float * 3), ("ior", C.c_float), ("priority", C.c_int32), ("real", C.c_int32),
                ("med_idx", C.c_int32)]


class Medium(C.Structure):
    _fields_ = [("absorption", C.c_float * 3), ("emission", C.c_float * 3), ("scattering", C.c_float * 3),
                ("continuation_prob", C.c_float), ("mean_cosine", C.c_float)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("mat_id", C.c_int32), ("med_id", C.c_int32), ("p0", C.c_float * 3),
                ("p1", C.c_float * 3), ("p2", C.c_float * 3), ("position", C.c_float * 3),
                ("intensity", C.c_float * 3), ("scale", C.c_float), ("uses_env_map", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("target", C.c_float * 3), ("roll", C.c_float * 3),
                ("horizontal_fov_deg", C.c_float), ("focal_dist", C.c_float), ("mat_id", C.c_int32),
                ("med_id", C.c_int32)]


class Triangles(C.Structure):
    _fields_ = [("count", C.c_uint32), ("verts", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("mat_id", C.POINTER(C.c_int32)), ("light_id", C.POINTER(C.c_int32))]


class Spheres(C.Structure):
    _fields_ = [("count", C.c_uint32), ("spheres", C.POINTER(C.c_float)), ("mat_id", C.POINTER(C.c_int32)),
                ("light_id", C.POINTER(C.c_int32))]


class EnvMap(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.POINTER(C.c_float))]


class Scene(C.Structure):
    _fields_ = [("real", Triangles), ("imag", Triangles), ("mat_count", C.c_uint32),
                ("materials", C.POINTER(Material)), ("med_count", C.c_uint32), ("media", C.POINTER(Medium)),
                ("light_count", C.c_uint32), ("lights", C.POINTER(Light)), ("global_medium_idx", C.c_uint32),
                ("background_light_idx", C.c_int32), ("camera", Camera), ("scene_center", C.c_float * 3),
                ("scene_radius", C.c_float), ("sph_real", Spheres), ("sph_imag", Spheres), ("env_map", EnvMap)]


class Tuning(C.Structure):
    _fields_ = [("subbeam_factor", C.c_float), ("query_piece_factor", C.c_float),
                ("merge_blocks_per_cu", C.c_int32), ("piece_order", C.c_int32), ("packet_bounces", C.c_int32),
                ("block_size", C.c_int32), ("beam_leaf", C.c_int32), ("graph", C.c_int32),
                ("bundle_bounces", C.c_int32), ("bundle_factor", C.c_float), ("bundle_piece_factor", C.c_float)]


class IterStats(C.Structure):
    _fields_ = [("ms_light", C.c_double), ("ms_build", C.c_double), ("ms_camera", C.c_double),
                ("ms_total", C.c_double), ("ms_bounce", C.c_double), ("ms_merge_pb2d", C.c_double),
                ("ms_merge_bb1d", C.c_double), ("ms_dominant", C.c_double), ("dominant_launches", C.c_uint64),
                ("dominant_kernel", C.c_char * 32),
                ("light_vertices", C.c_uint64), ("surface_vertices", C.c_uint64),
                ("medium_vertices", C.c_uint64), ("beams", C.c_uint64), ("subbeams", C.c_uint64),
                ("light_extend_steps", C.c_uint64), ("camera_extend_steps", C.c_uint64),
                ("bpt_connections", C.c_uint64), ("point_queries", C.c_uint64), ("segment_queries", C.c_uint64),
                ("merge_accepts", C.c_uint64), ("pb2d_accepts", C.c_uint64), ("bb1d_accepts", C.c_uint64),
                ("seg_nodes", C.c_uint64), ("seg_candidates", C.c_uint64), ("seg_pieces", C.c_uint64),
                ("seg_wave_trips", C.c_uint64), ("seg_cand_batches", C.c_uint64), ("seg_pair_batches", C.c_uint64),
                ("vertex_drops", C.c_uint64), ("beam_drops", C.c_uint64), ("path_errors", C.c_uint64),
                ("algorithmic_bytes", C.c_double), ("dominant_bytes", C.c_double),
                ("guide_samples", C.c_uint64), ("guide_loss", C.c_double), ("ms_guide", C.c_double),
                ("merge_stack_overflows", C.c_uint64), ("photon_list_resizes", C.c_uint64),
                ("graph_launches", C.c_uint64), ("bundles", C.c_uint64), ("bundle_pieces_left", C.c_uint64),
                ("bundle_nodes", C.c_uint64), ("bundle_leaves", C.c_uint64), ("bundle_tests", C.c_uint64)]


class GuideConfig(C.Structure):
    _fields_ = [("alpha", C.c_float), ("grid_res", C.c_int32), ("train", C.c_int32),
                ("steps_per_iteration", C.c_int32), ("batch", C.c_int32), ("learning_rate", C.c_float),
                ("seed", C.c_uint32), ("max_samples", C.c_uint32), ("surface_alpha", C.c_float),
                ("history", C.c_uint32)]


_lib = None

# (name, restype, argtypes); declared once so that every call is type checked
_SIGS = [
    ("upbp_parse_commandline", C.c_int, [C.c_int, C.POINTER(C.c_char_p), C.POINTER(Config)]),
    ("upbp_config_defaults", None, [C.POINTER(Config)]),
    ("upbp_scene_load_obj", C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("upbp_scene_load_predefined", C.c_int, [C.c_int32, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("upbp_predefined_scene_count", C.c_int32, []),
    ("upbp_predefined_scene_name", C.c_char_p, [C.c_int32, C.c_int32]),
    ("upbp_scene_get", C.POINTER(Scene), [C.c_void_p]),
    ("upbp_scene_free", None, [C.c_void_p]),
    ("upbp_save_exr", C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    ("upbp_save_bmp", C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float]),
    ("upbp_save_hdr", C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    ("upbp_save_image", C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    ("upbp_memory_info", C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("upbp_generate_scene", C.c_int, [C.c_char_p, C.c_uint32, C.c_char_p]),
    ("upbp_global_error", C.c_char_p, []),
    ("upbp_create", C.c_int, [C.POINTER(Scene), C.POINTER(Config), C.POINTER(C.c_void_p)]),
    ("upbp_run_iteration", C.c_int, [C.c_void_p, C.c_int, C.POINTER(IterStats)]),
    ("upbp_read_accum", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("upbp_destroy", None, [C.c_void_p]),
    ("upbp_last_error", C.c_char_p, [C.c_void_p]),
    ("upbp_sync", C.c_int, [C.c_void_p]),
    ("upbp_test_range_query", C.c_int, [C.c_int, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_float,
                                        C.c_float, C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("upbp_test_range_query_wide", C.c_int, [C.c_int, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_float,
                                             C.c_float, C.c_int, C.c_void_p, C.c_void_p, C.c_uint64,
                                             C.POINTER(C.c_uint64)]),
    ("upbp_test_copy", C.c_int, [C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    ("upbp_tuning_defaults", None, [C.c_void_p]),
    ("upbp_set_tuning", C.c_int, [C.c_void_p, C.c_void_p]),
    ("upbp_autotune", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    ("upbp_run_batch", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]),
    ("upbp_batch_create", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    ("upbp_batch_run", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    ("upbp_batch_counts", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("upbp_batch_last_error", C.c_char_p, [C.c_void_p]),
    ("upbp_batch_destroy", None, [C.c_void_p]),
    ("upbp_test_closest_hit", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("upbp_test_sphere_hits", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("upbp_test_env_samples", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("upbp_test_guide_samples", C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_int32, C.c_void_p, C.c_uint32,
                                          C.c_void_p]),
    ("upbp_mlp_param_count", C.c_uint32, []),
    ("upbp_mlp_create", C.c_int, [C.c_uint32, C.POINTER(C.c_void_p)]),
    ("upbp_mlp_destroy", None, [C.c_void_p]),
    ("upbp_mlp_set_params", C.c_int, [C.c_void_p, C.c_void_p]),
    ("upbp_mlp_get_params", C.c_int, [C.c_void_p, C.c_void_p]),
    ("upbp_mlp_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("upbp_mlp_loss_grad", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.c_void_p]),
    ("upbp_mlp_train_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.POINTER(C.c_float)]),
    ("upbp_guide_defaults", None, [C.c_void_p]),
    ("upbp_set_guide", C.c_int, [C.c_void_p, C.c_void_p]),
    ("upbp_set_guide_lobes", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_float]),
    ("upbp_read_exr", C.c_int, [C.c_char_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("upbp_mlp_forward_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("upbp_mlp_noalloc_max", C.c_uint32, []),
    ("upbp_set_mis_log", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("upbp_read_mis_log", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("upbp_read_guide_samples", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64), C.c_void_p]),
]


def lib():
    """Loads libupbp.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libupbp.so is not built: run python -m gpuupbp_amd.build")
        h = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            if hasattr(h, name):
                f = getattr(h, name)
                f.restype = res
                f.argtypes = args
        _lib = h
    return _lib


def exported_symbols():
    return [s[0] for s in _SIGS]


def parse_args(argv):
    """Runs the reference-compatible command-line parser; returns (rc, Config)."""
    cfg = Config()
    arr = (C.c_char_p * (len(argv) + 1))(*([b"gpuupbp"] + [a.encode() for a in argv]))
    rc = lib().upbp_parse_commandline(len(argv) + 1, arr, C.byref(cfg))
    return rc, cfg


class SceneHandle:
    """Owns a scene loaded by the C++ OBJ/MTL/AUX loader, or a predefined scene when path is an
    int (the -s id)."""

    def __init__(self, path, res_x=256, res_y=256):
        h = C.c_void_p()
        if isinstance(path, int):
            rc = lib().upbp_scene_load_predefined(path, res_x, res_y, C.byref(h))
        else:
            rc = lib().upbp_scene_load_obj(path.encode(), res_x, res_y, C.byref(h))
        if rc != 0:
            raise RuntimeError("scene load failed (%d): %s" % (rc, lib().upbp_global_error().decode()))
        self.h = h
        self.ptr = lib().upbp_scene_get(h)

    @property
    def scene(self):
        return self.ptr.contents

    def close(self):
        if self.h:
            lib().upbp_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_exr(path):
    """(H, W, 3) float32 image of an OpenEXR file (the environment-map reader)."""
    w, h = C.c_uint32(), C.c_uint32()
    rc = lib().upbp_read_exr(path.encode(), None, 0, C.byref(w), C.byref(h))
    if rc != 0:
        raise RuntimeError("upbp_read_exr failed (%d)" % rc)
    import numpy as np
    img = np.zeros((h.value, w.value, 3), np.float32)
    rc = lib().upbp_read_exr(path.encode(), img.ctypes.data, img.size, C.byref(w), C.byref(h))
    if rc != 0:
        raise RuntimeError("upbp_read_exr failed (%d)" % rc)
    return img


def predefined_scenes():
    """(short name, long name) of every predefined scene, indexed by -s id."""
    n = lib().upbp_predefined_scene_count()
    return [(lib().upbp_predefined_scene_name(i, 0).decode(), lib().upbp_predefined_scene_name(i, 1).decode())
            for i in range(n)]


def generate_scene(kind, seed, obj_path):
    rc = lib().upbp_generate_scene(kind.encode(), seed, obj_path.encode())
    if rc != 0:
        raise RuntimeError("scene generation failed (%d)" % rc)
    return obj_path
