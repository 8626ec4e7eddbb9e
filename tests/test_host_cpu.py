"""Host-side boundary tests (no GPU): command-line parser, OBJ/MTL/AUX loader, EXR writer and
the C ABI exports of libupbp.so."""
import os
import re
import struct
import subprocess
import zipfile

import numpy as np
import pytest

from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SCENES = "/root/reference/scenes"


def parse(*args):
    rc, cfg = capi.parse_args(list(args))
    return rc, cfg


# ---- command line: gpuupbp/misc/config.cpp:735-1370 ------------------------------------
def test_defaults():
    rc, c = parse("-s", "-1", "x/scene.obj")
    assert rc == 0
    assert c.algorithm_flags == capi.UPBP_BPT | capi.UPBP_SURF | capi.UPBP_PP3D | capi.UPBP_PB2D | capi.UPBP_BB1D
    assert (c.res_x, c.res_y, c.max_path_length, c.min_path_length, c.iteration_count) == (256, 256, 10, 0, 1)
    assert c.path_count_per_iter == 256 * 256
    assert c.query_beam_type == capi.UPBP_BEAM_LONG and c.photon_beam_type == capi.UPBP_BEAM_SHORT
    assert abs(c.surf_radius_initial + 0.0015) < 1e-9 and abs(c.surf_radius_alpha - 0.75) < 1e-9
    for t in ("pp3d", "pb2d", "bb1d", "bp2d"):
        assert abs(getattr(c, t + "_radius_initial") + 0.001) < 1e-9 and getattr(c, t + "_radius_alpha") == 1.0
    assert c.use_shading_normal == 1 and c.bb1d_used_light_subpath_count == -1.0
    assert c.output_name == b"s-scene_i01_l10_a+bpt+surf+pp3d+pb2d+bb1d_qbtL_pbtS.exr"


def test_algorithm_selection():
    assert parse("-s", "-1", "a.obj", "-a", "upbp_bp2d")[1].algorithm_flags == capi.UPBP_BP2D
    # "upbp+" adds bpt and surf to the listed techniques
    assert parse("-s", "-1", "a.obj", "-a", "upbp+pp3d")[1].algorithm_flags == (
        capi.UPBP_BPT | capi.UPBP_SURF | capi.UPBP_PP3D)
    assert parse("-s", "-1", "a.obj", "-a", "upbp_foo")[0] < 0
    rc, c = parse("-s", "-1", "a.obj", "-a", "upbp_pb2d+bb1d", "-previous", "-qbt", "S", "-pbt", "L")
    assert c.algorithm_flags == capi.UPBP_PB2D | capi.UPBP_BB1D | capi.UPBP_PREVIOUS
    assert c.output_name == b"s-a_i01_l10_a+pb2d+bb1d_qbtS_pbtL_prev.exr"


def test_radius_option_priorities():
    # a technique-specific option wins over the shared one regardless of order (config.cpp:934-1083)
    rc, c = parse("-s", "-1", "a.obj", "-r_initial_surf", "0.5", "-r_initial", "0.2", "-r_alpha", "0.9",
                  "-r_alpha_bb1d", "0.6")
    assert rc == 0
    assert c.surf_radius_initial == pytest.approx(0.5) and c.pp3d_radius_initial == pytest.approx(0.2)
    assert c.bb1d_radius_alpha == pytest.approx(0.6) and c.surf_radius_alpha == pytest.approx(0.9)
    assert parse("-s", "-1", "a.obj", "-r_initial", "0")[0] < 0
    assert parse("-s", "-1", "a.obj", "-r_alpha", "-1")[0] < 0


def test_path_counts_and_resolution():
    rc, c = parse("-s", "-1", "a.obj", "-r", "320x200", "-pcpi", "-0.5")
    assert (c.res_x, c.res_y) == (320, 200) and c.path_count_per_iter == 32000
    rc, c = parse("-s", "-1", "a.obj", "-pcpi", "1234.7")
    assert c.path_count_per_iter == 1234
    assert parse("-s", "-1", "a.obj", "-pcpi", "0.2")[0] < 0
    assert parse("-s", "-1", "a.obj", "-r", "0x10")[0] < 0


def test_output_name_rules():
    assert parse("-s", "-1", "a.obj", "-o", "img")[1].output_name == b"img.exr"
    assert parse("-s", "-1", "a.obj", "-o", "img.bmp")[1].output_name == b"img.bmp"
    rc, c = parse("-s", "-1", "dir/a.obj", "-i", "7", "-l", "5", "-r", "64x32", "-t", "3", "-pbc", "100", "-log")
    # extra-argument suffixes follow the command-line order (config.cpp:889,932)
    assert c.output_name == b"s-a_i07_l05_a+bpt+surf+pp3d+pb2d+bb1d_qbtL_pbtS_r64x32_t3_pbc100_log.exr"


def test_errors_and_help(capfd):
    assert parse("-s", "3")[0] < 0
    assert parse("-l", "0")[0] < 0
    assert parse("-ignorespec", "2")[0] < 0
    assert parse("-nosuchoption")[0] < 0
    err = capfd.readouterr().err
    assert "invalid option \"-nosuchoption\", please see help (-hf)" in err
    assert parse("-h")[0] == capi.UPBP_HELP


# ---- scene loader: gpuupbp/misc/scene_loader_obj.cpp -------------------------------------
def test_box_scene_loads(box_obj):
    sh = capi.SceneHandle(box_obj, 64, 64)
    s = sh.scene
    assert s.real.count == 12 and s.imag.count == 1280
    # two emissive triangles -> two area lights after the (zero) AUX lights
    assert s.light_count == 2 and all(s.lights[i].type == 0 for i in range(2))
    assert abs(s.scene_radius - np.sqrt(12) / 2) < 1e-5
    # sphere medium: continuation = max(albedo, 0.8) = max(0.5/0.6, 0.8) (scene_loader_obj.cpp:1075-1088)
    med = [s.media[i] for i in range(s.med_count)]
    assert s.med_count == 2 and s.global_medium_idx == 1
    assert med[0].continuation_prob == pytest.approx(0.5 / 0.6, rel=1e-6) and med[1].continuation_prob == 1.0
    # camera from the AUX transform rows
    cam = s.camera
    assert list(cam.origin) == pytest.approx([0, 0, 3.2]) and cam.horizontal_fov_deg == pytest.approx(40.0, rel=1e-5)
    mats = {s.materials[i].priority for i in range(s.mat_count)}
    assert 1 in mats


def test_loader_errors(tmp_path):
    h = __import__("ctypes").c_void_p()
    assert capi.lib().upbp_scene_load_obj(str(tmp_path / "missing.obj").encode(), 8, 8, __import__("ctypes").byref(h)) \
        == -2
    (tmp_path / "m.mtl").write_text("newmtl a\nKd 1 1 1\n")
    (tmp_path / "s.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl b\nf 1 2 3\n")
    (tmp_path / "s.obj.aux").write_text("")
    with pytest.raises(RuntimeError, match="material not found"):
        capi.SceneHandle(str(tmp_path / "s.obj"))
    # a material library that names a directory (a directory opens for reading on Linux, and
    # its size reads as 2^63 - 1) is an error code, not an exception through the C ABI
    (tmp_path / "sub").mkdir()
    (tmp_path / "d.obj").write_text("mtllib sub\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    (tmp_path / "d.obj.aux").write_text("")
    with pytest.raises(RuntimeError, match="could not open"):
        capi.SceneHandle(str(tmp_path / "d.obj"))


def test_polygon_fan_and_normals(tmp_path):
    (tmp_path / "m.mtl").write_text("newmtl a\nKd 0.5 0.5 0.5\nnewmtl l\nKe 1 2 3\n")
    (tmp_path / "s.obj").write_text(
        "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 -1\nusemtl a\nf 1//1 2//1 3//1 4//1\n"
        "usemtl l\nf 1 2 3\n")
    (tmp_path / "s.obj.aux").write_text("light_point 1 2 3 4 5 6\nmaterial a\nmirror 0.5 0.5 0.5\nior 1.5\n")
    sh = capi.SceneHandle(str(tmp_path / "s.obj"))
    s = sh.scene
    assert s.real.count == 3
    nrm = np.ctypeslib.as_array(s.real.normals, (3 * 12,)).reshape(3, 4, 3)
    # face normal flipped to agree with the vertex normals (scene_loader_obj.cpp:661-667)
    assert np.allclose(nrm[0, 3], [0, 0, -1]) and np.allclose(nrm[2, 3], [0, 0, 1])
    assert s.light_count == 2 and s.lights[0].type == 3 and s.lights[1].type == 0
    assert list(s.lights[1].intensity) == [1, 2, 3]
    assert s.materials[0].ior == pytest.approx(1.5) and list(s.materials[0].mirror) == pytest.approx([0.5] * 3)


@pytest.mark.skipif(not os.path.exists(REF_SCENES + "/bunny/bunny.zip"), reason="reference scenes not mounted")
def test_reference_bunny_scene(tmp_path):
    # the reference's own scene files load: 4 media plus an appended clear global medium (the AUX
    # file names no globalMediumId), Ke light enclosed by matGlobal
    for f in ("bunny.mtl", "bunny.obj.aux"):
        (tmp_path / f).write_bytes(open(os.path.join(REF_SCENES, "bunny", f), "rb").read())
    with zipfile.ZipFile(REF_SCENES + "/bunny/bunny.zip") as z:
        z.extract("bunny.obj", str(tmp_path))
    sh = capi.SceneHandle(str(tmp_path / "bunny.obj"), 128, 128)
    s = sh.scene
    assert s.med_count == 5 and s.global_medium_idx == 4 and s.real.count > 10000 and s.light_count > 0
    # "scaterring" in the reference file still sets the scattering coefficient
    assert list(s.media[0].scattering) == pytest.approx([0.06, 0.1, 0.1])
    assert s.camera.horizontal_fov_deg == pytest.approx(0.2 * 57.295779, rel=1e-5)


# ---- EXR writer: gpuupbp/misc/tinyexr_wrapper.cpp:34-95 ----------------------------------
def _read_exr(path):
    b = open(path, "rb").read()
    assert b[:4] == b"\x76\x2f\x31\x01" and b[4] == 2
    pos, attrs = 8, {}
    while b[pos] != 0:
        e = b.index(b"\0", pos)
        name = b[pos:e].decode()
        e2 = b.index(b"\0", e + 1)
        typ = b[e + 1:e2].decode()
        size = struct.unpack("<i", b[e2 + 1:e2 + 5])[0]
        attrs[name] = (typ, b[e2 + 5:e2 + 5 + size])
        pos = e2 + 5 + size
    pos += 1
    xmin, ymin, xmax, ymax = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = xmax - xmin + 1, ymax - ymin + 1
    offs = struct.unpack("<%dQ" % h, b[pos:pos + 8 * h])
    img = np.zeros((h, w, 3), np.float32)
    for y in range(h):
        o = offs[y]
        yy, size = struct.unpack("<2i", b[o:o + 8])
        assert yy == y and size == 6 * w
        px = np.frombuffer(b[o + 8:o + 8 + size], np.float16).reshape(3, w).astype(np.float32)
        img[y, :, 2], img[y, :, 1], img[y, :, 0] = px[0], px[1], px[2]  # stored B, G, R
    return attrs, img


def test_exr_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    img = (rng.uniform(0, 4, (5, 7, 3)) * (rng.uniform(0, 1, (5, 7, 3)) > 0.2)).astype(np.float32)
    img[0, 0] = [70000.0, 1e-7, 0.333333]
    p = str(tmp_path / "o.exr")
    gpu.save_exr(p, img)
    attrs, back = _read_exr(p)
    assert attrs["compression"] == ("compression", b"\x00")
    ch = attrs["channels"][1]
    assert ch.startswith(b"B\x00") and b"G\x00" in ch and b"R\x00" in ch
    assert struct.unpack("<i", ch[2:6])[0] == 1  # HALF
    ref = img.astype(np.float16).astype(np.float32)  # IEEE round-to-nearest-even half
    assert np.array_equal(back, ref)


def test_bmp_writer(tmp_path):
    # 24-bit BI_RGB, bottom-up rows padded to 4 bytes, B,G,R, gamma 2.2, clamped
    # (framebuffer.cpp:129-200); width 5 needs one padding byte per row
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1.5, (3, 5, 3)).astype(np.float32)
    p = str(tmp_path / "o.bmp")
    gpu.save_image(p, img)
    b = open(p, "rb").read()
    assert b[:2] == b"BM"
    size, _, off, hs, w, h, planes, bpp, comp, data = struct.unpack("<IIIIiiHHII", b[2:38])
    assert (size, off, hs, w, h, planes, bpp, comp, data) == (54 + 16 * 3, 54, 40, 5, 3, 1, 24, 0, 16 * 3)
    ref = np.clip(np.power(img, 1 / 2.2) * 255.0, 0, 255).astype(np.uint8)
    for y in range(3):
        row = np.frombuffer(b[54 + 16 * y:54 + 16 * y + 15], np.uint8).reshape(5, 3)
        assert np.array_equal(row[:, ::-1], ref[2 - y])  # bottom-up, B,G,R
        assert b[54 + 16 * y + 15] == 0


def test_hdr_writer(tmp_path):
    # Radiance RGBE: shared exponent, mantissas in [128, 256) for the largest channel
    rng = np.random.default_rng(4)
    img = (rng.uniform(0, 8, (4, 6, 3)) * (rng.uniform(0, 1, (4, 6, 3)) > 0.1)).astype(np.float32)
    img[0, 0] = 0.0
    p = str(tmp_path / "o.hdr")
    gpu.save_image(p, img)
    b = open(p, "rb").read()
    head = b"#?RADIANCE\n"
    assert b.startswith(head) and b"FORMAT=32-bit_rle_rgbe\n\n-Y 4 +X 6\n" in b
    px = np.frombuffer(b[b.index(b"+X 6\n") + 5:], np.uint8).reshape(4, 6, 4).astype(np.float64)
    dec = px[..., :3] * np.ldexp(1.0, (px[..., 3:4] - 136).astype(int))
    dec[px[..., 3] == 0] = 0.0
    big = img.max(axis=2) > 1e-3
    assert np.all(np.abs(dec - img)[big] <= img.max(axis=2)[big][:, None] / 128.0 + 1e-6)
    assert np.all(px[0, 0] == 0)
    with pytest.raises(RuntimeError):
        gpu.save_image(str(tmp_path / "o.png"), img)


# ---- C ABI ---------------------------------------------------------------------------------
def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "upbp_capi.h")).read()
    declared = set(re.findall(r"\b(upbp_[a-z_0-9]+)\s*\(", hdr))
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (upbp_[a-z_0-9]+)", out))
    assert declared and declared <= exported, declared - exported
    lib = capi.lib()
    for name in declared:
        assert hasattr(lib, name)


def test_no_gpu_reports_error():
    # on a machine without a GPU the renderer fails loudly instead of falling back to the CPU
    import ctypes as C
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    cfg = capi.Config()
    capi.lib().upbp_config_defaults(C.byref(cfg))
    out = C.c_void_p()
    assert capi.lib().upbp_test_range_query(0, None, 0, None, 0, 0.1, 0.0, (C.c_uint32 * 1)(), None, 0, None) < 0


def test_cli_help_and_no_gpu_error(box_obj):
    import subprocess
    exe = os.path.join(ROOT, "gpuupbp_amd", "gpuupbp")
    if not os.path.exists(exe):
        pytest.skip("CLI not built")
    h = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=60)
    assert h.returncode == 0 and "Usage:" in h.stdout
    if not os.path.exists("/dev/kfd"):
        r = subprocess.run([exe, "-s", "-1", box_obj, "-r", "16x16", "-i", "1"], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 1 and "no GPU visible" in r.stderr


def test_batch_autotune_and_guide_reject_bad_arguments():
    # argument checks come before any GPU call, so they hold on a machine without one
    import ctypes as C
    E_ARG = -1  # UPBP_E_ARG
    lib = capi.lib()
    its = (C.c_int32 * 2)(0, 1)
    assert lib.upbp_run_batch(None, its, 2, 4) == E_ARG
    assert lib.upbp_run_batch((C.c_void_p * 2)(None, None), its, 2, 4) == E_ARG
    assert lib.upbp_run_batch(None, None, 0, 0) == E_ARG  # no threads
    assert lib.upbp_run_batch(None, None, -1, 1) == E_ARG
    # the same renderer twice (the handles are never dereferenced: the check comes first)
    dup = (C.c_void_p * 3)(0x1000, 0x2000, 0x1000)
    assert lib.upbp_run_batch(dup, (C.c_int32 * 3)(0, 0, 1), 3, 2) == E_ARG
    assert b"twice" in lib.upbp_last_error(None)
    # batched rounds: the same checks before any stream or device memory is made
    out = C.c_void_p()
    assert lib.upbp_batch_create(None, 2, C.byref(out)) == E_ARG
    assert lib.upbp_batch_create((C.c_void_p * 2)(None, None), 2, C.byref(out)) == E_ARG
    assert lib.upbp_batch_create((C.c_void_p * 1)(0x1000), 0, C.byref(out)) == E_ARG
    assert lib.upbp_batch_create((C.c_void_p * 1)(0x1000), 1, None) == E_ARG
    assert lib.upbp_batch_create(dup, 3, C.byref(out)) == E_ARG and not out.value
    assert b"twice" in lib.upbp_batch_last_error(None)
    assert lib.upbp_batch_run(None, its, 1) == E_ARG
    lib.upbp_batch_destroy(None)
    assert lib.upbp_autotune(None, 0, 1, None, None) == E_ARG
    g = capi.GuideConfig()
    lib.upbp_guide_defaults(C.byref(g))
    assert (g.alpha, g.surface_alpha, g.grid_res, g.steps_per_iteration) == (0.0, pytest.approx(0.65), 64, 10)
    assert lib.upbp_set_guide(None, C.byref(g)) == E_ARG


def test_guide_options():
    rc, cfg = capi.parse_args(["-s", "-1", "x.obj", "-guide", "0.65", "-guide_media", "0.2"])
    assert rc == 0 and cfg.guide_surface_alpha == pytest.approx(0.65) and cfg.guide_media_alpha == pytest.approx(0.2)
    assert b"_guide0.65_guidem0.2" in cfg.output_name
    rc, cfg = capi.parse_args(["-s", "-1", "x.obj"])
    assert rc == 0 and cfg.guide_surface_alpha == 0.0 and cfg.guide_media_alpha == 0.0
    for bad in ("2", "-0.1", "abc"):
        rc, _ = capi.parse_args(["-s", "-1", "x.obj", "-guide", bad])
        assert rc != 0
