"""Environment-map background light, host side (no GPU): the OpenEXR reader, the AUX
light_background_em path (gpuupbp/misc/scene_loader_obj.cpp:994-1013,
gpuupbp/misc/tinyexr_wrapper.cpp:97-136) and the oracle's luminance distribution
(gpuupbp/host/env_map.hpp:42-144, gpuupbp/kernel/env_map.cuh:36-137), pinned by what the
mathematics fixes: the reported pdf must be the density the samples follow (E[1/pdf] = 4 pi and
E[f/pdf] = the quadrature of f), a constant map gives the closed-form pdf, and sampling and lookup
agree on pdf and radiance.
"""
import os
import struct
import zlib

import numpy as np
import pytest

from conftest import env_scene, sky
from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu
from oracle import oracle_py


# ---- OpenEXR files written from the published layout -----------------------------------------
def _attr(name, typ, data):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data


def _predict_split(raw):
    # even bytes first, then odd bytes, then byte-wise deltas + 128 (the RLE / ZIP pre-pass)
    t = bytearray(raw[0::2] + raw[1::2])
    out = bytearray(t)
    for i in range(len(t) - 1, 0, -1):
        out[i] = (t[i] - t[i - 1] + 128) & 255
    return bytes(out)


def _rle(data):
    out, i, n = bytearray(), 0, len(data)
    while i < n:
        j = i
        while j + 1 < n and data[j + 1] == data[i] and j - i < 127:
            j += 1
        if j > i + 1:  # a run of >= 3 bytes
            out += bytes([j - i, data[i]])
            i = j + 1
        else:
            k = i
            while k < n and k - i < 127 and not (k + 2 < n and data[k] == data[k + 1] == data[k + 2]):
                k += 1
            out += struct.pack("b", -(k - i)) + bytes(data[i:k])
            i = k
    return bytes(out)


def write_exr(path, img, compression, pixel=2, channels="BGR"):
    """Single-part scanline OpenEXR of an (H, W, 3) image; pixel 1 HALF, 2 FLOAT."""
    h, w, _ = img.shape
    names = sorted(channels)
    chl = b"".join(c.encode() + b"\0" + struct.pack("<iB3xii", pixel, 0, 1, 1) for c in names) + b"\0"
    hdr = struct.pack("<II", 20000630, 2)
    hdr += _attr("channels", "chlist", chl)
    hdr += _attr("compression", "compression", bytes([compression]))
    hdr += _attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    hdr += _attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    hdr += _attr("lineOrder", "lineOrder", b"\0")
    hdr += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    hdr += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
    hdr += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    hdr += b"\0"
    lines = {0: 1, 1: 1, 2: 1, 3: 16}[compression]
    idx = {"R": 0, "G": 1, "B": 2, "Y": 0}
    chunks = []
    for y0 in range(0, h, lines):
        raw = b""
        for y in range(y0, min(h, y0 + lines)):
            for c in names:
                v = img[y, :, idx[c]]
                raw += (v.astype(np.float16) if pixel == 1 else v.astype(np.float32)).tobytes()
        if compression == 0:
            data = raw
        elif compression == 1:
            data = _rle(_predict_split(raw))
        else:
            data = zlib.compress(_predict_split(raw))
        if len(data) >= len(raw):
            data = raw
        chunks.append(struct.pack("<ii", y0, len(data)) + data)
    off = len(hdr) + 8 * len(chunks)
    table = b""
    for ch in chunks:
        table += struct.pack("<Q", off)
        off += len(ch)
    with open(path, "wb") as f:
        f.write(hdr + table + b"".join(chunks))


def _test_image(h=12, w=24, seed=0):
    rng = np.random.default_rng(seed)
    img = rng.uniform(0, 4, (h, w, 3)).astype(np.float32)
    img[:, 5:9] = 1.5  # flat runs so that RLE has something to do
    return img


@pytest.mark.parametrize("compression", [0, 1, 2, 3])
@pytest.mark.parametrize("pixel", [1, 2])
def test_exr_reader_compressions(tmp_path, compression, pixel):
    img = _test_image()
    p = str(tmp_path / "e.exr")
    write_exr(p, img, compression, pixel)
    got = capi.read_exr(p)
    want = img.astype(np.float16).astype(np.float32) if pixel == 1 else img
    assert got.shape == img.shape and np.array_equal(got, want)


def test_exr_reader_reads_the_writer_and_rejects_bad_files(tmp_path):
    img = _test_image(8, 16, 1)
    p = str(tmp_path / "w.exr")
    gpu.save_exr(p, img)  # the renderer's own HALF writer
    assert np.array_equal(capi.read_exr(p), img.astype(np.float16).astype(np.float32))
    grey = str(tmp_path / "y.exr")
    write_exr(grey, img, 3, 2, channels="Y")
    g = capi.read_exr(grey)
    assert np.array_equal(g[..., 0], g[..., 1]) and np.array_equal(g[..., 0], img[..., 0])
    bad = str(tmp_path / "bad.exr")
    with open(bad, "wb") as f:
        f.write(b"not an exr file at all")
    with pytest.raises(RuntimeError):
        capi.read_exr(bad)
    with pytest.raises(RuntimeError):
        capi.read_exr(str(tmp_path / "missing.exr"))
    piz = str(tmp_path / "piz.exr")
    write_exr(piz, img, 0)
    data = bytearray(open(piz, "rb").read())
    data[data.index(b"compression\0compression\0") + 28] = 4  # PIZ: not read
    open(piz, "wb").write(bytes(data))
    with pytest.raises(RuntimeError):
        capi.read_exr(piz)


@pytest.mark.parametrize("offset", [2 ** 64 - 4, 2 ** 63, 10 ** 9, -8])
def test_exr_reader_rejects_corrupt_offset_tables(tmp_path, offset):
    # a chunk offset near 2^64 would wrap an unchecked "offset + size" bounds test; every entry
    # must be rejected before a chunk is read (offset -8: eight bytes before the end of the file)
    img = _test_image(4, 8, 2)
    p = str(tmp_path / "o.exr")
    write_exr(p, img, 0)
    data = bytearray(open(p, "rb").read())
    table = len(data) - sum(len(img[0]) * 3 * 4 + 8 for _ in range(4)) - 8 * 4
    assert struct.unpack("<Q", bytes(data[table:table + 8]))[0] == table + 8 * 4  # first chunk follows the table
    bad = offset if offset >= 0 else len(data) - 4
    data[table + 8:table + 16] = struct.pack("<Q", bad)
    open(p, "wb").write(bytes(data))
    with pytest.raises(RuntimeError):
        capi.read_exr(p)


def test_exr_reader_rejects_an_empty_compression_attribute(tmp_path):
    img = _test_image(4, 8, 3)
    p = str(tmp_path / "c.exr")
    write_exr(p, img, 0)
    data = bytearray(open(p, "rb").read())
    at = data.index(b"compression\0compression\0") + len(b"compression\0compression\0")
    assert struct.unpack("<i", bytes(data[at:at + 4]))[0] == 1
    data[at:at + 5] = struct.pack("<i", 0)  # size 0, value byte dropped
    open(p, "wb").write(bytes(data))
    with pytest.raises(RuntimeError):
        capi.read_exr(p)


def test_exr_reader_rejects_a_data_window_the_file_cannot_hold(tmp_path):
    # a 65536 x 65536 ZIP data window whose 4096 offsets all name the one real chunk: every
    # offset lies in the file, so only the size bound (deflate expands at most ~1032:1) keeps the
    # reader from allocating 48 GiB of pixels
    img = _test_image(4, 8, 4)
    p = str(tmp_path / "w.exr")
    write_exr(p, img, 3)
    data = bytes(open(p, "rb").read())
    table = next(i for i in range(len(data) - 8) if struct.unpack("<Q", data[i:i + 8])[0] == i + 8)
    hdr = bytearray(data[:table])
    dw = hdr.index(b"dataWindow\0box2i\0") + len(b"dataWindow\0box2i\0") + 4
    hdr[dw:dw + 16] = struct.pack("<iiii", 0, 0, 65535, 65535)
    chunks = 65536 // 16
    first = len(hdr) + 8 * chunks
    open(p, "wb").write(bytes(hdr) + struct.pack("<Q", first) * chunks + data[table + 8:])
    with pytest.raises(RuntimeError):
        capi.read_exr(p)


# ---- AUX light_background_em ------------------------------------------------------------------
def test_aux_environment_map_loads_rotated_and_scaled(box_obj, tmp_path):
    env = sky()
    obj = env_scene(box_obj, str(tmp_path / "env"), env, scale=2.0, rotate=0.25)
    sh = capi.SceneHandle(obj, 32, 32)
    s = sh.scene
    assert s.env_map.width == 32 and s.env_map.height == 16
    got = np.ctypeslib.as_array(s.env_map.rgb, (16 * 32 * 3,)).reshape(16, 32, 3)
    src = env.astype(np.float16).astype(np.float32) * 2.0
    assert np.array_equal(got, np.roll(src, 8, axis=1))  # column x moves to x + 0.25 * 32
    bg = [s.lights[i] for i in range(s.light_count) if s.lights[i].type == capi.UPBP_LIGHT_BACKGROUND]
    assert len(bg) == 1 and bg[0].uses_env_map == 1 and s.background_light_idx >= 0


def test_aux_environment_map_errors(box_obj, tmp_path):
    obj = env_scene(box_obj, str(tmp_path / "narrow"), np.ones((16, 16, 3), np.float32))
    with pytest.raises(RuntimeError):
        capi.SceneHandle(obj, 32, 32)  # must be twice as wide as high
    obj = env_scene(box_obj, str(tmp_path / "gone"), sky())
    os.remove(os.path.join(os.path.dirname(obj), "env.exr"))
    with pytest.raises(RuntimeError):
        capi.SceneHandle(obj, 32, 32)


# ---- oracle distribution pins ---------------------------------------------------------------------
def _grid(n):
    k = int(np.sqrt(n))
    a = (np.arange(k) + 0.5) / k
    u0, u1 = np.meshgrid(a, a)
    return np.stack([u0.ravel(), u1.ravel()], 1).astype(np.float32)


def test_oracle_constant_map_has_the_closed_form_pdf():
    h = 10
    out = oracle_py.env_samples(np.full((h, 2 * h, 3), 0.7, np.float32), _grid(400))
    mean_sin = np.mean(np.sin(np.pi * (np.arange(h) + 0.5) / h))
    want = 1.0 / (2 * np.pi ** 2 * mean_sin)  # sin(theta) cancels against the row weight
    assert np.allclose(out[:, 3], want, rtol=2e-6) and np.allclose(out[:, 7], want, rtol=2e-6)
    assert np.allclose(out[:, 4:7], 0.7, rtol=1e-6)
    assert np.allclose(np.linalg.norm(out[:, :3], axis=1), 1, atol=1e-6)


def test_oracle_sampling_density_matches_reported_pdf():
    env = sky()
    u = _grid(256 * 256)
    out = oracle_py.env_samples(env, u)
    pdf = out[:, 3].astype(np.float64)
    assert (pdf > 0).all()
    # E[1 / pdf] is the sphere's solid angle only if the samples follow the reported density
    assert abs(np.mean(1 / pdf) - 4 * np.pi) / (4 * np.pi) < 2e-3
    # E[f / pdf] against the integral of f(w) = (1 + z)^2 over the sphere, 16 pi / 3
    z = out[:, 2].astype(np.float64)
    est = np.mean((1 + z) ** 2 / pdf)
    assert abs(est - 16 * np.pi / 3) / (16 * np.pi / 3) < 3e-3
    # the sun patch (6 of 512 pixels) draws about half of the samples, as its luminance says
    assert 0.3 < np.mean(out[:, 4] > 20) < 0.95


def test_oracle_sample_and_lookup_agree():
    env = sky()
    out = oracle_py.env_samples(env, np.random.default_rng(4).uniform(0, 1, (5000, 2)).astype(np.float32))
    # away from pixel-row boundaries the lookup lands in the sampled cell
    same = np.abs(out[:, 3] - out[:, 7]) <= 1e-4 * out[:, 3]
    assert same.mean() > 0.97
    assert np.allclose(out[same, 4:7], out[same, 8:11], rtol=1e-3, atol=1e-4)
