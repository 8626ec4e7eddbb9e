"""Predefined scenes and the analytic sphere primitive, host side (no GPU).

The scene table and builder follow gpuupbp/misc/config.cpp:100-518 and
gpuupbp/misc/scene_loader_predef.cpp:126-603; the expected counts and positions below are worked
out by hand from those rules (box corners, sphere placement formulas, material keying by
priority). The oracle's sphere test (gpuupbp/kernel/intersector.cu:35-95 restated) is pinned
against the closed-form roots in float64.
"""
import math

import numpy as np
import pytest

from gpuupbp_amd import capi
from oracle import oracle_py

# box corners of scene_loader_predef.cpp:296-303
CB = np.array([[-1.27029, 1.30455, -1.28002], [1.28975, 1.30455, -1.28002], [1.28975, 1.30455, 1.28002],
               [-1.27029, 1.30455, 1.28002], [-1.27029, -1.25549, -1.28002], [1.28975, -1.25549, -1.28002],
               [1.28975, -1.25549, 1.28002], [-1.27029, -1.25549, 1.28002]], np.float32)


def load(i):
    h = capi.SceneHandle(i, 64, 64)
    return h, h.scene


def spheres(s, real):
    S = s.sph_real if real else s.sph_imag
    return np.ctypeslib.as_array(S.spheres, (S.count * 4,)).reshape(-1, 4) if S.count else np.zeros((0, 4))


def sphere_mats(s, real):
    S = s.sph_real if real else s.sph_imag
    return [S.mat_id[i] for i in range(S.count)]


def test_table_names_and_ids():
    t = capi.predefined_scenes()
    assert len(t) == 41
    assert t[0] == ("abssph", "large absorbing sphere + big area light")
    assert t[21][0] == "overgwi" and t[31][0] == "backg" and t[35][0] == "smallsb"
    assert t[39][0] == t[40][0] == "glasssphbck"  # the reference repeats the short name
    with pytest.raises(RuntimeError):
        capi.SceneHandle(41, 64, 64)


# (real tris, imag tris, real spheres, imag spheres, materials, lights) per scene id, from the
# builder's rules: box walls 2 triangles each, the big light replaces the ceiling, the small
# light adds a 4-wall housing (8) and its emitting floor (2), the distant one only the floor
EXPECTED = {
    0: (10, 0, 0, 1, 11, 2), 1: (20, 0, 0, 1, 11, 2), 2: (20, 0, 1, 0, 11, 2), 5: (10, 0, 2, 0, 12, 1),
    7: (10, 0, 2, 0, 12, 1), 11: (2, 0, 0, 0, 10, 2), 13: (0, 0, 0, 0, 10, 1), 17: (2, 0, 0, 0, 10, 2),
    21: (10, 0, 3, 0, 13, 2), 22: (10, 0, 3, 1, 14, 2), 23: (10, 0, 4, 0, 13, 2), 26: (10, 0, 4, 0, 14, 2),
    29: (2, 0, 0, 3, 13, 2), 30: (10, 0, 0, 4, 13, 1), 31: (0, 0, 0, 0, 10, 1), 35: (2, 12, 0, 0, 11, 2),
    39: (10, 0, 1, 0, 11, 1), 40: (10, 0, 1, 0, 11, 2),
}


@pytest.mark.parametrize("sid", sorted(EXPECTED))
def test_scene_structure(sid):
    _, s = load(sid)
    got = (s.real.count, s.imag.count, s.sph_real.count, s.sph_imag.count, s.mat_count, s.light_count)
    assert got == EXPECTED[sid]


def test_sphere_placement_and_materials():
    floor_c = (CB[0] + CB[1] + CB[4] + CB[5]) / 4
    box_c = CB.mean(axis=0)
    _, s = load(0)  # large imaginary red-absorbing sphere
    sp = spheres(s, False)
    assert np.allclose(sp[0, :3], floor_c + [0, 0, 0.81], atol=1e-6) and sp[0, 3] == np.float32(0.8)
    m = s.materials[sphere_mats(s, False)[0]]
    assert (m.priority, m.real, m.med_idx) == (8, 0, 7)  # (1 + -1) * 10 + 1 + 7
    _, s = load(21)  # glass, water (+ water medium), ice: bottom-left, bottom-right, top
    sp = spheres(s, True)
    ov = 0.25
    assert np.allclose(sp[:, :3], box_c + np.array([[-ov, 0, -ov], [ov, 0, -ov], [0, 0, ov]]), atol=1e-6)
    mats = [s.materials[i] for i in sphere_mats(s, True)]
    assert [m.priority for m in mats] == [30, 12, 20]
    assert mats[0].ior == pytest.approx(1.6) and mats[1].ior == pytest.approx(1.33) and mats[1].med_idx == 1
    assert all(m.real == 1 for m in mats)
    _, s = load(23)  # the big water sphere shares the small water sphere's material (same priority)
    ids = sphere_mats(s, True)
    assert ids[1] == ids[3]
    _, s = load(5)  # left/right small spheres 2/7 of the way in from the walls
    sp = spheres(s, True)
    lw = (CB[0] + CB[4]) / 2 + [0, 0, 0.51]
    rw = (CB[1] + CB[5]) / 2 + [0, 0, 0.51]
    xl = rw[0] - lw[0]
    assert np.allclose(sp[0, :3], lw + [2 * xl / 7, 0, 0], atol=1e-6)
    assert np.allclose(sp[1, :3], rw - [2 * xl / 7, 0, 0], atol=1e-6)


def test_lights_camera_and_scene_sphere():
    _, s = load(0)
    L = s.lights[0]
    assert L.type == capi.UPBP_LIGHT_AREA and L.intensity[0] == pytest.approx(0.95492965)
    assert np.allclose(list(L.p0), CB[2]) and np.allclose(list(L.p1), CB[6]) and np.allclose(list(L.p2), CB[7])
    # scene sphere: half the diagonal of the box (all geometry inside it)
    assert s.scene_radius == pytest.approx(float(np.linalg.norm(CB[2] - CB[4]) / 2), rel=1e-6)
    assert list(s.camera.origin) == pytest.approx([-0.0439815, -4.12529, 0.222539])
    assert s.camera.horizontal_fov_deg == 45.0 and s.camera.mat_id == -1
    _, s = load(6)
    assert s.lights[0].type == capi.UPBP_LIGHT_POINT
    assert s.lights[0].intensity[0] == pytest.approx(70 / (4 * math.pi))
    _, s = load(5)
    assert s.lights[0].type == capi.UPBP_LIGHT_DIRECTIONAL and list(s.lights[0].intensity) == pytest.approx(
        [10, 4, 0])
    _, s = load(7)
    assert s.lights[0].type == capi.UPBP_LIGHT_BACKGROUND and s.background_light_idx == 0
    assert list(s.lights[0].intensity) == pytest.approx([135 / 255, 206 / 255, 250 / 255])
    _, s = load(13)  # no geometry at all: unit diagonal, centre at the origin
    assert s.scene_radius == 0.5 and list(s.scene_center) == [0, 0, 0]
    _, s = load(31)
    assert s.global_medium_idx == 11 and s.media[11].mean_cosine == pytest.approx(0.6)
    assert s.media[0].continuation_prob == 1.0 and s.media[1].continuation_prob == pytest.approx(0.8)


def test_command_line_scene_ids(capfd):
    rc, c = capi.parse_args([])
    assert rc == 0 and c.scene_id == 0 and c.output_name.startswith(b"s00_")
    for bad in ("3", "4", "9", "10", "41"):
        assert capi.parse_args(["-s", bad])[0] < 0
    assert capi.parse_args(["-s", "40"])[0] == 0
    assert capi.parse_args(["-hf"])[0] == capi.UPBP_HELP
    out = capfd.readouterr().out
    assert "large glass medium sphere + large area light" in out and "no geometry + anisoscattering" in out


# ---- analytic sphere: oracle pinned by the closed form -----------------------------------------------------
def _sphere_rays(rng, n_sph=6, n_rays=6000, dist=100.0):
    c = rng.uniform(-2, 2, (n_sph, 3))
    r = rng.uniform(0.05, 1.0, (n_sph, 1))
    sph = np.hstack([c, r]).astype(np.float32)
    o = rng.uniform(-dist, dist, (n_rays, 3))
    o[: n_rays // 6] = c[rng.integers(0, n_sph, n_rays // 6)] + rng.normal(scale=0.1, size=(n_rays // 6, 3))
    k = rng.integers(0, n_sph, n_rays)
    tgt = c[k] + rng.normal(scale=0.5, size=(n_rays, 3)) * r[k]
    d = tgt - o
    d[: n_rays // 6] = rng.normal(size=(n_rays // 6, 3))  # from inside or near a sphere
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tmax = np.where(rng.uniform(size=n_rays) < 0.2, rng.uniform(0.5, dist, n_rays), 1e30)
    rays = np.hstack([o, d, np.full((n_rays, 1), 1e-3), tmax[:, None]]).astype(np.float32)
    return sph, rays


def _closed_form(sph, rays):
    O = rays[:, None, :3].astype(np.float64) - sph[None, :, :3].astype(np.float64)
    D = rays[:, None, 3:6].astype(np.float64)
    D = D / np.linalg.norm(D, axis=2, keepdims=True)
    b = (O * D).sum(2)
    disc = b * b - ((O * O).sum(2) - sph[None, :, 3].astype(np.float64) ** 2)
    sq = np.sqrt(np.maximum(disc, 0))
    t0, t1 = -b - sq, -b + sq
    tmin, tmax = rays[:, None, 6].astype(np.float64), rays[:, None, 7].astype(np.float64)
    in0 = (disc > 0) & (t0 > tmin) & (t0 < tmax)
    in1 = (disc > 0) & (t1 > tmin) & (t1 < tmax)
    return in0, in1, t0, t1, disc


def test_oracle_sphere_roots_match_closed_form():
    rng = np.random.default_rng(11)
    sph, rays = _sphere_rays(rng)
    out = oracle_py.sphere_hits(sph, rays)
    in0, in1, t0, t1, disc = _closed_form(sph, rays)
    n = in0.astype(int) + in1.astype(int)
    # counts differ only where a root sits within rounding of tmin/tmax or the ray grazes the sphere
    r = sph[None, :, 3].astype(np.float64)
    # the first discriminant b^2 - c is formed in FP32 before any refinement: its error grows
    # like eps * b^2 for distant origins, so rays that graze within that band may go either way
    b = -(t0 + t1) / 2
    graze = np.abs(disc) <= 1e-4 * r * r + 2e-6 * b * b
    near_end = (np.abs(t0 - rays[:, None, 6]) <= 1e-4) | (np.abs(t1 - rays[:, None, 6]) <= 1e-4)
    near_end |= (np.abs(t0 - rays[:, None, 7]) <= 1e-3) | (np.abs(t1 - rays[:, None, 7]) <= 1e-3)
    safe = ~(graze | ((disc > 0) & near_end))
    assert safe.mean() > 0.97 and (n[safe] > 0).sum() > 2000 and (n[safe] == 1).sum() > 300
    assert np.array_equal(out[:, :, 0][safe], n[safe])
    # near root first, then the far one; the refinement keeps distant rays accurate to a few ulps
    both = safe & (n == 2)
    assert np.max(np.abs(out[:, :, 1][both] - t0[both]) / t0[both]) < 8e-6
    assert np.max(np.abs(out[:, :, 2][both] - t1[both]) / t1[both]) < 8e-6
    far_only = safe & ~in0 & in1
    assert far_only.sum() > 100
    assert np.max(np.abs(out[:, :, 1][far_only] - t1[far_only]) / np.abs(t1[far_only])) < 8e-6
    assert np.isinf(out[:, :, 1][safe & (n == 0)]).all()


def test_oracle_renders_predefined_scenes():
    # every renderable scene runs through the oracle (spheres, all light types, global media)
    for sid in (0, 2, 5, 7, 13, 21, 29, 35, 39):
        rc, cfg = capi.parse_args(["-s", str(sid), "-r", "16x16", "-l", "5", "-a",
                                   "upbp_bpt+surf+pp3d+pb2d+bb1d+bp2d"])
        assert rc == 0
        sh = capi.SceneHandle(sid, cfg.res_x, cfg.res_y)
        img, st = oracle_py.render(sh.ptr, cfg, 0, 1)
        assert np.isfinite(img).all() and (img >= 0).all() and st.path_errors == 0, sid
        assert img.mean() > 0, sid
