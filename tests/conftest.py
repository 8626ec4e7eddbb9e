import os
import shutil
import sys

import numpy as np

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def scene_dir(tmp_path_factory):
    from gpuupbp_amd import capi
    d = tmp_path_factory.mktemp("scenes")
    for kind in ("box", "furnace", "boxfog"):
        os.makedirs(str(d / kind), exist_ok=True)
        capi.generate_scene(kind, 0, str(d / kind / (kind + ".obj")))
    return d


@pytest.fixture(scope="session")
def box_obj(scene_dir):
    return str(scene_dir / "box" / "box.obj")


@pytest.fixture(scope="session")
def boxfog_obj(scene_dir):
    return str(scene_dir / "boxfog" / "boxfog.obj")


def box_args(res="64x64", alg="upbp_bpt+surf+pp3d+pb2d+bb1d+bp2d", extra=()):
    """Command line of the box workload (box [-1,1]^3, medium sphere, l=8, r=0.01)."""
    return ["-r", res, "-l", "8", "-a", alg, "-r_initial", "0.01", "-r_alpha", "1"] + list(extra)


def env_scene(box_obj, dst, env, scale=2.0, rotate=0.25):
    """The box workload lit additionally by an environment map through its open front."""
    os.makedirs(dst, exist_ok=True)
    base = os.path.dirname(box_obj)
    for f in os.listdir(base):
        shutil.copy(os.path.join(base, f), dst)
    obj = os.path.join(dst, os.path.basename(box_obj))
    from gpuupbp_amd import renderer as gpu
    gpu.save_exr(os.path.join(dst, "env.exr"), env)
    with open(obj + ".aux", "a") as f:
        f.write("\nlight_background_em %g %g env.exr\n" % (scale, rotate))
    return obj


def sky(h=16, w=32):
    """Positive everywhere, a bright sun patch and a horizon gradient."""
    v = (np.arange(h) + 0.5) / h
    img = np.zeros((h, w, 3), np.float32)
    img[...] = (0.2 + 0.6 * (1 - v))[:, None, None] * np.array([0.6, 0.8, 1.0], np.float32)
    img[3:5, 20:23] = [40.0, 35.0, 25.0]
    return img


def relmse(a, b):
    return float(np.mean((a - b) ** 2 / (b ** 2 + 1e-2)))


def oracle_independent(scene_ptr, cfg, first=0, iters=1):
    """The oracle's render of the same iterations on other random streams (iteration offset 2^16):
    the error of independent sampling, against which same-stream differences are judged."""
    from oracle import oracle_py
    cfg.iteration_offset = 1 << 16
    try:
        img, _ = oracle_py.render(scene_ptr, cfg, first, iters)
    finally:
        cfg.iteration_offset = 0
    return img


def assert_whole_image(g, o, o_indep, ratio=0.1):
    """Same-stream images: beyond the pixels that agree to FP32 rounding, the image as a whole must
    stay far closer to the oracle's than an independent render of the same iterations is."""
    r_same, r_ind = relmse(g, o), relmse(o_indep, o)
    assert r_same < ratio * r_ind, (r_same, r_ind)
