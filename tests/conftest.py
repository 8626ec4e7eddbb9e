import os
import sys

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
