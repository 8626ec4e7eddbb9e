"""The iteration captured in a HIP graph and replayed (GPU), and the device-resident iteration
under it: photon counts that never leave the device, and photon lists that outgrow their
capacity.

upbp_tuning.graph = 1 captures an iteration's work once (after one eager iteration with the same
settings) and replays it, reading each iteration's parameters from a staging slot. Every replay
must give the eager renderer's image after the same iterations, checked after each replay with
the program's own synchronisations in between; a settings change between replays must be picked
up (the graph is captured again). The same kernels run in both, so only the order of
floating-point atomic sums may differ.
"""
import numpy as np
import pytest

from conftest import box_args
from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu

pytestmark = pytest.mark.gpu


def _same(a, b):
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    return np.mean(rel < 1e-4) > 0.99 and abs(a.mean() - b.mean()) <= 1e-5 * abs(b.mean()) + 1e-12


@pytest.mark.parametrize("kind", ["box", "boxfog"])
def test_graph_replays_match_eager_iterations(scene_dir, kind):
    obj = str(scene_dir / kind / (kind + ".obj"))
    rc, cfg = capi.parse_args(["-s", "-1", obj] + box_args("64x48", extra=("-r_initial", "0.03", "-r_alpha", "0.75")))
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    eager = gpu.Renderer(sh.ptr, cfg)
    graph = gpu.Renderer(sh.ptr, cfg)
    graph.set_tuning(graph=1)
    for i in range(8):
        if i == 5:  # a settings change between replays: captured again
            eager.set_tuning(packet_bounces=64)
            graph.set_tuning(graph=1, packet_bounces=64)
        eager.run_iteration(i, want_stats=False)
        graph.run_iteration(i, want_stats=False)
        a, b = graph.read_accum(i + 1), eager.read_accum(i + 1)
        assert np.isfinite(a).all() and _same(a, b), (kind, i)
    st = graph.run_iteration(8)  # statistics: an eager iteration between replays
    graph.run_iteration(9, want_stats=False)
    eager.run_iteration(8, want_stats=False)
    eager.run_iteration(9, want_stats=False)
    assert st.path_errors == 0 and _same(graph.read_accum(10), eager.read_accum(10))
    eager.close()
    graph.close()
    sh.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_photon_lists_outgrow_their_capacity(scene_dir, graph):
    # The B-B1D radius shrinks with the iteration (-r_alpha_bb1d 0.5: r0 / sqrt(1 + i)), and
    # without B-P2D (whose radius would otherwise set the sub-beam length) beams are cut into more
    # sub-beams every iteration and outgrow the capacity the first iteration sized (a quarter to
    # spare): those iterations are discarded on the device, the lists resized and the iterations
    # rendered again. The same iterations in reverse order size for the largest count first and
    # never overflow; both renderers must end with the same image.
    obj = str(scene_dir / "boxfog" / "boxfog.obj")
    args = box_args("48x48", alg="upbp_bpt+surf+pp3d+pb2d+bb1d",
                    extra=("-r_initial", "0.03", "-r_alpha_bb1d", "0.5"))
    rc, cfg = capi.parse_args(["-s", "-1", obj] + args)
    assert rc == 0
    sh = capi.SceneHandle(obj, cfg.res_x, cfg.res_y)
    fwd, rev = gpu.Renderer(sh.ptr, cfg), gpu.Renderer(sh.ptr, cfg)
    fwd.set_tuning(graph=graph)
    n = 10
    for i in range(n - 1):
        fwd.run_iteration(i, want_stats=False)
    st_f = fwd.run_iteration(n - 1)
    for i in reversed(range(n)):
        st_r = rev.run_iteration(i)
    print("resizes forward %d, reverse %d" % (st_f.photon_list_resizes, st_r.photon_list_resizes))
    assert st_f.photon_list_resizes >= 2 and st_r.photon_list_resizes == 1
    a, b = fwd.read_accum(n), rev.read_accum(n)
    assert np.isfinite(a).all() and _same(a, b)
    fwd.close()
    rev.close()
    sh.close()
