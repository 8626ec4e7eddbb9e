"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU): the scene loader
(OBJ/MTL/AUX, and an environment map's EXR through the loader), the OpenEXR reader and the
command-line parser are built with g++ and the
sanitizers into a standalone driver (tests/fuzz/host_fuzz.cpp) and fed corrupted inputs. Every
call must come back with a result or an error code: a sanitizer report, or a C++ exception through
the C ABI (which would end the caller's process), fails the test."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = ["config", "scene_loader", "exr_writer", "image_writers", "scene_gen", "scene_predef", "exr_reader"]
FLAGS = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
         "-fno-omit-frame-pointer", "-I" + os.path.join(ROOT, "include")]


def _build(out_dir):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    srcs = [os.path.join(ROOT, "tests", "fuzz", "host_fuzz.cpp")] + \
           [os.path.join(ROOT, "gpuupbp_amd", "csrc", "host", n + ".cpp") for n in HOST]

    def cc(src):
        obj = os.path.join(out_dir, os.path.basename(src) + ".o")
        r = subprocess.run([gxx] + FLAGS + ["-c", src, "-o", obj], capture_output=True, text=True)
        return obj, r

    with ThreadPoolExecutor(max_workers=8) as ex:
        res = list(ex.map(cc, srcs))
    for obj, r in res:
        if r.returncode != 0:
            if "asan" in r.stderr.lower() or "ubsan" in r.stderr.lower():
                pytest.skip("the sanitizer runtimes are not installed")
            raise AssertionError(r.stderr[-2000:])
    exe = os.path.join(out_dir, "host_fuzz")
    r = subprocess.run([gxx] + FLAGS + [o for o, _ in res] + ["-o", exe, "-lpthread", "-lz"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        if "asan" in r.stderr.lower() or "ubsan" in r.stderr.lower():
            pytest.skip("the sanitizer runtimes are not installed")
        raise AssertionError(r.stderr[-2000:])
    return exe


def test_host_parsers_survive_corrupted_inputs(tmp_path):
    exe = _build(str(tmp_path))
    work = tmp_path / "work"
    work.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, "60", str(work), "box,boxfog,furnace"], capture_output=True, text=True, env=env,
                       timeout=600)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out, out[-4000:]
    assert r.returncode == 0 and "fuzz done" in r.stdout, out[-4000:]
    # the corruptions reach both outcomes: inputs that still load and inputs that are rejected
    for what in ("scene loads", "env scene loads", "exr", "argv"):
        line = next(l for l in r.stdout.splitlines() if l.startswith(what))
        ok, bad = [int(t) for t in line.replace(",", " ").split() if t.isdigit()]
        assert ok > 0 and bad > 0, line
