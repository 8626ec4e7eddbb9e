"""Build script for libupbp.so (HIP kernels for gfx950 + C++ host) and the gpuupbp CLI.

Every HIP source is compiled with hipcc for gfx950 only. Packed FP32 instructions
are disabled (-Xclang -target-feature -Xclang -packed-fp32-ops). Objects go to
gpuupbp_amd/_build/, the library and CLI into gpuupbp_amd/.

Usage: python -m gpuupbp_amd.build [--jobs N] [--oracle]
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libupbp.so")
CLI = os.path.join(HERE, "gpuupbp")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = "gfx950"

HOST_SOURCES = ["host/config.cpp", "host/scene_loader.cpp", "host/exr_writer.cpp", "host/image_writers.cpp",
                "host/scene_gen.cpp", "host/scene_predef.cpp",
                "host/exr_reader.cpp"]
HIP_SOURCES = ["device/renderer.hip", "device/lbvh.hip", "device/light_pass.hip", "device/camera_pass.hip",
               "device/merge_pass.hip", "device/test_hooks.hip", "device/guide_mlp.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]
HIP_FLAGS = COMMON + ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                      "-munsafe-fp-atomics", "-ffp-contract=off"]


def run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout)
        raise RuntimeError("compile failed: " + cmd[-1])
    return res.stdout


def obj_of(src):
    return os.path.join(BUILD, src.replace("/", "_").rsplit(".", 1)[0] + ".o")


def compile_one(src):
    path = os.path.join(CSRC, src)
    out = obj_of(src)
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", "-x", "hip", path, "-o", out]
    else:
        cmd = [HIPCC] + COMMON + ["-c", path, "-o", out]
    run(cmd)
    return out


def sources():
    return [s for s in HOST_SOURCES + HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(jobs=8, oracle=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, 16))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread", "-lz"])
    os.replace(tmp, LIB)
    main_src = os.path.join(CSRC, "host", "main.cpp")
    if os.path.exists(main_src):
        run([HIPCC] + COMMON + [main_src, "-o", CLI, "-L" + HERE, "-lupbp", "-Wl,-rpath,$ORIGIN"])
    if oracle:
        build_oracle()
    return LIB


def build_oracle():
    """The CPU oracle is test infrastructure; built beside the product for tests and bench."""
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, not a.no_oracle))
