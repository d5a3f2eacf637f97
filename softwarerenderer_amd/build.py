"""Build recipe for libswr_hip.so (gfx950), in-tree.

Every source is compiled by hipcc for gfx950 only, with contraction off (the draw path
reproduces the reference's float evaluation order) and packed-fp32 instruction selection
disabled.  Objects go to ``_build/``; the shared library lands next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libswr_hip.so")
SOURCES = ["swr_geometry.hip", "swr_sort.hip", "swr_raster.hip", "swr_lines.hip", "swr_diff.hip",
           "swr_api.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip", *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build_library(verbose: bool = False) -> str:
    """Compile every HIP source and link libswr_hip.so.  Returns the library path."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp, *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_pixel_module(src: str, out: str) -> str:
    """Compile a program module (a .hip file using include/swr/pixel_program.h or
    include/swr/vertex_program.h) into the shared library `out`, with the library's own flags
    (the raster chains and vertex arithmetic must round as the library's do).  Returns
    `out`."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    tmp = out + ".tmp"
    r = subprocess.run([HIPCC, *FLAGS, "-shared", src, "-o", tmp], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    os.replace(tmp, out)
    return out


build_vertex_module = build_pixel_module

TEST_MODULES = os.path.join(ROOT, "tests", "modules")


def build_test_modules() -> list:
    """tests/modules/*.hip -> tests/modules/lib*.so (the pixel and vertex modules the tests
    load)."""
    outs = []
    if not os.path.isdir(TEST_MODULES):
        return outs
    srcs = sorted(f for f in os.listdir(TEST_MODULES) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=max(1, min(8, len(srcs)))) as ex:
        outs = list(ex.map(lambda f: build_pixel_module(
            os.path.join(TEST_MODULES, f),
            os.path.join(TEST_MODULES, "lib" + os.path.splitext(f)[0] + ".so")), srcs))
    return outs


def build_cpp_example() -> str:
    """tests/cpp/render_scene: the C++ host API (include/swr/Renderer.h) driving the library."""
    src = os.path.join(ROOT, "tests", "cpp", "render_scene.cpp")
    exe = os.path.join(ROOT, "tests", "cpp", "render_scene")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
                        src, "-L" + HERE, "-lswr_hip",
                        "-Wl,-rpath,$ORIGIN/../../softwarerenderer_amd", "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("render_scene build failed:\n" + r.stderr)
    return exe


if __name__ == "__main__":
    build_library(verbose=True)
    build_cpp_example()
    build_test_modules()
