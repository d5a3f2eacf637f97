"""The C ABI boundary: the library exports every entry point include/swr_hip.h declares,
the ctypes mirrors have the C layouts, and the C++ host API compiles (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

import oracle as O
from softwarerenderer_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "swr_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(swr_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    lib = C.CDLL(_abi.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.SIGNATURES), "ctypes signature table out of date"


def _c_layout(code: str) -> dict:
    tmp = tempfile.mkdtemp(prefix="swr_layout_")
    src = os.path.join(tmp, "probe.c")
    exe = os.path.join(tmp, "probe")
    with open(src, "w") as f:
        f.write('#include <stdio.h>\n#include <stddef.h>\n#include "swr_hip.h"\n'
                '#include "swr_oracle.h"\nint main(void){\n' + code + "return 0;}\n")
    subprocess.run(["gcc", "-std=c99", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "oracle"), src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


@pytest.mark.parametrize("cname,py", [("swr_pixel_shader_desc", _abi.PixelShaderDesc),
                                      ("swr_vertex_shader_desc", _abi.VertexShaderDesc),
                                      ("swr_draw_stats", _abi.DrawStats),
                                      ("swr_diff_desc", _abi.DiffDesc),
                                      ("swr_scene", _abi.Scene),
                                      ("swr_tuning", _abi.Tuning),
                                      ("swr_pixel_module", _abi.PixelModule),
                                      ("swr_vertex_module", _abi.VertexModule),
                                      ("swr_pixel_shader_desc", O.PixelShaderDesc),
                                      ("swr_vertex_shader_desc", O.VertexShaderDesc),
                                      ("swr_pipeline_state", O.PipelineState)])
def test_struct_layouts(cname, py):
    code = f'printf("size %zu\\n", sizeof({cname}));\n'
    for name, _ in py._fields_:
        code += f'printf("{name} %zu\\n", offsetof({cname}, {name}));\n'
    lay = _c_layout(code)
    assert lay["size"] == C.sizeof(py)
    for name, _ in py._fields_:
        assert lay[name] == getattr(py, name).offset, name


def test_rasterizer_vertex_layout():
    lay = _c_layout('printf("size %zu\\n", sizeof(swr_rasterizer_vertex));\n'
                    'printf("avar %zu\\n", offsetof(swr_rasterizer_vertex, avar));\n'
                    'printf("pvar %zu\\n", offsetof(swr_rasterizer_vertex, pvar));\n')
    from softwarerenderer_amd.api import RASTERIZER_VERTEX
    assert lay["size"] == RASTERIZER_VERTEX.itemsize == 144  # IRasterizer.h:42-53
    assert lay["avar"] == RASTERIZER_VERTEX.fields["avar"][1]
    assert lay["pvar"] == RASTERIZER_VERTEX.fields["pvar"][1]


def test_cpp_host_api_compiles(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "render_scene.cpp")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                    "-I" + os.path.join(ROOT, "include"), src], check=True)


def test_context_needs_gfx950_device():
    """Without a GPU, swr_create fails cleanly (no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _abi.load()
    ctx = C.c_void_p()
    assert lib.swr_create(C.byref(ctx), 0) == -4  # SWR_ERR_NO_DEVICE
    assert not ctx.value
