"""The texture-sampling pixel program (the Box demo's shader, Box.cpp:27-53) on the GPU
against the oracle, bit for bit: mip-mapped trilinear taps with the anisotropic footprint
of Texture::sample over derivatives from the triangle's planes.  The oracle's sampler is
pinned in test_oracle_texture.py (the reference's Texture class needs SDL and cannot be
built here)."""
import os

import numpy as np
import pytest

import oracle as O
from softwarerenderer_amd import scenes
from helpers import assert_same, run_cpu, run_gpu

pytestmark = pytest.mark.gpu

CRATE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "obj", "crate.obj")
TEXTURES = [("noise", 64, 64), ("checker", 256, 128), ("noise", 37, 23)]


def _pipe(sc, tex, mode, aniso=8):
    return O.Pipeline(sc.width, sc.height, raster_mode=mode, mvp=sc.mvp,
                      vs_program=O.VS_MVP_POS_UV, vs_counts=(1, 0, 2), ps_program=O.PS_TEXTURE,
                      ps_z=False, ps_w=True, ps_a=0, ps_p=2,
                      texture=O.OracleTexture(tex, max_anisotropy=aniso))


@pytest.mark.parametrize("texture", TEXTURES, ids=lambda t: f"{t[0]}{t[1]}x{t[2]}")
@pytest.mark.parametrize("angle", [0.0, 0.6, 2.2])
@pytest.mark.parametrize("mode", [O.RASTER_SPAN, O.RASTER_BLOCK])
def test_box_textured(gpu, cpu_oracle, texture, angle, mode):
    sc = scenes.obj_scene(CRATE, angle)
    tex = scenes.procedural_texture(*texture, seed=1)
    pipe = _pipe(sc, tex, mode)
    a = run_cpu(cpu_oracle, pipe, sc)
    b = run_gpu(pipe, sc)
    assert (a.t0 != 0).sum() > 20_000
    assert_same(a, b, f"box texture={texture} angle={angle} mode={mode}")


@pytest.mark.parametrize("aniso", [1, 4, 16])
@pytest.mark.parametrize("mode", [O.RASTER_SPAN, O.RASTER_ADAPTIVE])
def test_floor_anisotropic(gpu, cpu_oracle, aniso, mode):
    sc = scenes.floor_scene()
    tex = scenes.procedural_texture("checker", 128, 128)
    pipe = _pipe(sc, tex, mode, aniso)
    a = run_cpu(cpu_oracle, pipe, sc)
    b = run_gpu(pipe, sc)
    assert (a.t0 != 0).sum() > 50_000
    assert_same(a, b, f"floor aniso={aniso} mode={mode}")


def test_texture_rejects_lines(gpu):
    import torch
    from softwarerenderer_amd import _abi, api
    r = api.Rasterizer(0)
    try:
        tex = api.Texture(r, scenes.procedural_texture("noise", 8, 8))
        color = torch.zeros((16, 16), dtype=torch.int32, device=gpu)
        r.setScissorRect(0, 0, 16, 16)
        r.setPixelShader(api.TextureShader(color, tex))
        v = np.zeros(2, O.RASTERIZER_VERTEX)
        v["x"], v["y"], v["w"] = [1.0, 9.0], [1.0, 5.0], 1.0
        vt = torch.from_numpy(v.view(np.uint8)).to(gpu)
        with pytest.raises(_abi.SwrError) as e:
            r.drawLineList(vt, torch.tensor([0, 1], dtype=torch.int32, device=gpu))
        assert e.value.code == -2
        tex.close()
    finally:
        r.close()
