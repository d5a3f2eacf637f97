"""The reference's demos, headless, on the GPU against the oracle bit for bit: the
VertexProcessorTest triangle (RGB shader, viewport offset, cull None) and the Box demo's
pipeline over an OBJ crate loaded with softwarerenderer_amd/obj.py (Span, cull CW, MVP
vertex program, texture coordinates as perspective varyings).  The oracle is pinned on the
same scenes against the reference build in test_oracle_vs_ref.py."""
import os

import numpy as np
import pytest

import oracle as O
from softwarerenderer_amd import scenes
from helpers import assert_same, run_cpu, run_gpu

pytestmark = pytest.mark.gpu

CRATE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "obj", "crate.obj")


@pytest.mark.parametrize("prog", [O.PS_RGB, O.PS_LOG])
def test_vertex_processor_test(gpu, cpu_oracle, prog):
    sc = scenes.vertex_processor_test_scene()
    pipe = O.Pipeline(640, 480, cull_mode=O.CULL_NONE, viewport=(100, 100, 440, 280),
                      vs_program=O.VS_POS_RGB, vs_counts=(1, 3, 0), ps_program=prog,
                      ps_z=False, ps_w=False, ps_a=3, ps_p=0)
    a = run_cpu(cpu_oracle, pipe, sc, log_capacity=200_000)
    b = run_gpu(pipe, sc, log_capacity=200_000)
    assert_same(a, b, f"VertexProcessorTest prog={prog}")
    if prog == O.PS_RGB:
        assert (b.t0 != 0).sum() > 20_000


@pytest.mark.parametrize("angle", [0.0, 0.6, 2.2, 4.0])
@pytest.mark.parametrize("mode", [O.RASTER_SPAN, O.RASTER_BLOCK])
def test_box(gpu, cpu_oracle, angle, mode):
    sc = scenes.obj_scene(CRATE, angle)
    pipe = O.Pipeline(640, 480, raster_mode=mode, mvp=sc.mvp, vs_program=O.VS_MVP_POS_UV,
                      vs_counts=(1, 0, 2), ps_program=O.PS_LOG, ps_z=False, ps_w=True,
                      ps_a=0, ps_p=2)
    a = run_cpu(cpu_oracle, pipe, sc, log_capacity=400_000)
    b = run_gpu(pipe, sc, log_capacity=400_000)
    assert int(a.log_count[0]) > 10_000
    assert_same(a, b, f"box angle={angle} mode={mode}")
