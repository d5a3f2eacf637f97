"""GPU parity: the gfx950 draw path against the oracle on identical inputs.

Every comparison is bit-exact: target images are compared as raw 32-bit words, and the
hash / log programs record, per pixel and in invocation order, the primitive key and the
bits of z, w, 1/w and every varying the pixel shader declares.
"""
import numpy as np
import pytest

import oracle as O
from softwarerenderer_amd import scenes
from helpers import assert_same, run_cpu, run_gpu

pytestmark = pytest.mark.gpu

MODES = [O.RASTER_SPAN, O.RASTER_BLOCK, O.RASTER_ADAPTIVE]


@pytest.fixture(scope="module")
def sphere():
    return scenes.sphere_scene()


@pytest.fixture(scope="module")
def clip():
    return scenes.clip_scene(ntri=3000, seed=3)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("prog", [O.PS_HASH, O.PS_DEPTH_CHECKER])
def test_sphere(gpu, cpu_oracle, sphere, mode, prog):
    pipe = O.Pipeline(640, 480, raster_mode=mode, mvp=sphere.mvp, ps_program=prog)
    a = run_cpu(cpu_oracle, pipe, sphere)
    b = run_gpu(pipe, sphere)
    assert_same(a, b, f"sphere mode={mode} prog={prog}")
    if prog == O.PS_HASH:
        assert int(a.t1.sum()) == 67832


@pytest.mark.parametrize("mode", MODES)
def test_sphere_log(gpu, cpu_oracle, sphere, mode):
    pipe = O.Pipeline(640, 480, raster_mode=mode, mvp=sphere.mvp, ps_program=O.PS_LOG)
    a = run_cpu(cpu_oracle, pipe, sphere, log_capacity=100_000)
    b = run_gpu(pipe, sphere, log_capacity=100_000)
    assert int(a.log_count[0]) == 67832
    assert_same(a, b, f"sphere log mode={mode}")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("cull", [O.CULL_NONE, O.CULL_CW, O.CULL_CCW])
def test_clip_scene(gpu, cpu_oracle, clip, mode, cull):
    pipe = O.Pipeline(clip.width, clip.height, raster_mode=mode, cull_mode=cull,
                      vs_program=O.VS_RAW, vs_counts=(1, 3, 2))
    a = run_cpu(cpu_oracle, pipe, clip)
    b = run_gpu(pipe, clip)
    assert_same(a, b, f"clip mode={mode} cull={cull}")
    assert a.t1.sum() > 0


@pytest.mark.parametrize("mode", [O.RASTER_SPAN, O.RASTER_BLOCK])
def test_soup_subset(gpu, cpu_oracle, mode):
    sc = scenes.soup_scene(60_000, seed=5)
    pipe = O.Pipeline(1920, 1080, raster_mode=mode, mvp=sc.mvp)
    a = run_cpu(cpu_oracle, pipe, sc)
    b = run_gpu(pipe, sc)
    assert_same(a, b, f"soup mode={mode}")


def test_host_buffers(gpu, cpu_oracle, sphere):
    pipe = O.Pipeline(640, 480, raster_mode=O.RASTER_BLOCK, mvp=sphere.mvp)
    a = run_cpu(cpu_oracle, pipe, sphere)
    b = run_gpu(pipe, sphere, device_buffers=False)
    assert_same(a, b, "host buffers")
