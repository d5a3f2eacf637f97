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


# ---- state variants ------------------------------------------------------------------------

@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("scissor", [(0, 0, 640, 480), (37, 21, 301, 250), (-40, -8, 500, 300),
                                     (600, 400, 200, 200), (0, 0, 0, 0), (100, 50, 3, 7)])
def test_scissor(gpu, cpu_oracle, sphere, mode, scissor):
    """Block mode shades whole blocks past the scissor (PixelShaderBase.h:68-93) and walks
    blocks upward when both step counts go negative (Rasterizer.h:254-258); Span clips x
    only (Rasterizer.h:377-378)."""
    pipe = O.Pipeline(640, 480, raster_mode=mode, mvp=sphere.mvp, scissor=scissor)
    a = run_cpu(cpu_oracle, pipe, sphere)
    b = run_gpu(pipe, sphere)
    assert_same(a, b, f"scissor {scissor} mode={mode}")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("viewport", [(100, 100, 440, 280), (-50, 30, 700, 500), (0, 0, 320, 480)])
def test_viewport_depth_range(gpu, cpu_oracle, sphere, mode, viewport):
    pipe = O.Pipeline(640, 480, raster_mode=mode, mvp=sphere.mvp, viewport=viewport,
                      depth_range=(0.25, 0.75))
    a = run_cpu(cpu_oracle, pipe, sphere)
    b = run_gpu(pipe, sphere)
    assert_same(a, b, f"viewport {viewport} mode={mode}")


@pytest.mark.parametrize("counts", [(0, 0), (3, 0), (0, 2), (1, 1), (2, 3), (4, 4), (5, 0),
                                    (0, 5), (16, 16)])
@pytest.mark.parametrize("zw", [(False, False), (True, True), (True, False)])
def test_varying_counts(gpu, cpu_oracle, counts, zw):
    """The pixel shader's statics choose what is interpolated (PixelShaderBase.h:44-53); the
    vertex shader's choose what the clipper lerps (PolyClipper.h:42-45)."""
    na, np_ = counts
    sc = scenes.clip_scene(ntri=800, seed=11, navar=na, npvar=np_)
    for mode in MODES:
        pipe = O.Pipeline(sc.width, sc.height, raster_mode=mode, vs_program=O.VS_RAW,
                          vs_counts=(1, na, np_), ps_a=na, ps_p=np_, ps_z=zw[0], ps_w=zw[1])
        a = run_cpu(cpu_oracle, pipe, sc)
        b = run_gpu(pipe, sc)
        assert_same(a, b, f"counts {counts} zw {zw} mode {mode}")


def test_big_triangles(gpu, cpu_oracle):
    """Screen-sized triangles and slivers: cooperative pair expansion and long Span replays."""
    rng = np.random.default_rng(7)
    n = 40
    v = np.zeros((3 * n, 9), np.float32)
    v[:, 0:2] = rng.uniform(-3, 3, (3 * n, 2))
    v[:, 2] = rng.uniform(-0.9, 0.9, 3 * n)
    v[:, 3] = 1.0
    v[:, 4:9] = rng.random((3 * n, 5))
    sc = scenes.Scene(v, np.arange(3 * n, dtype=np.int32), np.eye(4, dtype=np.float32), 320, 240)
    for mode in MODES:
        for cull in (O.CULL_NONE, O.CULL_CW):
            pipe = O.Pipeline(320, 240, raster_mode=mode, cull_mode=cull, vs_program=O.VS_RAW,
                              vs_counts=(1, 3, 2))
            a = run_cpu(cpu_oracle, pipe, sc)
            b = run_gpu(pipe, sc)
            assert_same(a, b, f"big mode={mode} cull={cull}")


def test_batch_boundaries(gpu, cpu_oracle):
    """Clipped fans are drawn after the rest of their 1024-triangle batch
    (VertexProcessor.cpp:254-261, flush at :110); keys follow the index-list slots."""
    sc = scenes.clip_scene(ntri=3 * 1024 + 517, seed=21, width=160, height=120, shrink=(0.01, 0.2))
    for mode in MODES:
        pipe = O.Pipeline(160, 120, raster_mode=mode, cull_mode=O.CULL_NONE,
                          vs_program=O.VS_RAW, vs_counts=(1, 3, 2), ps_program=O.PS_LOG)
        a = run_cpu(cpu_oracle, pipe, sc, log_capacity=400_000)
        b = run_gpu(pipe, sc, log_capacity=400_000)
        assert_same(a, b, f"batches mode={mode}")


def test_partial_and_empty(gpu, cpu_oracle, sphere):
    pipe = O.Pipeline(640, 480, raster_mode=O.RASTER_SPAN, mvp=sphere.mvp)
    for n in (0, 1, 2, 3, 3 * 1024, 3 * 1024 + 2, 3 * 2000 + 1):
        idx = sphere.indices[:n]
        a = run_cpu(cpu_oracle, pipe, sphere, indices=idx)
        b = run_gpu(pipe, sphere, indices=idx)
        assert_same(a, b, f"{n} indices")


def test_benchmark_scene(gpu, cpu_oracle):
    """The reference Benchmark (Benchmark.cpp:66-105): 40,960 large triangles, Span, cull
    None, {x,y,z,r,g,b} vertices with w = 1."""
    sc = scenes.benchmark_scene()
    pipe = O.Pipeline(640, 480, raster_mode=O.RASTER_SPAN, cull_mode=O.CULL_NONE,
                      vs_program=O.VS_POS_RGB, vs_counts=(1, 3, 0), ps_a=3, ps_p=0,
                      ps_z=False, ps_w=False)
    a = run_cpu(cpu_oracle, pipe, sc)
    b = run_gpu(pipe, sc)
    assert int(a.t1.astype(np.int64).sum()) == 240_235_639
    assert_same(a, b, "benchmark scene")


# ---- the IRasterizer seam (window-space triangle lists) -------------------------------------

def _window_tris(n, seed, w=96, h=64, grid=False):
    rng = np.random.default_rng(seed)
    v = np.zeros(3 * n, O.RASTERIZER_VERTEX)
    if grid:  # vertices on pixel centres / integer corners: exercise the tie rule
        v["x"] = rng.integers(-4, 2 * w + 8, 3 * n) * 0.5
        v["y"] = rng.integers(-4, 2 * h + 8, 3 * n) * 0.5
    else:
        v["x"] = rng.uniform(-10, w + 10, 3 * n)
        v["y"] = rng.uniform(-10, h + 10, 3 * n)
    v["z"] = rng.random(3 * n)
    v["w"] = rng.uniform(0.5, 2.0, 3 * n)
    v["avar"] = rng.random((3 * n, 16))
    v["pvar"] = rng.random((3 * n, 16))
    idx = np.arange(3 * n, dtype=np.int32)
    skip = rng.random(n) < 0.1
    idx.reshape(-1, 3)[skip, 0] = -1
    return v, idx


@pytest.mark.parametrize("grid", [False, True])
@pytest.mark.parametrize("mode", MODES)
def test_triangle_list(gpu, cpu_oracle, grid, mode):
    v, idx = _window_tris(400, seed=5 + grid, grid=grid)
    pipe = O.Pipeline(96, 64, raster_mode=mode, scissor=(0, 0, 96, 64), ps_program=O.PS_LOG,
                      ps_a=4, ps_p=3)
    a = run_cpu(cpu_oracle, pipe, vertices=v, indices=idx, log_capacity=200_000)
    b = run_gpu(pipe, vertices=v, indices=idx, log_capacity=200_000)
    assert_same(a, b, f"triangle list grid={grid} mode={mode}")


def test_rasterizer_test_triangle(gpu, cpu_oracle):
    """RasterizerTest.cpp:53-81: one RGB triangle drawn through Rasterizer::drawTriangle."""
    v = np.zeros(3, O.RASTERIZER_VERTEX)
    v["x"], v["y"] = [320, 480, 120], [100, 200, 300]
    v["avar"][0, :3], v["avar"][1, :3], v["avar"][2, :3] = [1, 0, 0], [0, 1, 0], [0, 0, 1]
    for mode in MODES:
        pipe = O.Pipeline(640, 480, raster_mode=mode, ps_program=O.PS_RGB, ps_a=3, ps_p=0,
                          ps_z=False, ps_w=False)
        a = run_cpu(cpu_oracle, pipe, vertices=v, indices=np.arange(3, dtype=np.int32))
        b = run_gpu(pipe, vertices=v, indices=np.arange(3, dtype=np.int32))
        assert_same(a, b, f"RasterizerTest mode={mode}")
        assert (a.t0 != 0).sum() > 20000


@pytest.mark.parametrize("mode", [O.RASTER_SPAN, O.RASTER_BLOCK])
@pytest.mark.parametrize("counts", [(0, 0), (2, 1), (3, 2), (4, 4), (8, 8), (16, 16)])
def test_clip_varying_classes(gpu, cpu_oracle, counts, mode):
    """Clipped triangles through every varying class of the geometry stage (0, 3, 5, 8, 16
    and 32 varyings: the register clippers, the clip queue's polygon store in LDS, and its
    per-thread store for wide varyings), full invocation logs against the oracle."""
    na, np_ = counts
    sc = scenes.clip_scene(ntri=3000, seed=17 + na, width=160, height=120, navar=na, npvar=np_,
                           shrink=(0.01, 0.25))
    pipe = O.Pipeline(160, 120, raster_mode=mode, cull_mode=O.CULL_NONE, vs_program=O.VS_RAW,
                      vs_counts=(1, na, np_), ps_program=O.PS_LOG, ps_a=na, ps_p=np_)
    a = run_cpu(cpu_oracle, pipe, sc, log_capacity=400_000)
    b = run_gpu(pipe, sc, log_capacity=400_000)
    assert int(a.log_count[0]) > 10_000
    assert_same(a, b, f"clip varyings={counts} mode={mode}")
