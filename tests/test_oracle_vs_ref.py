"""The oracle against the reference itself (oracle/_ref: trenki2/SoftwareRenderer's own
src/renderer sources compiled in this container), call for call, bit for bit.

These tests pin the C restatement: every case the GPU parity tests use is drawn here by
both CPU implementations and compared as raw 32-bit words (images) or full invocation
logs.  They skip where the reference build is absent (it is built from /root/reference).
"""
import numpy as np
import pytest

import oracle as O
from softwarerenderer_amd import scenes
from helpers import assert_same, run_cpu

MODES = [O.RASTER_SPAN, O.RASTER_BLOCK, O.RASTER_ADAPTIVE]


@pytest.fixture(scope="module")
def sphere():
    return scenes.sphere_scene()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("prog", [O.PS_HASH, O.PS_DEPTH_CHECKER])
def test_sphere(cpu_oracle, cpu_ref, sphere, mode, prog):
    pipe = O.Pipeline(640, 480, raster_mode=mode, mvp=sphere.mvp, ps_program=prog)
    assert_same(run_cpu(cpu_ref, pipe, sphere), run_cpu(cpu_oracle, pipe, sphere), "sphere")


@pytest.mark.parametrize("mode", MODES)
def test_sphere_log(cpu_oracle, cpu_ref, sphere, mode):
    pipe = O.Pipeline(640, 480, raster_mode=mode, mvp=sphere.mvp, ps_program=O.PS_LOG)
    a = run_cpu(cpu_ref, pipe, sphere, log_capacity=100_000)
    b = run_cpu(cpu_oracle, pipe, sphere, log_capacity=100_000)
    assert int(a.log_count[0]) == 67832
    assert_same(a, b, "sphere log")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("cull", [O.CULL_NONE, O.CULL_CW, O.CULL_CCW])
def test_clip_scene(cpu_oracle, cpu_ref, mode, cull):
    sc = scenes.clip_scene(ntri=3000, seed=3)
    pipe = O.Pipeline(sc.width, sc.height, raster_mode=mode, cull_mode=cull,
                      vs_program=O.VS_RAW, vs_counts=(1, 3, 2))
    assert_same(run_cpu(cpu_ref, pipe, sc), run_cpu(cpu_oracle, pipe, sc), "clip")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("scissor", [(37, 21, 301, 250), (-40, -8, 500, 300),
                                     (600, 400, 200, 200), (0, 0, 0, 0), (100, 50, 3, 7)])
def test_scissor(cpu_oracle, cpu_ref, sphere, mode, scissor):
    pipe = O.Pipeline(640, 480, raster_mode=mode, mvp=sphere.mvp, scissor=scissor)
    assert_same(run_cpu(cpu_ref, pipe, sphere), run_cpu(cpu_oracle, pipe, sphere), "scissor")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("viewport", [(100, 100, 440, 280), (-50, 30, 700, 500)])
def test_viewport_depth_range(cpu_oracle, cpu_ref, sphere, mode, viewport):
    pipe = O.Pipeline(640, 480, raster_mode=mode, mvp=sphere.mvp, viewport=viewport,
                      depth_range=(0.25, 0.75))
    assert_same(run_cpu(cpu_ref, pipe, sphere), run_cpu(cpu_oracle, pipe, sphere), "viewport")


@pytest.mark.parametrize("counts", [(0, 0), (3, 0), (0, 2), (1, 1), (2, 3), (4, 4), (16, 16)])
@pytest.mark.parametrize("zw", [(False, False), (True, True), (True, False)])
def test_varying_counts(cpu_oracle, cpu_ref, counts, zw):
    na, np_ = counts
    sc = scenes.clip_scene(ntri=800, seed=11, navar=na, npvar=np_)
    for mode in MODES:
        pipe = O.Pipeline(sc.width, sc.height, raster_mode=mode, vs_program=O.VS_RAW,
                          vs_counts=(1, na, np_), ps_a=na, ps_p=np_, ps_z=zw[0], ps_w=zw[1])
        a, b = run_cpu(cpu_ref, pipe, sc), run_cpu(cpu_oracle, pipe, sc)
        if mode != O.RASTER_SPAN and np_ > 0 and not zw[1]:
            # Block traversal copies 1/w only when InterpolateW (PixelShaderBase.h:122-136): the
            # reference's pvar values are undefined there; coverage and order still agree
            assert np.array_equal(a.t1, b.t1)
            continue
        assert_same(a, b, f"counts {counts}")


def test_batch_boundaries(cpu_oracle, cpu_ref):
    sc = scenes.clip_scene(ntri=3 * 1024 + 517, seed=21, width=160, height=120, shrink=(0.01, 0.2))
    for mode in MODES:
        pipe = O.Pipeline(160, 120, raster_mode=mode, cull_mode=O.CULL_NONE,
                          vs_program=O.VS_RAW, vs_counts=(1, 3, 2), ps_program=O.PS_LOG)
        a = run_cpu(cpu_ref, pipe, sc, log_capacity=400_000)
        b = run_cpu(cpu_oracle, pipe, sc, log_capacity=400_000)
        assert int(a.log_count[0]) < 400_000
        assert_same(a, b, "batches")


def test_partial_and_empty(cpu_oracle, cpu_ref, sphere):
    pipe = O.Pipeline(640, 480, raster_mode=O.RASTER_SPAN, mvp=sphere.mvp)
    for n in (0, 1, 2, 3, 3 * 1024, 3 * 1024 + 2, 3 * 2000 + 1):
        idx = sphere.indices[:n]
        assert_same(run_cpu(cpu_ref, pipe, sphere, indices=idx),
                    run_cpu(cpu_oracle, pipe, sphere, indices=idx), f"{n} indices")


def test_benchmark_scene(cpu_oracle, cpu_ref):
    sc = scenes.benchmark_scene()
    pipe = O.Pipeline(640, 480, raster_mode=O.RASTER_SPAN, cull_mode=O.CULL_NONE,
                      vs_program=O.VS_POS_RGB, vs_counts=(1, 3, 0), ps_a=3, ps_p=0,
                      ps_z=False, ps_w=False)
    a = run_cpu(cpu_ref, pipe, sc)
    b = run_cpu(cpu_oracle, pipe, sc)
    assert int(a.t1.astype(np.int64).sum()) == 240_235_639
    assert_same(a, b, "benchmark")


@pytest.mark.parametrize("grid", [False, True])
@pytest.mark.parametrize("mode", MODES)
def test_triangle_list(cpu_oracle, cpu_ref, grid, mode):
    from test_gpu_parity import _window_tris
    v, idx = _window_tris(400, seed=5 + grid, grid=grid)
    pipe = O.Pipeline(96, 64, raster_mode=mode, scissor=(0, 0, 96, 64), ps_program=O.PS_LOG,
                      ps_a=4, ps_p=3)
    a = run_cpu(cpu_ref, pipe, vertices=v, indices=idx, log_capacity=200_000)
    b = run_cpu(cpu_oracle, pipe, vertices=v, indices=idx, log_capacity=200_000)
    assert_same(a, b, "triangle list")


def test_soup(cpu_oracle, cpu_ref):
    sc = scenes.soup_scene(30_000, seed=5)
    for mode in MODES:
        pipe = O.Pipeline(1920, 1080, raster_mode=mode, mvp=sc.mvp)
        assert_same(run_cpu(cpu_ref, pipe, sc), run_cpu(cpu_oracle, pipe, sc), "soup")


# ---- lines and points ---------------------------------------------------------------------
# drawElements(Line / Point): LineClipper / clipPoints, transformVertices, then the DDA of
# Rasterizer.h:175-222 and drawPointTemplate (:149-158); the list entry points directly.

# (program, InterpolateZ, InterpolateW, AVarCount, PVarCount); the vertex program writes the
# same counts (a clipped endpoint leaves the others undefined, PolyClipper.h:34-48)
LP_PROGS = [(O.PS_LOG, True, True, 3, 2), (O.PS_LOG, False, False, 2, 2),
            (O.PS_HASH, True, True, 4, 4), (O.PS_DEPTH_CHECKER, True, True, 3, 2)]


def _lp_pipe(w, h, prog, z, wv, a, p, **kw):
    return O.Pipeline(w, h, ps_program=prog, ps_z=z, ps_w=wv, ps_a=a, ps_p=p, **kw)


@pytest.mark.parametrize("prog", LP_PROGS, ids=lambda t: f"p{t[0]}z{int(t[1])}w{int(t[2])}a{t[3]}p{t[4]}")
@pytest.mark.parametrize("mode", [O.DRAW_LINE, O.DRAW_POINT])
def test_lines_points_clip(cpu_oracle, cpu_ref, prog, mode):
    sc = scenes.clip_scene(ntri=1500, seed=31, width=160, height=120, navar=prog[3],
                           npvar=prog[4])
    idx = scenes.edge_indices(sc.indices) if mode == O.DRAW_LINE else sc.indices
    pipe = _lp_pipe(160, 120, *prog, vs_program=O.VS_RAW, vs_counts=(1, prog[3], prog[4]))
    cap = 400_000 if prog[0] == O.PS_LOG else 0
    a = run_cpu(cpu_ref, pipe, sc, mode=mode, indices=idx, log_capacity=cap)
    b = run_cpu(cpu_oracle, pipe, sc, mode=mode, indices=idx, log_capacity=cap)
    if prog[0] == O.PS_LOG:
        assert int(a.log_count[0]) > 1000
    assert_same(a, b, f"mode={mode} prog={prog}")


@pytest.mark.parametrize("mode", [O.DRAW_LINE, O.DRAW_POINT])
def test_wire_sphere(cpu_oracle, cpu_ref, sphere, mode):
    idx = scenes.edge_indices(sphere.indices) if mode == O.DRAW_LINE else sphere.indices
    for prog in (O.PS_DEPTH_CHECKER, O.PS_HASH):
        pipe = O.Pipeline(640, 480, mvp=sphere.mvp, ps_program=prog,
                          viewport=(-20, 10, 640, 480), scissor=(5, 3, 600, 470))
        a = run_cpu(cpu_ref, pipe, sphere, mode=mode, indices=idx)
        b = run_cpu(cpu_oracle, pipe, sphere, mode=mode, indices=idx)
        assert_same(a, b, f"wire sphere mode={mode} prog={prog}")


@pytest.mark.parametrize("grid", [False, True])
@pytest.mark.parametrize("prim", ["line", "point"])
def test_line_point_list(cpu_oracle, cpu_ref, grid, prim):
    v = scenes.window_lines(500, seed=7 + grid, width=96, height=64, grid=grid)
    idx = np.arange(v.size, dtype=np.int32)
    idx[::8] = -1  # skipped primitives: a line by its first index (Rasterizer.h:119, 128)
    for scissor in ((0, 0, 96, 64), (-8, -8, 120, 90), (10, 6, 50, 40)):
        pipe = O.Pipeline(96, 64, scissor=scissor, ps_program=O.PS_LOG, ps_a=4, ps_p=3)
        a = run_cpu(cpu_ref, pipe, vertices=v, indices=idx, prim=prim, log_capacity=200_000)
        b = run_cpu(cpu_oracle, pipe, vertices=v, indices=idx, prim=prim, log_capacity=200_000)
        assert int(a.log_count[0]) > 100
        assert_same(a, b, f"{prim} list scissor={scissor}")


# ---- the reference's demos, headless ----------------------------------------------------------

def test_vertex_processor_test_scene(cpu_oracle, cpu_ref):
    """VertexProcessorTest.cpp:70-118: RGB shader without z / w, viewport (100, 100, 440,
    280), cull None."""
    sc = scenes.vertex_processor_test_scene()
    for prog in (O.PS_RGB, O.PS_LOG):
        pipe = O.Pipeline(640, 480, cull_mode=O.CULL_NONE, viewport=(100, 100, 440, 280),
                          vs_program=O.VS_POS_RGB, vs_counts=(1, 3, 0), ps_program=prog,
                          ps_z=False, ps_w=False, ps_a=3, ps_p=0)
        a = run_cpu(cpu_ref, pipe, sc, log_capacity=200_000)
        b = run_cpu(cpu_oracle, pipe, sc, log_capacity=200_000)
        assert_same(a, b, f"VertexProcessorTest prog={prog}")


@pytest.mark.parametrize("angle", [0.0, 0.6, 2.2])
def test_box_scene(cpu_oracle, cpu_ref, angle):
    """Box.cpp:152-196 with an OBJ crate: Span, cull CW, MVP vertex program with the texture
    coordinate as 2 perspective varyings, InterpolateW (the texture program's inputs)."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "obj", "crate.obj")
    sc = scenes.obj_scene(path, angle)
    pipe = O.Pipeline(640, 480, mvp=sc.mvp, vs_program=O.VS_MVP_POS_UV, vs_counts=(1, 0, 2),
                      ps_program=O.PS_LOG, ps_z=False, ps_w=True, ps_a=0, ps_p=2)
    a = run_cpu(cpu_ref, pipe, sc, log_capacity=400_000)
    b = run_cpu(cpu_oracle, pipe, sc, log_capacity=400_000)
    assert int(a.log_count[0]) > 10_000
    assert_same(a, b, f"box angle={angle}")
