"""Batched draws (swr_batch_create / swr_batch_draw): many scenes with their own state and
targets in one pass of the pipeline.  Each scene must come out exactly as its own
drawElements call (and the oracle) draws it: the order-sensitive hash program checks
coverage, every interpolated value and the per-pixel order of invocations."""
import numpy as np
import pytest

import oracle as O
from helpers import assert_same, run_cpu, run_gpu
from softwarerenderer_amd import scenes as S

pytestmark = pytest.mark.gpu


def pipe_for(sc, mode, cull=O.CULL_CW, scissor=None, vs=O.VS_MVP_POS_RGB_UV, vs_counts=(1, 3, 2),
             ps=O.PS_HASH):
    return O.Pipeline(width=sc.width, height=sc.height, raster_mode=mode, cull_mode=cull,
                      scissor=scissor, vs_program=vs, vs_counts=vs_counts, mvp=sc.mvp,
                      ps_program=ps, ps_z=True, ps_w=ps != O.PS_VISIBILITY,
                      ps_a=3 if ps != O.PS_VISIBILITY else 0,
                      ps_p=2 if ps != O.PS_VISIBILITY else 0)


def run_batch(cases, rasterizer=None, draws=1):
    """cases: [(scene, pipe)] sharing vertex and pixel programs.  Returns Targets per case."""
    import torch
    from softwarerenderer_amd import api

    def tens(a):
        if a is None:
            return None
        return torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a.copy()).cuda()

    r = rasterizer or api.Rasterizer(0)
    p0 = cases[0][1]
    ps = api.PixelShaderBase()
    ps.Program = p0.ps_program
    ps.InterpolateZ, ps.InterpolateW = p0.ps_z, p0.ps_w
    ps.AVarCount, ps.PVarCount = p0.ps_a, p0.ps_p
    r.setPixelShader(ps)
    vp = api.VertexProcessor(r)
    vs = api.VertexShaderBase()
    vs.Program = p0.vs_program
    vs.AttribCount, vs.AVarCount, vs.PVarCount = p0.vs_counts
    vp.setVertexShader(vs)
    descs, outs = [], []
    for sc, pipe in cases:
        init = pipe.targets()
        t0, t1 = tens(init.t0), tens(init.t1)
        outs.append((init, t0, t1))
        descs.append({
            "indices": torch.from_numpy(np.ascontiguousarray(sc.indices, np.int32)).cuda(),
            "vertices": torch.from_numpy(np.ascontiguousarray(sc.vertices)).cuda(),
            "stride": sc.stride, "mvp": pipe.mvp, "viewport": pipe.viewport,
            "depth_range": pipe.depth_range, "cull_mode": pipe.cull_mode,
            "raster_mode": pipe.raster_mode, "scissor": pipe.scissor,
            "targets": (t0, t1), "size": (pipe.height, pipe.width)})
    b = api.Batch(r, descs)
    for _ in range(draws):
        b.draw()
    res = []
    for init, t0, t1 in outs:
        out = O.Targets(init.program, init.width, init.height)
        out.t0 = t0.cpu().numpy().view(init.t0.dtype).copy()
        out.t1 = t1.cpu().numpy().view(init.t1.dtype).copy() if t1 is not None else None
        res.append(out)
    b.close()
    if rasterizer is None:
        r.close()
    return res


def mvp_cases(ps=O.PS_HASH):
    sph = S.sphere_scene(64, 128, seed=1, width=640, height=480)
    sph_small = S.sphere_scene(16, 32, seed=2, width=256, height=256)
    sph_big = S.sphere_scene(24, 48, seed=3, width=1024, height=1024)
    tiny = S.sphere_scene(3, 4, seed=4, width=64, height=64)
    soup = S.soup_scene(30_000, seed=5, width=1920, height=1080)
    soup_s = S.soup_scene(5_000, seed=6, width=512, height=512)
    empty = S.Scene(sph.vertices, np.zeros(0, np.int32), sph.mvp, 96, 80)
    return [
        (sph, pipe_for(sph, O.RASTER_SPAN, ps=ps)),
        (soup, pipe_for(soup, O.RASTER_BLOCK, ps=ps)),
        (sph_small, pipe_for(sph_small, O.RASTER_BLOCK, cull=O.CULL_NONE, ps=ps)),
        (empty, pipe_for(empty, O.RASTER_SPAN, ps=ps)),
        (sph_big, pipe_for(sph_big, O.RASTER_ADAPTIVE, scissor=(100, 50, 700, 900), ps=ps)),
        (tiny, pipe_for(tiny, O.RASTER_SPAN, ps=ps)),
        (soup_s, pipe_for(soup_s, O.RASTER_SPAN, cull=O.CULL_CCW, ps=ps)),
    ]


def test_batch_matches_single_draws_and_oracle(gpu, cpu_oracle):
    cases = mvp_cases()
    got = run_batch(cases)
    for i, ((sc, pipe), g) in enumerate(zip(cases, got)):
        single = run_gpu(pipe, sc)
        assert_same(single, g, f"scene {i}: batch vs single draw")
        if sc.triangles <= 40_000:
            assert_same(run_cpu(cpu_oracle, pipe, sc), g, f"scene {i}: batch vs oracle")
        assert int(g.t1.sum()) > 0 or sc.triangles == 0


@pytest.mark.parametrize("tune", [dict(bin_size=16, waves_per_bin=2),
                                  dict(bin_size=32, waves_per_bin=4, raster_group=2),
                                  dict(bin_size=16, waves_per_bin=4, raster_queue=1),
                                  dict(raster_waves=4, raster_queue=1)],
                         ids=["bin16", "bin32", "bin16-queue", "bin8-queue"])
@pytest.mark.parametrize("ps", [O.PS_HASH, O.PS_DEPTH_CHECKER])
def test_batch_larger_bins(gpu, cpu_oracle, tune, ps):
    """Every scene of a mixed batch with 16x16 and 32x32 screen bins (each scene's bins
    numbered after the previous scenes'), against the oracle."""
    from softwarerenderer_amd import api
    cases = mvp_cases(ps=ps)
    r = api.Rasterizer(0)
    r.setTuning(**tune)
    try:
        got = run_batch(cases, rasterizer=r)
    finally:
        r.close()
    for i, ((sc, pipe), g) in enumerate(zip(cases, got)):
        if sc.triangles <= 40_000:
            assert_same(run_cpu(cpu_oracle, pipe, sc), g, f"scene {i}: {tune} vs oracle")


def test_batch_depth_checker(gpu, cpu_oracle):
    """The depth-checker program in a batch: each pixel chains its winning fragment's planes
    again when its tile is written, with its own scene's scissor and pitch."""
    cases = mvp_cases(ps=O.PS_DEPTH_CHECKER)
    got = run_batch(cases)
    for i, ((sc, pipe), g) in enumerate(zip(cases, got)):
        ref = run_cpu(cpu_oracle, pipe, sc) if sc.triangles <= 40_000 else run_gpu(pipe, sc)
        assert_same(ref, g, f"scene {i}: depth checker in a batch")
        assert int((g.t0 != 0).sum()) > 0 or sc.triangles == 0


def test_batch_redraw_and_reorder(gpu):
    """Scene order inside a batch changes no scene's result, and drawing a batch twice
    hashes on top of the first draw exactly as a second single draw does."""
    import torch
    from softwarerenderer_amd import api

    cases = mvp_cases()
    base = run_batch(cases)
    rev = run_batch(cases[::-1])[::-1]
    for i in range(len(cases)):
        assert_same(base[i], rev[i], f"scene {i}: reversed batch")
    twice = run_batch(cases, draws=2)
    for i in (0, 2, 5):
        sc, pipe = cases[i]
        r = api.Rasterizer(0)
        t0 = torch.from_numpy(base[i].t0.view(np.int32).copy()).cuda()
        t1 = torch.from_numpy(base[i].t1.view(np.int32).copy()).cuda()
        ps = api.PixelShaderBase()
        ps.Program = pipe.ps_program
        ps.InterpolateZ, ps.InterpolateW = pipe.ps_z, pipe.ps_w
        ps.AVarCount, ps.PVarCount = pipe.ps_a, pipe.ps_p
        ps.targets = lambda: (t0, t1, None, pipe.width, pipe.height, pipe.width)
        r.setRasterMode(pipe.raster_mode)
        r.setScissorRect(*pipe.scissor)
        r.setPixelShader(ps)
        vp = api.VertexProcessor(r)
        vp.setViewport(*pipe.viewport)
        vp.setCullMode(pipe.cull_mode)
        vs = api.VertexShaderBase()
        vs.Program = pipe.vs_program
        vs.AttribCount, vs.AVarCount, vs.PVarCount = pipe.vs_counts
        vs.mvp = pipe.mvp
        vp.setVertexShader(vs)
        vp.setVertexAttribPointer(0, sc.stride, torch.from_numpy(sc.vertices).cuda())
        vp.drawElements(O.DRAW_TRIANGLE, torch.from_numpy(sc.indices).cuda())
        assert np.array_equal(t0.cpu().numpy().view(np.uint32), twice[i].t0), i
        assert np.array_equal(t1.cpu().numpy().view(np.uint32), twice[i].t1), i
        r.close()


@pytest.mark.parametrize("sizes", [(2500, 700, 1024), (3000, 1, 2047)])
def test_batch_clip_scenes_vs_oracle(gpu, cpu_oracle, sizes):
    """Raw clip-space scenes: clipping fans and partial last batches inside a batch."""
    cases = []
    for k, n in enumerate(sizes):
        sc = S.clip_scene(n, seed=40 + k, width=200 + 37 * k, height=150 + 11 * k,
                          shrink=(0.01, 0.3))
        mode = (O.RASTER_SPAN, O.RASTER_BLOCK, O.RASTER_ADAPTIVE)[k % 3]
        cases.append((sc, pipe_for(sc, mode, cull=O.CULL_NONE, vs=O.VS_RAW)))
    got = run_batch(cases)
    for i, ((sc, pipe), g) in enumerate(zip(cases, got)):
        assert_same(run_cpu(cpu_oracle, pipe, sc), g, f"clip scene {i}")


def test_visibility_single_and_batch_vs_oracle(gpu, cpu_oracle):
    """The visibility program: nearest source triangle + 1 and its depth, including the
    fan triangles of clipped input triangles (they report the triangle they came from)."""
    sph = S.sphere_scene(32, 64, seed=7, width=320, height=240)
    clip = S.clip_scene(2200, seed=9, width=256, height=192, shrink=(0.02, 0.5))
    cases = [(sph, pipe_for(sph, O.RASTER_SPAN, ps=O.PS_VISIBILITY)),
             (clip, pipe_for(clip, O.RASTER_BLOCK, cull=O.CULL_NONE, vs=O.VS_RAW,
                             vs_counts=(1, 3, 2), ps=O.PS_VISIBILITY))]
    got = run_batch(cases[:1]) + run_batch(cases[1:])
    for i, ((sc, pipe), g) in enumerate(zip(cases, got)):
        ref = run_cpu(cpu_oracle, pipe, sc)
        assert_same(ref, g, f"visibility scene {i}")
        assert_same(ref, run_gpu(pipe, sc), f"visibility scene {i} single draw")
        assert (ref.t0 > 0).sum() > 1000


def test_many_scene_workload(gpu, cpu_oracle):
    """The 256-scene batching workload at full size (UV spheres of 1k..200k triangles,
    256x256 to 1920x1080, Span and Block; 9.0 M triangles and 202 M pixels): one batched pass
    draws every scene exactly as its own drawElements call does (order-sensitive hash
    program), and the smallest scenes exactly as the oracle."""
    from softwarerenderer_amd import api

    work = S.batch_workload(256, seed=3)
    modes = (O.RASTER_SPAN, O.RASTER_BLOCK)
    cases = [(sc, pipe_for(sc, modes[sc.info["raster_mode"]])) for sc in work]
    got = run_batch(cases)
    r = api.Rasterizer(0)
    try:
        for i, ((sc, pipe), g) in enumerate(zip(cases, got)):
            assert_same(run_gpu(pipe, sc, rasterizer=r), g, f"scene {i}: batch vs single draw")
            assert int(g.t1.astype(np.int64).sum()) > 0, f"scene {i} drew nothing"
    finally:
        r.close()
    small = sorted(range(len(cases)), key=lambda i: cases[i][0].triangles *
                   cases[i][1].width * cases[i][1].height)[:4]
    for i in small:
        sc, pipe = cases[i]
        assert_same(run_cpu(cpu_oracle, pipe, sc), got[i], f"scene {i}: batch vs oracle")
