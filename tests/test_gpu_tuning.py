"""Launch-shape tuning: every swr_tuning setting draws the same bits as the oracle, invalid
settings are refused, and the autotuner leaves a valid setting behind."""
import numpy as np
import pytest

import oracle as O
from softwarerenderer_amd import scenes
from helpers import assert_same, run_cpu, run_gpu

pytestmark = pytest.mark.gpu

TUNINGS = [
    dict(raster_waves=1, raster_group=1, xcd_order=-1, expand_chunk=256, vertex_pass=1),
    dict(raster_waves=2, raster_group=2, xcd_order=1, expand_chunk=512, vertex_pass=-1),
    dict(raster_waves=8, raster_group=4, xcd_order=1, expand_chunk=2048, bin_size=16,
         waves_per_bin=1),
    dict(raster_waves=4, raster_group=8, xcd_order=-1, expand_chunk=4096, bin_size=32,
         waves_per_bin=4),
    dict(bin_size=16, waves_per_bin=4, raster_group=2, vertex_pass=1),
    dict(bin_size=32, waves_per_bin=2, expand_chunk=256),
    # persistent raster workgroups taking bins from the per-XCD queues
    dict(raster_waves=2, raster_queue=1),
    dict(bin_size=16, waves_per_bin=4, expand_chunk=512, raster_queue=1),
    dict(bin_size=32, waves_per_bin=2, raster_group=2, raster_queue=1),
]
KNOB_DEFAULTS = dict(raster_waves=0, raster_group=0, xcd_order=0, expand_chunk=0, vertex_pass=0,
                     bin_size=0, waves_per_bin=0, raster_queue=0)


def _tid(t):
    return "-".join(f"{k[:3]}{v}" for k, v in t.items())


@pytest.fixture(scope="module")
def sphere():
    return scenes.sphere_scene()


@pytest.fixture(scope="module")
def clip():
    return scenes.clip_scene(ntri=3000, seed=3)


def _rasterizer(tune):
    from softwarerenderer_amd import api
    r = api.Rasterizer(0)
    r.setTuning(**tune)
    assert r.tuning() == dict(KNOB_DEFAULTS, **tune)
    return r


@pytest.mark.parametrize("tune", TUNINGS, ids=_tid)
@pytest.mark.parametrize("prog", [O.PS_LOG, O.PS_DEPTH_CHECKER, O.PS_VISIBILITY])
def test_sphere_tuned(gpu, cpu_oracle, sphere, tune, prog):
    pipe = O.Pipeline(640, 480, raster_mode=O.RASTER_BLOCK, mvp=sphere.mvp, ps_program=prog)
    cap = 100_000 if prog == O.PS_LOG else 0
    a = run_cpu(cpu_oracle, pipe, sphere, log_capacity=cap)
    r = _rasterizer(tune)
    try:
        b = run_gpu(pipe, sphere, log_capacity=cap, rasterizer=r)
    finally:
        r.close()
    assert_same(a, b, f"sphere prog={prog} tuning={tune}")


@pytest.mark.parametrize("tune", TUNINGS, ids=_tid)
@pytest.mark.parametrize("mode", [O.RASTER_SPAN, O.RASTER_ADAPTIVE])
def test_sphere_span_tuned(gpu, cpu_oracle, sphere, tune, mode):
    """Span and Adaptive rows under every bin size: the hash program sees every fragment's
    values in order."""
    pipe = O.Pipeline(640, 480, raster_mode=mode, mvp=sphere.mvp, ps_program=O.PS_HASH,
                      scissor=(13, 7, 600, 450))
    a = run_cpu(cpu_oracle, pipe, sphere)
    r = _rasterizer(tune)
    try:
        b = run_gpu(pipe, sphere, rasterizer=r)
    finally:
        r.close()
    assert_same(a, b, f"sphere span mode={mode} tuning={tune}")


@pytest.mark.parametrize("tune", TUNINGS, ids=_tid)
def test_clip_tuned(gpu, cpu_oracle, clip, tune):
    pipe = O.Pipeline(clip.width, clip.height, raster_mode=O.RASTER_SPAN, cull_mode=O.CULL_NONE,
                      vs_program=O.VS_RAW, vs_counts=(1, 3, 2), ps_program=O.PS_HASH)
    a = run_cpu(cpu_oracle, pipe, clip)
    r = _rasterizer(tune)
    try:
        b = run_gpu(pipe, clip, rasterizer=r)
    finally:
        r.close()
    assert_same(a, b, f"clip tuning={tune}")


def test_soup_tuned(gpu, cpu_oracle):
    sc = scenes.soup_scene(40_000, seed=9)
    pipe = O.Pipeline(1920, 1080, raster_mode=O.RASTER_BLOCK, mvp=sc.mvp,
                      ps_program=O.PS_DEPTH_CHECKER)
    a = run_cpu(cpu_oracle, pipe, sc)
    for tune in TUNINGS:
        r = _rasterizer(tune)
        try:
            b = run_gpu(pipe, sc, rasterizer=r)
        finally:
            r.close()
        assert_same(a, b, f"soup tuning={tune}")


@pytest.mark.parametrize("bad", [dict(raster_waves=3), dict(raster_waves=16),
                                 dict(raster_group=5), dict(expand_chunk=128),
                                 dict(expand_chunk=3000), dict(xcd_order=2),
                                 dict(vertex_pass=2), dict(bin_size=24), dict(bin_size=64),
                                 dict(waves_per_bin=8), dict(waves_per_bin=3),
                                 dict(raster_queue=2), dict(raster_queue=-2)])
def test_invalid_tuning(gpu, bad):
    from softwarerenderer_amd import _abi, api
    r = api.Rasterizer(0)
    try:
        before = r.tuning()
        with pytest.raises(_abi.SwrError):
            r.setTuning(**bad)
        assert r.tuning() == before
    finally:
        r.close()


def test_autotune_small(gpu, cpu_oracle, sphere):
    import torch
    from softwarerenderer_amd import api
    from softwarerenderer_amd.autotune import autotune
    W, H = 640, 480
    color = torch.zeros((H, W), dtype=torch.int32, device=gpu)
    depth = torch.ones((H, W), dtype=torch.float32, device=gpu)
    r = api.Rasterizer(0)
    try:
        v = api.VertexProcessor(r)
        r.setRasterMode(api.RasterMode.Block)
        r.setScissorRect(0, 0, W, H)
        r.setPixelShader(api.DepthCheckerShader(color, depth))
        v.setViewport(0, 0, W, H)
        v.setCullMode(api.CullMode.CW)
        v.setVertexShader(api.MvpVertexShader(sphere.mvp))
        v.setVertexAttribPointer(0, sphere.stride, torch.from_numpy(sphere.vertices).to(gpu))
        idx = torch.from_numpy(sphere.indices).to(gpu)

        def draw():
            color.zero_()
            depth.fill_(1.0)
            v.drawElementsAsync(api.DrawMode.Triangle, idx)

        v.drawElements(api.DrawMode.Triangle, idx)
        res = autotune(r, draw, space={"raster_waves": (2, 4), "raster_group": (0, 2),
                                       "bin_size": (8, 16)}, reps=2, warmup=1, rounds=1)
        assert len(res["table"]) == 2 * 2 * 2
        assert r.tuning() == res["best"]
        assert res["ms"] > 0
        draw()
        r.synchronize()
        pipe = O.Pipeline(W, H, raster_mode=O.RASTER_BLOCK, mvp=sphere.mvp,
                          ps_program=O.PS_DEPTH_CHECKER)
        a = run_cpu(cpu_oracle, pipe, sphere)
        assert np.array_equal(a.t0.view(np.uint32), color.cpu().numpy().view(np.uint32))
        assert np.array_equal(a.t1.view(np.uint32), depth.cpu().numpy().view(np.uint32))
    finally:
        r.close()
