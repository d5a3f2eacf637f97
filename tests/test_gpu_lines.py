"""GPU parity of points and lines: drawElements(Point / Line) and the IRasterizer line and
point lists against the oracle, bit for bit (targets as raw words, full invocation logs).

The oracle's line and point paths are pinned against the reference build in
test_oracle_vs_ref.py (test_lines_points_clip, test_wire_sphere, test_line_point_list) and
by the golden logs in test_golden.py."""
import numpy as np
import pytest

import oracle as O
from softwarerenderer_amd import scenes
from helpers import assert_same, run_cpu, run_gpu

pytestmark = pytest.mark.gpu

PROGS = [(O.PS_LOG, True, True, 3, 2), (O.PS_LOG, False, False, 2, 2),
         (O.PS_HASH, True, True, 4, 4), (O.PS_HASH, True, False, 16, 16),
         (O.PS_DEPTH_CHECKER, True, True, 3, 2), (O.PS_RGB, False, False, 3, 0),
         (O.PS_COVERAGE, False, False, 0, 0), (O.PS_VISIBILITY, True, False, 0, 0)]


def _ids(t):
    return f"p{t[0]}z{int(t[1])}w{int(t[2])}a{t[3]}p{t[4]}"


def _pipe(w, h, prog, z, wv, a, p, **kw):
    return O.Pipeline(w, h, ps_program=prog, ps_z=z, ps_w=wv, ps_a=a, ps_p=p, **kw)


@pytest.fixture(scope="module")
def sphere():
    return scenes.sphere_scene()


@pytest.mark.parametrize("prog", PROGS, ids=_ids)
@pytest.mark.parametrize("mode", [O.DRAW_LINE, O.DRAW_POINT])
def test_clip_lines_points(gpu, cpu_oracle, prog, mode):
    na, np_ = max(prog[3], 1), max(prog[4], 1)
    sc = scenes.clip_scene(ntri=1500, seed=31, width=160, height=120, navar=na, npvar=np_)
    idx = scenes.edge_indices(sc.indices) if mode == O.DRAW_LINE else sc.indices
    pipe = _pipe(160, 120, *prog, vs_program=O.VS_RAW, vs_counts=(1, na, np_))
    cap = 400_000 if prog[0] == O.PS_LOG else 0
    a = run_cpu(cpu_oracle, pipe, sc, mode=mode, indices=idx, log_capacity=cap)
    b = run_gpu(pipe, sc, mode=mode, indices=idx, log_capacity=cap)
    assert_same(a, b, f"mode={mode} prog={prog}")


@pytest.mark.parametrize("mode", [O.DRAW_LINE, O.DRAW_POINT])
@pytest.mark.parametrize("prog", [O.PS_DEPTH_CHECKER, O.PS_HASH, O.PS_LOG])
def test_wire_sphere(gpu, cpu_oracle, sphere, mode, prog):
    idx = scenes.edge_indices(sphere.indices) if mode == O.DRAW_LINE else sphere.indices
    pipe = O.Pipeline(640, 480, mvp=sphere.mvp, ps_program=prog,
                      viewport=(-20, 10, 640, 480), scissor=(5, 3, 600, 470))
    cap = 400_000 if prog == O.PS_LOG else 0
    a = run_cpu(cpu_oracle, pipe, sphere, mode=mode, indices=idx, log_capacity=cap)
    b = run_gpu(pipe, sphere, mode=mode, indices=idx, log_capacity=cap)
    if prog == O.PS_LOG:
        assert int(a.log_count[0]) > 10_000
    assert_same(a, b, f"wire sphere mode={mode} prog={prog}")


@pytest.mark.parametrize("grid", [False, True])
@pytest.mark.parametrize("prim", ["line", "point"])
@pytest.mark.parametrize("scissor", [(0, 0, 96, 64), (-8, -8, 120, 90), (10, 6, 50, 40)])
def test_line_point_list(gpu, cpu_oracle, grid, prim, scissor):
    v = scenes.window_lines(500, seed=7 + grid, width=96, height=64, grid=grid)
    idx = np.arange(v.size, dtype=np.int32)
    idx[::8] = -1  # skipped primitives: a line by its first index (Rasterizer.h:119, 128)
    pipe = O.Pipeline(96, 64, scissor=scissor, ps_program=O.PS_LOG, ps_a=4, ps_p=3)
    a = run_cpu(cpu_oracle, pipe, vertices=v, indices=idx, prim=prim, log_capacity=200_000)
    b = run_gpu(pipe, vertices=v, indices=idx, prim=prim, log_capacity=200_000)
    assert int(a.log_count[0]) > 100
    assert_same(a, b, f"{prim} list scissor={scissor}")


def test_many_lines_one_pixel(gpu, cpu_oracle):
    """Thousands of lines through a few pixels: long per-tile fragment lists, ordered."""
    rng = np.random.default_rng(4)
    n = 3000
    v = np.zeros(2 * n, O.RASTERIZER_VERTEX)
    c = np.array([20.5, 17.5]) + rng.uniform(-1.5, 1.5, (n, 2))
    d = rng.normal(0, 4, (n, 2))
    p = np.stack([c - d, c + d], 1).reshape(-1, 2)
    v["x"], v["y"] = p[:, 0], p[:, 1]
    v["z"] = rng.random(2 * n)
    v["w"] = rng.uniform(0.5, 2, 2 * n)
    v["avar"] = rng.random((2 * n, 16))
    v["pvar"] = rng.random((2 * n, 16))
    idx = np.arange(2 * n, dtype=np.int32)
    for prog in (O.PS_LOG, O.PS_HASH, O.PS_DEPTH_CHECKER):
        pipe = O.Pipeline(40, 32, scissor=(0, 0, 40, 32), ps_program=prog)
        cap = 200_000 if prog == O.PS_LOG else 0
        a = run_cpu(cpu_oracle, pipe, vertices=v, indices=idx, prim="line", log_capacity=cap)
        b = run_gpu(pipe, vertices=v, indices=idx, prim="line", log_capacity=cap)
        assert_same(a, b, f"dense lines prog={prog}")


def test_batches_and_partial(gpu, cpu_oracle):
    """More than one 1024-primitive batch (keys (batch << 15) + slot), an odd index count
    (the partial last line is dropped) and host-side buffers."""
    sc = scenes.clip_scene(ntri=2100, seed=8, width=128, height=96, spread=1.1)
    idx = scenes.edge_indices(sc.indices)[: 2 * 3000 + 1]
    pipe = O.Pipeline(128, 96, vs_program=O.VS_RAW, vs_counts=(1, 3, 2), ps_program=O.PS_LOG)
    a = run_cpu(cpu_oracle, pipe, sc, mode=O.DRAW_LINE, indices=idx, log_capacity=300_000)
    for dev in (True, False):
        b = run_gpu(pipe, sc, mode=O.DRAW_LINE, indices=idx, log_capacity=300_000,
                    device_buffers=dev)
        assert_same(a, b, f"line batches device_buffers={dev}")
    a = run_cpu(cpu_oracle, pipe, sc, mode=O.DRAW_POINT, indices=idx, log_capacity=300_000)
    b = run_gpu(pipe, sc, mode=O.DRAW_POINT, indices=idx, log_capacity=300_000)
    assert_same(a, b, "point batches")


def test_empty_and_errors(gpu):
    import torch
    from softwarerenderer_amd import _abi, api
    r = api.Rasterizer(0)
    try:
        buf = torch.zeros((8, 8), dtype=torch.int32, device=gpu)
        r.setScissorRect(0, 0, 8, 8)
        r.setPixelShader(api.CoverageShader(buf))
        v = np.zeros(2, O.RASTERIZER_VERTEX)
        v["x"], v["y"], v["w"] = [1.0, 3.0e7], [1.0, 2.0], 1.0
        vt = torch.from_numpy(v.view(np.uint8)).to(gpu)
        r.drawLineList(vt, torch.zeros(0, dtype=torch.int32, device=gpu))   # nothing to draw
        r.drawPointList(vt, torch.tensor([-1, -1], dtype=torch.int32, device=gpu))
        assert int(buf.sum()) == 0
        with pytest.raises(_abi.SwrError) as e:  # 3e7 steps: beyond SWR_MAX_LINE_STEPS
            r.drawLineList(vt, torch.tensor([0, 1], dtype=torch.int32, device=gpu))
        assert e.value.code == -5
        with pytest.raises(_abi.SwrError) as e:  # index past the vertex array
            r.drawPointList(vt, torch.tensor([0, 5], dtype=torch.int32, device=gpu))
        assert e.value.code == -6
        assert int(buf.sum()) == 0
        r.drawPointList(vt, torch.tensor([0], dtype=torch.int32, device=gpu))
        assert int(buf.sum()) == 1 and int(buf[1, 1]) == 1
    finally:
        r.close()
