"""Golden fixtures written from the reference itself (tests/golden/make_golden.py):
the oracle must reproduce each one bit for bit (CPU tests), and so must the gfx950 path
(GPU tests).

Fixtures: triangle-setup coefficients (TriangleEquations.h:47-71) for 1,200 triangles
including ties, axis-aligned edges and slivers; clipper polygons (PolyClipper.cpp:45-81)
for 600 clip-space triangles including vertices and edges exactly on planes; image
digests of the sphere, a 20k-triangle 1080p soup and the reference Benchmark scene; full
invocation logs (per pixel, in order) of batch-boundary fans, window-space lists, cull
modes and the default scissor.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle as O
from softwarerenderer_amd import scenes
from helpers import assert_same, run_cpu, run_gpu, sorted_log

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = [O.RASTER_SPAN, O.RASTER_BLOCK, O.RASTER_ADAPTIVE]


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).view(np.uint8).tobytes())
    return h.hexdigest()


def scene_digest(sc):
    return digest(sc.vertices, sc.indices, sc.mvp)


# ---- CPU: the oracle reproduces the reference's vectors ---------------------------------------

def test_setup_coefficients(cpu_oracle):
    d = load("setup.npz")
    v, coeffs, valid = d["vertices"], d["coeffs"], d["valid"]
    na, np_ = int(d["na"]), int(d["np"])
    # both outcomes covered: area2 <= 0 stops after the edges (10 floats), else the planes
    assert (valid == 10).sum() > 100 and (valid > 10).sum() > 100
    for i in range(len(v)):
        r = cpu_oracle.triangle_setup(v[i, 0], v[i, 1], v[i, 2], na, np_)
        assert r.size == valid[i]
        assert np.array_equal(r.view(np.uint32), coeffs[i, :r.size].view(np.uint32)), i


def test_clip_polygons(cpu_oracle):
    d = load("clip.npz")
    v, poly, count = d["vertices"], d["poly"], d["count"]
    na, np_ = int(d["na"]), int(d["np"])
    assert (count < 3).sum() > 10 and (count > 3).sum() > 100
    for i in range(len(v)):
        p, n = cpu_oracle.clip_triangle(v[i, 0], v[i, 1], v[i, 2], na, np_, poly.shape[1])
        assert n == count[i], i
        ref = poly[i, :len(p)]
        for f in ("x", "y", "z", "w"):
            assert np.array_equal(p[f].view(np.uint32), ref[f].view(np.uint32)), (i, f)
        assert np.array_equal(p["avar"][:, :na].view(np.uint32), ref["avar"][:, :na].view(np.uint32))
        assert np.array_equal(p["pvar"][:, :np_].view(np.uint32), ref["pvar"][:, :np_].view(np.uint32))


def _scene_cases():
    sph = scenes.sphere_scene()
    soup = scenes.soup_scene(20_000, seed=7)
    bench = scenes.benchmark_scene()
    cases = []
    for mode in MODES:
        for prog in (O.PS_HASH, O.PS_DEPTH_CHECKER):
            cases.append((f"sphere_m{mode}_p{prog}", sph,
                          O.Pipeline(640, 480, raster_mode=mode, mvp=sph.mvp, ps_program=prog)))
    for mode in MODES:
        cases.append((f"soup20k_m{mode}", soup, O.Pipeline(1920, 1080, raster_mode=mode, mvp=soup.mvp)))
    cases.append(("benchmark", bench,
                  O.Pipeline(640, 480, raster_mode=O.RASTER_SPAN, cull_mode=O.CULL_NONE,
                             vs_program=O.VS_POS_RGB, vs_counts=(1, 3, 0), ps_a=3, ps_p=0,
                             ps_z=False, ps_w=False)))
    return sph, soup, bench, cases


@pytest.fixture(scope="module")
def scene_cases():
    d = load("scenes.npz")
    sph, soup, bench, cases = _scene_cases()
    assert scene_digest(sph) == str(d["sphere_input"]), "sphere generator changed"
    assert scene_digest(soup) == str(d["soup20k_input"]), "soup generator changed"
    assert scene_digest(bench) == str(d["benchmark_input"]), "benchmark generator changed"
    return d, cases


def test_scene_digests(cpu_oracle, scene_cases):
    d, cases = scene_cases
    for name, sc, pipe in cases:
        t = run_cpu(cpu_oracle, pipe, sc)
        assert digest(t.t0, t.t1) == str(d[name]), name
    assert int(d["benchmark_frags"]) == 240_235_639


# ---- the benchmark's 1M-triangle 1080p frame ------------------------------------------------

FRAME1M_CASES = [(mode, prog) for mode in (O.RASTER_SPAN, O.RASTER_BLOCK)
                 for prog in (O.PS_DEPTH_CHECKER, O.PS_HASH)]


@pytest.fixture(scope="module")
def frame1m():
    d = load("frame1m.npz")
    soup = scenes.soup_scene(1_000_000, seed=2)
    assert scene_digest(soup) == str(d["input"]), "soup generator changed"
    return d, soup


def test_frame1m_digests(cpu_oracle, frame1m):
    """The oracle draws the whole benchmark frame (1M triangles, 1080p) as the reference
    did, in Span and Block, with the depth-checker and the order-sensitive hash program."""
    d, soup = frame1m
    for mode, prog in FRAME1M_CASES:
        pipe = O.Pipeline(1920, 1080, raster_mode=mode, mvp=soup.mvp, ps_program=prog)
        t = run_cpu(cpu_oracle, pipe, soup)
        assert digest(t.t0, t.t1) == str(d[f"m{mode}_p{prog}"]), (mode, prog)
        if prog == O.PS_HASH:
            assert int(t.t1.astype(np.int64).sum()) == int(d[f"m{mode}_frags"])


@pytest.mark.gpu
@pytest.mark.parametrize("mode,prog", FRAME1M_CASES)
def test_gpu_frame1m(gpu, cpu_oracle, frame1m, mode, prog):
    """The gfx950 path on the benchmark's full frame: word for word against the oracle on
    the same inputs, and against the reference's digest.  At this size the per-tile lists
    run long (about 52 primitives per 8x8 tile on average) and the first synchronous draw
    grows its arenas and draws again."""
    d, soup = frame1m
    pipe = O.Pipeline(1920, 1080, raster_mode=mode, mvp=soup.mvp, ps_program=prog)
    b = run_gpu(pipe, soup)
    assert digest(b.t0, b.t1) == str(d[f"m{mode}_p{prog}"]), (mode, prog)
    a = run_cpu(cpu_oracle, pipe, soup)
    assert_same(a, b, f"frame1m mode={mode} prog={prog}")


def _log_cases():
    sph = scenes.sphere_scene()
    clip = scenes.clip_scene(ntri=3 * 1024 + 517, seed=21, width=160, height=120,
                             shrink=(0.01, 0.2))
    d = load("logs.npz")
    cases = []
    for mode in (O.RASTER_SPAN, O.RASTER_BLOCK):
        cases.append((f"sphere_log_m{mode}", dict(scene=sph),
                      O.Pipeline(640, 480, raster_mode=mode, mvp=sph.mvp, ps_program=O.PS_LOG)))
    for mode in MODES:
        cases.append((f"batch_m{mode}", dict(scene=clip),
                      O.Pipeline(160, 120, raster_mode=mode, cull_mode=O.CULL_NONE,
                                 vs_program=O.VS_RAW, vs_counts=(1, 3, 2), ps_program=O.PS_LOG)))
    v = d["list_vertices"]
    for mode in MODES:
        cases.append((f"list_m{mode}", dict(vertices=v, indices=np.arange(v.size, dtype=np.int32)),
                      O.Pipeline(96, 64, raster_mode=mode, ps_program=O.PS_LOG, ps_a=3, ps_p=2)))
    cv = d["cull_vertices"]
    csc = scenes.Scene(cv, np.arange(6, dtype=np.int32), np.eye(4, dtype=np.float32), 64, 48)
    for cull in (O.CULL_NONE, O.CULL_CCW, O.CULL_CW):
        for mode in MODES:
            cases.append((f"cull{cull}_m{mode}", dict(scene=csc),
                          O.Pipeline(64, 48, raster_mode=mode, cull_mode=cull, vs_program=O.VS_RAW,
                                     vs_counts=(1, 3, 2), ps_program=O.PS_LOG)))
    cases.append(("default_scissor", dict(scene=csc),
                  O.Pipeline(64, 48, raster_mode=O.RASTER_SPAN, cull_mode=O.CULL_NONE,
                             vs_program=O.VS_RAW, vs_counts=(1, 3, 2), ps_program=O.PS_LOG,
                             scissor=(0, 0, 0, 0))))
    assert scene_digest(clip) == str(d["batch_input"]), "clip scene generator changed"
    return d, cases


def _check_log(d, name, rec):
    if name + "_sha" in d.files:
        assert len(rec) == int(d[name + "_n"]), name
        k = len(d[name + "_head"])
        assert np.array_equal(rec[:k, :10], d[name + "_head"]), name
        assert np.array_equal(rec[:k, 23:25], d[name + "_headp"]), name
        assert digest(rec) == str(d[name + "_sha"]), name
    else:
        assert np.array_equal(rec, d[name]), name


def _draw_log(draw, kw, pipe):
    if "vertices" in kw:
        return sorted_log(draw(pipe, vertices=kw["vertices"], indices=kw["indices"],
                               log_capacity=300_000))
    return sorted_log(draw(pipe, kw["scene"], log_capacity=300_000))


def test_invocation_logs(cpu_oracle):
    d, cases = _log_cases()
    for name, kw, pipe in cases:
        rec = _draw_log(lambda *a, **k: run_cpu(cpu_oracle, *a, **k), kw, pipe)
        _check_log(d, name, rec)
        if name.startswith("sphere_log"):
            m = name[-1]
            crop = rec[(rec[:, 0] >= 288) & (rec[:, 0] < 352) & (rec[:, 1] >= 208) & (rec[:, 1] < 272)]
            assert np.array_equal(crop, d[f"sphere_crop_m{m}"])
    assert len(d["default_scissor"]) == 0  # Rasterizer.h:67-72: nothing drawn by default


# ---- GPU: the gfx950 path reproduces the same vectors ------------------------------------------

@pytest.mark.gpu
def test_gpu_scene_digests(gpu, scene_cases):
    d, cases = scene_cases
    for name, sc, pipe in cases:
        t = run_gpu(pipe, sc)
        assert digest(t.t0, t.t1) == str(d[name]), name


@pytest.mark.gpu
def test_gpu_invocation_logs(gpu):
    d, cases = _log_cases()
    for name, kw, pipe in cases:
        _check_log(d, name, _draw_log(run_gpu, kw, pipe))


# ---- points and lines (tests/golden/lines.npz) ------------------------------------------------

def _line_cases():
    import sys
    sys.path.insert(0, GOLDEN)
    from make_golden import line_cases
    d = load("lines.npz")
    cases, inputs = line_cases()
    for k, v in inputs.items():
        assert v == str(d["input_" + k]), f"{k} input generator changed"
    return d, cases


def _check_line_case(d, name, draw, kw, pipe):
    if name.endswith("_log"):
        _check_log(d, name, sorted_log(draw(pipe, log_capacity=300_000, **kw)))
    else:
        t = draw(pipe, **kw)
        assert digest(t.t0, t.t1) == str(d[name]), name


def test_line_point_fixtures(cpu_oracle):
    d, cases = _line_cases()
    for name, kw, pipe in cases:
        _check_line_case(d, name, lambda *a, **k: run_cpu(cpu_oracle, *a, **k), kw, pipe)
    assert int(d["clip_line_log_n"]) > 100_000 and int(d["list_point_log_n"]) > 400


@pytest.mark.gpu
def test_gpu_line_point_fixtures(gpu):
    d, cases = _line_cases()
    for name, kw, pipe in cases:
        _check_line_case(d, name, run_gpu, kw, pipe)
