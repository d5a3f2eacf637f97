"""Benchmark of the gfx950 draw path on the 1080p, 1M-triangle soup (Block mode).

One step = one drawElements of the whole frame (vertex program, clip, cull, setup,
binning, raster, depth-tested checker shading into RGBA8 + fp32 depth) preceded by the
target clear, with vertices and indices resident in HBM.  Prints one JSON line.

The line also carries the differentiation workload (64 views at 512x512 of a 40,962-vertex
icosphere, forward + backward of the image loss 1/2 |image - target|^2): views per second of the batched step
(one visibility draw for all views, captured in a HIP graph) against the same work issued
one view at a time; and the many-scene workload (256 UV spheres of 1k..200k triangles at
256x256..1920x1080 in Span and Block mode): scenes per second of one batched pass against
one draw per scene; and batched differentiation under load (32 fBm heightfields of 8k..2.1M
triangles, 4 views each at 1024x1024): views per second of the batched forward + backward
against one view at a time, and the HIP-graph Adam step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-diff] [--no-batch]
                    [--no-terrain] [--copy-probe]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--ntri", type=int, default=1_000_000)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-runs", type=int, default=3)
    p.add_argument("--copy-probe", action="store_true",
                   help="run one plain 256 MiB device copy first (counter calibration)")
    p.add_argument("--no-diff", action="store_true", help="skip the differentiation workload")
    p.add_argument("--diff-views", type=int, default=64)
    p.add_argument("--diff-size", type=int, default=512)
    p.add_argument("--diff-subdiv", type=int, default=6)
    p.add_argument("--no-batch", action="store_true", help="skip the many-scene workload")
    p.add_argument("--batch-scenes", type=int, default=256)
    p.add_argument("--no-terrain", action="store_true",
                   help="skip the batched differentiation workload under load")
    p.add_argument("--terrain-scenes", type=int, default=32)
    p.add_argument("--terrain-size", type=int, default=1024)
    p.add_argument("--terrain-reserve", type=float, default=8.0,
                   help="tile-pair arena of the terrain workload, in multiples of the target")
    p.add_argument("--no-span", action="store_true",
                   help="skip the Span-mode workloads (sphere, the reference's Benchmark scene)")
    p.add_argument("--no-tune", action="store_true",
                   help="skip the launch-shape autotuner (built-in defaults)")
    p.add_argument("--tuning", default="",
                   help="a fixed launch shape for the 1M frame instead of the autotuner's, "
                        "e.g. bin_size=16,waves_per_bin=4,expand_chunk=512")
    return p.parse_args()


def progress(msg: str) -> None:
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def source_revision() -> str:
    """A digest of the kernel and C-ABI sources (softwarerenderer_amd/csrc): counter figures
    recorded for other sources do not describe this build."""
    d = os.path.join(ROOT, "softwarerenderer_amd", "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(d)):
        with open(os.path.join(d, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def _counters(workload: str) -> dict:
    """The raster kernel's rocprofv3 counter figures recorded in profiles/traffic.json, when
    they were recorded for this workload and these sources (else nothing)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return {}
    if t.get("workload") != workload or t.get("source") != source_revision():
        return {}
    return t


# the knobs that select the raster kernel's launch and work split: counter figures measured
# under one setting of them do not describe another
def raster_shape(tuning: dict) -> dict:
    """The raster launch a tuning selects, with the built-in defaults filled in (0 = default;
    raster_queue -1 and 0 both mean one workgroup per bin)."""
    bins = int(tuning.get("bin_size", 0)) or 8
    shape = {"bin_size": bins, "raster_group": int(tuning.get("raster_group", 0)),
             "raster_queue": 1 if int(tuning.get("raster_queue", 0)) > 0 else 0}
    if bins == 8:
        shape["raster_waves"] = int(tuning.get("raster_waves", 0)) or 4
    else:
        shape["waves_per_bin"] = min(int(tuning.get("waves_per_bin", 0)) or 4, (bins // 8) ** 2)
    return shape


def _counters_for(workload: str, tuning: dict) -> dict:
    """The recorded counters, when they were also taken with this raster launch shape."""
    t = _counters(workload)
    if not t or raster_shape(t.get("tuning", {})) != raster_shape(tuning):
        return {}
    return t


def measured_traffic(workload: str, tuning: dict):
    """HBM bytes per raster launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled
    by the copy-probe factors recorded with them)."""
    return _counters_for(workload, tuning).get("raster_bytes")


def valu_issue(workload: str, tuning: dict, raster_ms: float):
    """The raster kernel's VALU issue rate: SQ_INSTS_VALU per launch (rocprofv3) over the
    launch time measured here, against the wave64 fp32 issue peak."""
    t = _counters_for(workload, tuning)
    insts, peak = t.get("valu_insts_per_launch"), t.get("valu_peak_winst_per_s")
    if not insts or not peak or raster_ms <= 0:
        return None
    rate = insts / (raster_ms * 1e-3)
    return {"achieved": round(rate / 1e9, 1), "peak": round(peak / 1e9, 1),
            "unit": "G wave-instructions/s", "frac": round(rate / peak, 4),
            "insts_per_launch": int(insts)}


def span_bench(args, dev):
    """Span-mode frames: the 640x480 UV sphere (16,128 triangles, CW cull, depth-checker
    program) and the reference's Benchmark scene (Benchmark.cpp:66-105: 40,960 large
    overlapping triangles, no cull, the RGB program, 240 M fragments).  Each is drawn K times
    back to back (targets cleared each time), inputs resident in HBM."""
    import torch

    from softwarerenderer_amd import api, scenes

    out = {}
    stream = torch.cuda.current_stream(dev)
    for name, sc, prog, cull in (
            ("sphere-640x480-span", scenes.sphere_scene(), "depth", api.CullMode.CW),
            ("benchmark-640x480-span", scenes.benchmark_scene(), "rgb", api.CullMode.None_)):
        H, W = sc.height, sc.width
        verts = torch.from_numpy(sc.vertices).to(dev)
        idx = torch.from_numpy(sc.indices).to(dev)
        color = torch.zeros((H, W), dtype=torch.int32, device=dev)
        depth = torch.ones((H, W), dtype=torch.float32, device=dev)
        r = api.Rasterizer(0)
        r.setStream(stream.cuda_stream)
        v = api.VertexProcessor(r)
        r.setRasterMode(api.RasterMode.Span)
        r.setScissorRect(0, 0, W, H)
        v.setViewport(0, 0, W, H)
        v.setCullMode(cull)
        if prog == "depth":
            v.setVertexShader(api.MvpVertexShader(sc.mvp))
        else:
            v.setVertexShader(api.PosRGBVertexShader())
        v.setVertexAttribPointer(0, sc.stride, verts)
        # fragments shaded: one hash-program draw with the same varyings
        hsh = torch.zeros((H, W), dtype=torch.int32, device=dev)
        cnt = torch.zeros((H, W), dtype=torch.int32, device=dev)
        if prog == "depth":
            r.setPixelShader(api.HashShader(hsh, cnt))
        else:
            r.setPixelShader(api.HashShader(hsh, cnt, z=False, w=False, avars=3, pvars=0))
        v.drawElements(api.DrawMode.Triangle, idx)
        frags = int(cnt.sum().item())
        r.setPixelShader(api.DepthCheckerShader(color, depth) if prog == "depth"
                         else api.RGBShader(color))

        def step(sync):
            color.zero_()
            depth.fill_(1.0)
            (v.drawElements if sync else v.drawElementsAsync)(api.DrawMode.Triangle, idx)

        step(True)
        for _ in range(args.warmup):
            step(False)
        r.synchronize()
        steps = max(args.steps, 20)
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0.record(stream)
        for _ in range(steps):
            step(False)
        t1.record(stream)
        r.synchronize()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / steps
        r.setProfiling(True)
        stage = np.zeros(8)
        for _ in range(10):
            step(True)
            stage += np.array(r.stats().stage_ms[:8], dtype=np.float64)
        r.setProfiling(False)
        out[name] = {"triangles": sc.triangles, "fragments": frags,
                     "pixel_program": "depth<, rgb*checker(uv,8)" if prog == "depth" else "rgb",
                     "ms_per_frame": round(ms, 4),
                     "mtri_per_s": round(sc.triangles / (ms * 1e-3) / 1e6, 2),
                     "gpix_per_s": round(frags / (ms * 1e-3) / 1e9, 3),
                     "raster_ms": round(stage[5] / 10, 4)}
        r.close()
    return out


def diff_bench(args, dev):
    """Batched forward + backward over all views (graph replay) against one view at a
    time, on the same renderer code."""
    import torch

    from softwarerenderer_amd import scenes
    from softwarerenderer_amd.diff import DiffRenderer

    V, S_, sub = args.diff_views, args.diff_size, args.diff_subdiv
    Vt, F = scenes.icosphere(sub)
    N = Vt.shape[0]
    rng = np.random.default_rng(0)
    mvps = torch.from_numpy(scenes.orbit_cameras(V, S_, S_, seed=0)).float().to(dev)
    pos = torch.from_numpy(Vt).float().to(dev)
    col = torch.from_numpy(rng.random((N, 3))).float().to(dev)
    hom = torch.cat([pos, torch.ones_like(pos[:, :1])], 1)
    tp = pos + torch.from_numpy(rng.normal(0, 0.02, Vt.shape)).float().to(dev)
    tc = col + torch.from_numpy(rng.normal(0, 0.1, (N, 3))).float().to(dev)
    thom = torch.cat([tp, torch.ones_like(tp[:, :1])], 1)

    r = DiffRenderer(F, N, V, S_, S_)
    target = r.forward(torch.einsum("vij,nj->vni", mvps, thom).contiguous(), tc, sync=True).clone()

    def step(rr, m, tgt):
        clip = torch.einsum("vij,nj->vni", m, hom).contiguous()
        rr.forward(clip, col)
        # L = 1/2 |image - target|^2: dL/dimage = image - target, formed in the backward
        gclip, gattr = rr.backward_l2(tgt)
        gpos = torch.einsum("vni,vij->nj", gclip, m)[:, :3]
        return gpos, gattr

    s = torch.cuda.Stream(dev)
    s.wait_stream(torch.cuda.current_stream(dev))
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        step(r, mvps, target)
        with torch.cuda.graph(graph, stream=s):
            step(r, mvps, target)
    torch.cuda.current_stream(dev).wait_stream(s)
    torch.cuda.synchronize()
    reps = 10
    graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    r.check()
    t_batch = e0.elapsed_time(e1) / reps

    # the same work one view at a time: a one-view renderer, eager launches per view
    r1 = DiffRenderer(F, N, 1, S_, S_)
    for v in range(V):  # synchronous draws size the arenas for every view
        r1.forward(torch.einsum("vij,nj->vni", mvps[v:v + 1], hom).contiguous(), col, sync=True)
    step(r1, mvps[:1], target[:1])
    torch.cuda.synchronize()
    e0.record()
    for v in range(V):
        step(r1, mvps[v:v + 1], target[v:v + 1])
    e1.record()
    torch.cuda.synchronize()
    r1.check()
    t_plain = e0.elapsed_time(e1)
    r.close()
    r1.close()
    return {"workload": f"icosphere-s{sub}-{V}views-{S_}px-fwd+bwd", "vertices": N,
            "triangles": int(F.shape[0]), "views": V, "size": S_,
            "views_per_s": round(V / (t_batch * 1e-3), 1),
            "plain_views_per_s": round(V / (t_plain * 1e-3), 1),
            "speedup": round(t_plain / t_batch, 2),
            "batched_ms_per_step": round(t_batch, 4), "plain_ms_per_step": round(t_plain, 4),
            "batched": "one visibility draw for all views + forward + backward, HIP graph replay",
            "plain": "per view: visibility draw + forward + backward, eager launches"}


def terrain_bench(args, dev):
    """Batched differentiation under load: 32 fBm heightfields of 8k..2.1M triangles, 4
    views each at 1024x1024, all in one SceneDiffRenderer.  Times (a) the forward + backward
    step of all views from a HIP graph against the same work one view at a time (eager
    launches), and (b) the graph-captured Adam step (lr 1e-2) on heights and colours, run
    for the workload's 100 steps."""
    import torch

    from softwarerenderer_amd import scenes
    from softwarerenderer_amd.diff import DiffRenderer, SceneDiffRenderer, render

    views, size = 4, args.terrain_size
    work = scenes.terrain_workload(args.terrain_scenes, seed=5, views=views, size=size)
    sr = SceneDiffRenderer([(w["tris"], w["pos"].shape[0], views) for w in work], size, size)
    mvps = [torch.from_numpy(w["mvps"]).float().to(dev) for w in work]
    xz = [torch.from_numpy(w["pos"][:, [0, 2]]).float().to(dev) for w in work]
    cuts = np.cumsum([0] + [w["pos"].shape[0] for w in work])
    rng = np.random.default_rng(1)
    h0 = torch.from_numpy(np.concatenate([w["pos"][:, 1] for w in work])).float().to(dev)
    c0 = torch.from_numpy(np.concatenate([w["col"] for w in work])).float().to(dev)

    def clip_scene(k, h):
        hk = h[cuts[k]:cuts[k + 1]]
        p = torch.stack([xz[k][:, 0], hk, xz[k][:, 1], torch.ones_like(hk)], 1)
        return torch.matmul(mvps[k], p.t()).transpose(1, 2)

    def clip_of(h):
        return SceneDiffRenderer.clip_rows([clip_scene(k, h) for k in range(len(work))])

    with torch.no_grad():
        th = h0 + torch.from_numpy(rng.normal(0, 0.02, h0.shape)).float().to(dev)
        tc = c0 + torch.from_numpy(rng.normal(0, 0.1, c0.shape)).float().to(dev)
        target = sr.forward(clip_of(th), tc, sync=True).clone()
        # headroom for the optimisation moving the terrain (captured draws cannot grow
        # their arenas): Adam's steps roughen the heights, and tile pairs grow with the
        # slivers; HBM is plentiful, so reserve several times the target's needs
        st = sr.r.stats()
        sr.reserve(primitives=int(st.input_primitives * 1.3),
                   pairs=int(st.pairs_emitted * args.terrain_reserve))
        sr.forward(clip_of(h0), c0, sync=True)
    pairs_first = int(sr.r.stats().pairs_emitted)
    sr.check()
    s = torch.cuda.Stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def replay_ms(graph, reps):
        graph.replay()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            graph.replay()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    # (a) forward + backward of every view, HIP graph
    clip0 = clip_of(h0).contiguous()

    def fwd_bwd():
        img = sr.forward(clip0, c0)
        return sr.backward(img - target)

    s.wait_stream(torch.cuda.current_stream(dev))
    g_fb = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        fwd_bwd()
        with torch.cuda.graph(g_fb, stream=s):
            fwd_bwd()
    torch.cuda.current_stream(dev).wait_stream(s)
    t_fb = replay_ms(g_fb, 5)
    sr.check()
    del g_fb

    # the same work one view at a time: a one-view renderer per scene, eager launches
    plain = [DiffRenderer(w["tris"], w["pos"].shape[0], 1, size, size) for w in work]
    with torch.no_grad():
        clips = [clip_scene(k, h0).contiguous() for k in range(len(work))]
    for k, r1 in enumerate(plain):  # synchronous draws size each renderer's arenas
        for v in range(views):
            r1.forward(clips[k][v:v + 1], c0[cuts[k]:cuts[k + 1]], sync=True)
    torch.cuda.synchronize()
    e0.record()
    for k, r1 in enumerate(plain):
        ck = c0[cuts[k]:cuts[k + 1]]
        for v in range(views):
            img = r1.forward(clips[k][v:v + 1], ck)
            r1.backward(img - target[k * views + v:k * views + v + 1])
    e1.record()
    torch.cuda.synchronize()
    t_plain = e0.elapsed_time(e1)
    for r1 in plain:
        r1.check()
        r1.close()

    # (b) the Adam step, HIP graph, 100 steps
    h = h0.clone().requires_grad_(True)
    c = c0.clone().requires_grad_(True)
    opt = torch.optim.Adam([h, c], lr=1e-2, capturable=True)

    def adam_step():
        opt.zero_grad(set_to_none=False)
        img = render(sr, clip_of(h), c)
        loss = 0.5 * ((img - target) ** 2).sum()
        loss.backward()
        opt.step()
        return loss.detach()

    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        first = float(adam_step())
    torch.cuda.current_stream(dev).wait_stream(s)
    torch.cuda.synchronize()
    g_adam = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g_adam, stream=s):
        out = adam_step()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(99):
        g_adam.replay()
    e1.record()
    torch.cuda.synchronize()
    t_adam = e0.elapsed_time(e1) / 99
    sr.check()
    last = float(out)
    with torch.no_grad():
        sr.forward(clip_of(h), c, sync=True)
    pairs_last = int(sr.r.stats().pairs_emitted)
    V = sr.V
    T = int(sum(w["tris"].shape[0] for w in work))
    sr.close()
    return {"workload": f"terrain-{len(work)}scenes-{views}views-{size}px-fwd+bwd",
            "scenes": len(work), "views": V, "triangles": T, "size": size,
            "vertices": int(cuts[-1]),
            "views_per_s": round(V / (t_fb * 1e-3), 1),
            "plain_views_per_s": round(V / (t_plain * 1e-3), 1),
            "speedup": round(t_plain / t_fb, 2),
            "batched_ms_per_step": round(t_fb, 3), "plain_ms_per_step": round(t_plain, 3),
            "adam_ms_per_step": round(t_adam, 3), "adam_steps": 100,
            "loss_first": round(first, 3), "loss_last": round(last, 3),
            "tile_pairs_first": pairs_first, "tile_pairs_last": pairs_last,
            "batched": "one visibility draw for all views + forward + backward, HIP graph replay",
            "plain": "per view: visibility draw + forward + backward, eager launches",
            "adam": "render + L2 loss + backward + Adam(lr 1e-2) on heights and colours, "
                    "HIP graph replay"}


def tuning_summary(tu: dict) -> dict:
    """The autotuner's pick, and the best time of each raster work distribution: one
    workgroup per bin (hardware dispatch) and persistent workgroups on the bin queues."""
    def best(queue: bool):
        ts = [t for c, t in tu["table"] if (c.get("raster_queue", 0) > 0) == queue]
        return round(min(ts), 4) if ts else None
    return {"chosen": tu["best"], "tuned_ms": round(tu["ms"], 4),
            "default_ms": round(tu["default_ms"], 4), "configs": len(tu["table"]),
            "best_dispatch_ms": best(False), "best_queue_ms": best(True)}


def batch_bench(args, dev):
    """The many-scene workload (256 UV spheres, 1k..200k triangles, 256x256..1920x1080,
    Span / Block, depth-checker program): one batched pass (swr_batch_draw, enqueued
    asynchronously) against one drawElements launch sequence per scene, same programs and
    targets, targets cleared in every step."""
    import torch

    from softwarerenderer_amd import api, scenes
    from softwarerenderer_amd.autotune import autotune

    work = scenes.batch_workload(args.batch_scenes, seed=3)
    px = np.concatenate([[0], np.cumsum([sc.width * sc.height for sc in work])])
    color = torch.zeros(int(px[-1]), dtype=torch.int32, device=dev)
    depth = torch.ones(int(px[-1]), dtype=torch.float32, device=dev)
    verts = [torch.from_numpy(sc.vertices).to(dev) for sc in work]
    idx = [torch.from_numpy(sc.indices).to(dev) for sc in work]
    tg = [(color[px[i]:px[i + 1]].view(sc.height, sc.width),
           depth[px[i]:px[i + 1]].view(sc.height, sc.width)) for i, sc in enumerate(work)]
    modes = (api.RasterMode.Span, api.RasterMode.Block)
    r = api.Rasterizer(0)
    r.setStream(torch.cuda.current_stream(dev).cuda_stream)
    vp = api.VertexProcessor(r)
    r.setPixelShader(api.DepthCheckerShader(*tg[0]))
    vp.setVertexShader(api.MvpVertexShader(work[0].mvp))
    descs = [{"indices": idx[i], "vertices": verts[i], "stride": sc.stride, "mvp": sc.mvp,
              "viewport": (0, 0, sc.width, sc.height), "scissor": (0, 0, sc.width, sc.height),
              "cull_mode": api.CullMode.CW, "raster_mode": modes[sc.info["raster_mode"]],
              "targets": tg[i], "size": (sc.height, sc.width)} for i, sc in enumerate(work)]
    b = api.Batch(r, descs)

    def clear():
        color.zero_()
        depth.fill_(1.0)

    def step_batch():
        clear()
        b.draw(async_=True)

    shaders = [api.DepthCheckerShader(*t) for t in tg]
    vshaders = [api.MvpVertexShader(sc.mvp) for sc in work]

    def step_plain(sync=False):
        clear()
        for i, sc in enumerate(work):
            r.setRasterMode(modes[sc.info["raster_mode"]])
            r.setScissorRect(0, 0, sc.width, sc.height)
            r.setPixelShader(shaders[i])
            vp.setViewport(0, 0, sc.width, sc.height)
            vp.setCullMode(api.CullMode.CW)
            vp.setVertexShader(vshaders[i])
            vp.setVertexAttribPointer(0, sc.stride, verts[i])
            if sync:
                vp.drawElements(api.DrawMode.Triangle, idx[i])
            else:
                vp.drawElementsAsync(api.DrawMode.Triangle, idx[i])

    b.draw()        # sizes the arenas for the batch
    step_plain(True)  # and for every single scene
    r.setPixelShader(api.DepthCheckerShader(*tg[0]))
    tuning = None
    if not args.no_tune:
        tu = autotune(r, step_batch, reps=3, warmup=1, rounds=1)
        tuning = tuning_summary(tu)
    reps = 5
    s = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        r.synchronize()
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        r.synchronize()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    t_batch = timed(step_batch)
    t_plain = timed(step_plain)
    r.setTuning()
    T = sum(sc.triangles for sc in work)
    b.close()
    r.close()
    return {"workload": f"spheres-{len(work)}scenes-1k..200k-tris-256px..1080p-span/block",
            "scenes": len(work), "triangles": T, "pixels": int(px[-1]),
            "scenes_per_s": round(len(work) / (t_batch * 1e-3), 1),
            "plain_scenes_per_s": round(len(work) / (t_plain * 1e-3), 1),
            "speedup": round(t_plain / t_batch, 2),
            "mtri_per_s": round(T / (t_batch * 1e-3) / 1e6, 1),
            "batched_ms_per_step": round(t_batch, 4), "plain_ms_per_step": round(t_plain, 4),
            "tuning": tuning,
            "batched": "one swr_batch_draw of all scenes (targets cleared each step)",
            "plain": "per scene: state + drawElementsAsync (targets cleared each step)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the draw path has no multi-GPU launch (asked for "
              f"--gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    import torch

    from softwarerenderer_amd import api, scenes

    dev = torch.device("cuda:0")
    # one stream for the clears, the draws and the timing events
    stream = torch.cuda.Stream(dev)
    torch.cuda.set_stream(stream)
    if args.copy_probe:
        n = 256 * 1024 * 1024 // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    sc = scenes.soup_scene(args.ntri, seed=args.seed)
    W, H, T = sc.width, sc.height, sc.triangles
    V = int(np.unique(sc.indices).size)
    verts = torch.from_numpy(sc.vertices).to(dev)
    idx = torch.from_numpy(sc.indices).to(dev)
    color = torch.zeros((H, W), dtype=torch.int32, device=dev)
    depth = torch.ones((H, W), dtype=torch.float32, device=dev)

    r = api.Rasterizer(0)
    r.setStream(stream.cuda_stream)
    v = api.VertexProcessor(r)
    r.setRasterMode(api.RasterMode.Block)
    r.setScissorRect(0, 0, W, H)
    r.setPixelShader(api.DepthCheckerShader(color, depth))
    v.setViewport(0, 0, W, H)
    v.setCullMode(api.CullMode.CW)
    v.setVertexShader(api.MvpVertexShader(sc.mvp))
    v.setVertexAttribPointer(0, sc.stride, verts)

    def step(sync: bool):
        color.zero_()
        depth.fill_(1.0)
        if sync:
            v.drawElements(api.DrawMode.Triangle, idx)
        else:
            v.drawElementsAsync(api.DrawMode.Triangle, idx)

    # the first synchronous draw sizes the arenas
    step(True)
    tuning = None
    if args.tuning:
        fixed = {k: int(v) for k, v in (kv.split("=") for kv in args.tuning.split(","))}
        r.setTuning(**fixed)
        tuning = {"chosen": r.tuning(), "fixed": True}
    elif not args.no_tune:
        from softwarerenderer_amd.autotune import autotune
        tu = autotune(r, lambda: step(False))
        tuning = tuning_summary(tu)
    shape = r.tuning()
    for _ in range(max(1, args.warmup)):
        step(True)
    st = r.stats()
    prims, pairs = int(st.raster_primitives), int(st.tile_pairs)
    torch.cuda.synchronize()

    # timed region: K draws enqueued back to back
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t0.record(stream)
    for _ in range(args.steps):
        step(False)
    t1.record(stream)
    r.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1) / args.steps
    # outside the timed region: the last timed frame against the reference's digest of it
    parity = frame_parity(sc, args, color, depth)

    # per-stage kernel time (HIP events on the draw's stream), synchronous draws
    r.setProfiling(True)
    nprof = min(args.steps, 20)
    stage = np.zeros(8)
    for _ in range(nprof):
        step(True)
        stage += np.array(r.stats().stage_ms[:8], dtype=np.float64)
    r.setProfiling(False)
    stage /= nprof

    # fragments shaded (depth-failing fragments count): one hash-program draw
    hsh = torch.zeros((H, W), dtype=torch.int32, device=dev)
    cnt = torch.zeros((H, W), dtype=torch.int32, device=dev)
    r.setPixelShader(api.HashShader(hsh, cnt))
    v.drawElements(api.DrawMode.Triangle, idx)
    frags = int(cnt.sum().item())
    r.setPixelShader(api.DepthCheckerShader(color, depth))

    raster_ms = float(stage[5])
    frame_bytes = 12 * T + 32 * V + 16 * W * H
    raster_bytes = 16 * W * H
    kern_ms = float(stage[:6].sum())
    raster_gbs = raster_bytes / (raster_ms * 1e-3) / 1e9 if raster_ms > 0 else 0.0
    out = {
        "metric": "Mtri/s + Gpix/s shaded (1080p, 1M tris)",
        "value": round(T / (ms * 1e-3) / 1e6, 3),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "soup-1M-1080p-block", "triangles": T, "vertices": V,
                   "width": W, "height": H, "raster_mode": "Block", "cull": "CW",
                   "pixel_program": "depth<, rgb*checker(uv,8) RGBA8 + f32 depth",
                   "seed": args.seed, "raster_primitives": prims, "tile_pairs": pairs,
                   "fragments": frags},
        "gpix_per_s": round(frags / (ms * 1e-3) / 1e9, 3),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "stage_ms": {"geometry": round(stage[0], 4), "scans": round(stage[1], 4),
                     "emit": round(stage[2], 4), "sort": round(stage[3], 4),
                     "ranges": round(stage[4], 4), "raster": round(stage[5], 4)},
        "roofline": {"bound": "hbm", "kernel": "k_raster<DepthChecker>",
                     "achieved": round(raster_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(raster_gbs / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic("soup-1M-1080p-block", shape),
                     "source": source_revision(),
                     "launch_shape": raster_shape(shape),
                     "algorithmic_bytes": raster_bytes,
                     "valu_issue": valu_issue("soup-1M-1080p-block", shape, raster_ms)},
        "parity": parity,
        "tuning": tuning,
        "frame": {"algorithmic_bytes": frame_bytes, "kernel_ms": round(kern_ms, 4),
                  "achieved_GBs": round(frame_bytes / (kern_ms * 1e-3) / 1e9, 2)
                  if kern_ms > 0 else None},
    }
    r.close()

    progress(f"frame: {out['ms_per_step']} ms per frame")
    if not args.no_span:
        out["span"] = span_bench(args, dev)
        progress("span legs done")
    if not args.no_diff:
        out["diff"] = diff_bench(args, dev)
        progress(f"diff: {out['diff']['batched_ms_per_step']} ms per step")
    if not args.no_batch:
        out["batch"] = batch_bench(args, dev)
        progress(f"batch: {out['batch']['batched_ms_per_step']} ms per step")
    if not args.no_terrain:
        out["terrain"] = terrain_bench(args, dev)
        progress(f"terrain: {out['terrain']['batched_ms_per_step']} ms per step")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_runs)
        progress("cpu baseline done")
    print(json.dumps(out))


def frame_parity(sc, args, color, depth) -> dict:
    """The drawn frame (RGBA8 words, then depth words) against the digest the reference
    wrote for the same scene (tests/golden/frame1m.npz, Block, depth-checker program)."""
    path = os.path.join(ROOT, "tests", "golden", "frame1m.npz")
    if not os.path.exists(path):
        return {"reference_digest": "absent"}
    gold = np.load(path, allow_pickle=False)
    inp = hashlib.sha256()
    for a in (sc.vertices, sc.indices, sc.mvp):
        inp.update(np.ascontiguousarray(a).view(np.uint8).tobytes())
    if inp.hexdigest() != str(gold["input"]):
        return {"reference_digest": "not this scene"}
    h = hashlib.sha256()
    for t in (color, depth):
        h.update(t.cpu().numpy().view(np.uint8).tobytes())
    ok = h.hexdigest() == str(gold["m1_p1"])
    return {"reference_digest": "match" if ok else "MISMATCH",
            "checked": "last timed frame, RGBA8 + depth words, sha256"}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, runs: int):
    """The reference CPU path timed on this host's cores on the same scene, as
    Benchmark.cpp:101-104 times drawElements: one warm-up, then the median of `runs` draws,
    targets cleared outside the timed region.  Two builds of the reference's own sources
    (oracle/_ref): serial, and with its OpenMP pragmas (Rasterizer.h:257, 369, 396) on
    OMP_NUM_THREADS threads (the usable cores when unset).  `value` is the faster one; the
    oracle (a C port) stands in where the reference build is absent."""
    import oracle as O

    ncores = len(os.sched_getaffinity(0))
    pipe = O.Pipeline(sc.width, sc.height, raster_mode=O.RASTER_BLOCK, mvp=sc.mvp,
                      ps_program=O.PS_DEPTH_CHECKER)

    def time_kind(kind):
        cpu = O.CpuRenderer(kind)
        times = []
        for i in range(runs + 1):
            t = pipe.targets()
            s = time.perf_counter()
            cpu.draw_elements(pipe, t, [(sc.vertices, sc.stride)], O.DRAW_TRIANGLE, sc.indices)
            if i:
                times.append(time.perf_counter() - s)
        progress(f"cpu {kind}: {statistics.median(times) * 1e3:.0f} ms per frame")
        return statistics.median(times)

    legs = {}
    if O.ref_available():
        legs["serial"] = (time_kind("ref"), 1, "reference")
        if os.path.exists(O.REF_OMP_SO):
            os.environ.setdefault("OMP_NUM_THREADS", str(ncores))
            legs["openmp"] = (time_kind("ref_omp"), int(os.environ["OMP_NUM_THREADS"]),
                              "reference")
    else:
        legs["serial"] = (time_kind("oracle"), 1, "port")
    best = min(legs, key=lambda k: legs[k][0])
    med, cores, kind = legs[best]
    return {"value": round(sc.triangles / med / 1e6, 4), "unit": "Mtri/s", "cores": cores,
            "kind": kind,
            "sample": f"full frame ({sc.triangles} triangles, 1920x1080 Block, depth-checker), "
                      f"median of {runs} draws after one warm-up, {best} build: "
                      f"{med * 1e3:.0f} ms per frame",
            "legs": {k: {"ms_per_frame": round(v[0] * 1e3, 1), "threads": v[1],
                         "mtri_per_s": round(sc.triangles / v[0] / 1e6, 4)}
                     for k, v in legs.items()},
            "host": {"nproc": ncores, "cpu_model": _cpu_model()}}


if __name__ == "__main__":
    main()
