"""Benchmark of the gfx950 draw path on the 1080p, 1M-triangle soup (Block mode).

One step = one drawElements of the whole frame (vertex program, clip, cull, setup,
binning, raster, depth-tested checker shading into RGBA8 + fp32 depth) preceded by the
target clear, with vertices and indices resident in HBM.  Prints one JSON line.

The line also carries the differentiation workload (64 views at 512x512 of a 40,962-vertex
icosphere, forward + backward of the image loss 1/2 |image - target|^2): views per second of the batched step
(one visibility draw for all views, captured in a HIP graph) against the same work issued
one view at a time.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-diff] [--copy-probe]
"""
from __future__ import annotations

import argparse
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
    p.add_argument("--no-tune", action="store_true",
                   help="skip the launch-shape autotuner (built-in defaults)")
    return p.parse_args()


def measured_traffic(workload: str):
    """HBM bytes per raster launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs recorded
    in profiles/traffic.json (scaled by the copy-probe factors recorded with them)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload:
        return None
    return t.get("raster_bytes")


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
        img = rr.forward(clip, col)
        g = img - tgt  # dL/dimage of L = 1/2 |image - target|^2
        gclip, gattr = rr.backward(g)
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
    for v in range(2):
        step(r1, mvps[v:v + 1], target[v:v + 1])
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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); sharding is described in DESIGN.md", file=sys.stderr)
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
    if not args.no_tune:
        from softwarerenderer_amd.autotune import autotune
        tu = autotune(r, lambda: step(False))
        tuning = {"chosen": tu["best"], "tuned_ms": round(tu["ms"], 4),
                  "default_ms": round(tu["default_ms"], 4), "configs": len(tu["table"])}
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
                     "traffic": measured_traffic("soup-1M-1080p-block"),
                     "algorithmic_bytes": raster_bytes},
        "tuning": tuning,
        "frame": {"algorithmic_bytes": frame_bytes, "kernel_ms": round(kern_ms, 4),
                  "achieved_GBs": round(frame_bytes / (kern_ms * 1e-3) / 1e9, 2)
                  if kern_ms > 0 else None},
    }
    r.close()

    if not args.no_diff:
        out["diff"] = diff_bench(args, dev)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_runs)
    print(json.dumps(out))


def cpu_baseline(sc, runs: int):
    """The reference CPU path (oracle/_ref, or the oracle where that is absent) timed on
    this host's cores on the same scene: one warm-up, then the median of `runs` draws."""
    import oracle as O

    kind = "ref" if O.ref_available() else "oracle"
    cpu = O.CpuRenderer(kind)
    pipe = O.Pipeline(sc.width, sc.height, raster_mode=O.RASTER_BLOCK, mvp=sc.mvp,
                      ps_program=O.PS_DEPTH_CHECKER)
    times = []
    for i in range(runs + 1):
        t = pipe.targets()
        s = time.perf_counter()
        cpu.draw_elements(pipe, t, [(sc.vertices, sc.stride)], O.DRAW_TRIANGLE, sc.indices)
        if i:
            times.append(time.perf_counter() - s)
    med = statistics.median(times)
    return {"value": round(sc.triangles / med / 1e6, 4), "unit": "Mtri/s", "cores": 1,
            "kind": "reference" if kind == "ref" else "port",
            "sample": f"full frame ({sc.triangles} triangles, 1920x1080 Block), median of "
                      f"{runs} serial draws after one warm-up; {med * 1e3:.0f} ms per frame"}


if __name__ == "__main__":
    main()
