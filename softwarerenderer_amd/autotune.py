"""Autotuner for the draw path's launch shapes.

The knobs (``swr_tuning``, set with :meth:`Rasterizer.setTuning`) change how the work is cut
into workgroups, never what is drawn: every setting produces the same bits, which the GPU
tests check.  The tuner times a caller-supplied draw under each setting of a search space and
leaves the fastest one set on the rasterizer.

    best = autotune(r, lambda: v.drawElementsAsync(DrawMode.Triangle, idx))

The draw must be asynchronous and enqueue on the rasterizer's stream; the tuner binds that
stream to torch's current stream for the duration, so its HIP events bracket the draws.

The default search runs in two stages.  First the binning shape: screen bins of 8 pixels with
1, 2, 4 or 8 bins (one wave each) per raster workgroup, or bins of 16 or 32 pixels with 1, 2
or 4 waves sharing a bin's 8x8 blocks, each with and without the vertex pass.  Then, for the
fastest shape, primitives per row-walk group, XCD-grouped bin order, the expansion chunk and
persistent raster workgroups fed by per-XCD bin queues.
"""
from __future__ import annotations

import itertools
from typing import Callable, Iterable, Optional

import torch

KNOBS = ("raster_waves", "raster_group", "xcd_order", "expand_chunk", "vertex_pass",
         "bin_size", "waves_per_bin", "raster_queue")

# stage 1: (bin size, raster workgroup shape) x vertex pass
SHAPES = ([dict(bin_size=8, raster_waves=w) for w in (1, 2, 4, 8)] +
          [dict(bin_size=b, waves_per_bin=w) for b in (16, 32) for w in (1, 2, 4)])
VERTEX_PASS = (1, -1)
# stage 2, for the fastest shape: primitives per raster row-walk group (0: the pixel
# program's default), XCD-grouped bin order on/off, tile pairs per expansion workgroup,
# persistent raster workgroups fed by per-XCD bin queues on/off
REFINE = {
    "raster_group": (0, 1, 2, 4),
    "xcd_order": (1, -1),
    "expand_chunk": (512, 1024, 2048),
    "raster_queue": (-1, 1),
}


def full(cfg: dict) -> dict:
    """A complete setting: every knob, 0 (the built-in default) where `cfg` has none."""
    return {k: cfg.get(k, 0) for k in KNOBS}


def configs(space: dict) -> Iterable[dict]:
    """Every combination of the knobs `space` lists (the others at their defaults)."""
    keys = [k for k in KNOBS if k in space]
    for vals in itertools.product(*(space[k] for k in keys)):
        yield full(dict(zip(keys, vals)))


def staged_configs(best_shape: Optional[dict] = None) -> Iterable[dict]:
    if best_shape is None:
        for shape in SHAPES:
            for vp in VERTEX_PASS:
                yield full(dict(shape, vertex_pass=vp))
    else:
        for cfg in configs(REFINE):
            yield full(dict(cfg, **{k: best_shape[k] for k in
                                    ("bin_size", "raster_waves", "waves_per_bin",
                                     "vertex_pass")}))


def time_draws(draw: Callable[[], None], reps: int, warmup: int) -> float:
    """Milliseconds per draw: HIP events on the current stream around `reps` queued draws."""
    for _ in range(warmup):
        draw()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        draw()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def _sweep(rasterizer, draw, cfgs, reps, warmup, rounds, best_t):
    for rnd in range(rounds):
        order = range(len(cfgs)) if rnd % 2 == 0 else reversed(range(len(cfgs)))
        for i in order:
            rasterizer.setTuning(**cfgs[i])
            key = tuple(sorted(cfgs[i].items()))
            best_t[key] = min(best_t.get(key, float("inf")), time_draws(draw, reps, warmup))


def autotune(rasterizer, draw: Callable[[], None], space: Optional[dict] = None,
             reps: int = 8, warmup: int = 2, rounds: int = 2) -> dict:
    """Times `draw` under every configuration of `space` (each listed knob's values, the
    others at their defaults), or of the two-stage default search when `space` is None,
    `rounds` passes in alternating order to cancel drift, keeps each configuration's best
    time, sets the fastest on `rasterizer` and returns
    ``{"best": cfg, "ms": t, "default_ms": t0, "table": [(cfg, t), ...]}``."""
    stream = torch.cuda.current_stream()
    rasterizer.setStream(stream.cuda_stream)
    rasterizer.synchronize()
    start = rasterizer.tuning()
    best_t: dict = {}
    try:
        rasterizer.setTuning()
        default_ms = time_draws(draw, reps, warmup)
        if space is not None:
            _sweep(rasterizer, draw, list(configs(space)), reps, warmup, rounds, best_t)
        else:
            _sweep(rasterizer, draw, list(staged_configs()), reps, warmup, rounds, best_t)
            shape = dict(min(best_t, key=best_t.get))
            _sweep(rasterizer, draw, list(staged_configs(shape)), reps, warmup, rounds, best_t)
    except BaseException:
        rasterizer.synchronize()
        rasterizer.setTuning(**start)
        raise
    rasterizer.synchronize()
    key = min(best_t, key=best_t.get)
    best = dict(key)
    rasterizer.setTuning(**best)
    table = sorted(((dict(k), t) for k, t in best_t.items()), key=lambda ct: ct[1])
    return {"best": best, "ms": best_t[key], "default_ms": default_ms, "table": table}
