"""Autotuner for the draw path's launch shapes.

The knobs (``swr_tuning``, set with :meth:`Rasterizer.setTuning`) change how the work is cut
into workgroups, never what is drawn: every setting produces the same bits, which the GPU
tests check.  The tuner times a caller-supplied draw under each setting of a search space and
leaves the fastest one set on the rasterizer.

    best = autotune(r, lambda: v.drawElementsAsync(DrawMode.Triangle, idx))

The draw must be asynchronous and enqueue on the rasterizer's stream; the tuner binds that
stream to torch's current stream for the duration, so its HIP events bracket the draws.
"""
from __future__ import annotations

import itertools
from typing import Callable, Iterable, Optional

import torch

KNOBS = ("raster_waves", "raster_group", "xcd_order", "expand_chunk")

# raster workgroup size in 8x8 tiles, primitives per raster row-walk group (0: the pixel
# program's default), XCD-grouped tile order on/off, tile pairs per expansion workgroup
DEFAULT_SPACE = {
    "raster_waves": (1, 2, 4, 8),
    "raster_group": (0, 1, 2, 4),
    "xcd_order": (1, -1),
    "expand_chunk": (512, 1024, 2048),
}


def configs(space: Optional[dict] = None) -> Iterable[dict]:
    space = dict(DEFAULT_SPACE, **(space or {}))
    for vals in itertools.product(*(space[k] for k in KNOBS)):
        yield dict(zip(KNOBS, vals))


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


def autotune(rasterizer, draw: Callable[[], None], space: Optional[dict] = None,
             reps: int = 8, warmup: int = 2, rounds: int = 2) -> dict:
    """Times `draw` under every configuration of `space` (``DEFAULT_SPACE`` overridden per
    knob), `rounds` passes in alternating order to cancel drift, keeps each configuration's
    best time, sets the fastest on `rasterizer` and returns
    ``{"best": cfg, "ms": t, "default_ms": t0, "table": [(cfg, t), ...]}``."""
    stream = torch.cuda.current_stream()
    rasterizer.setStream(stream.cuda_stream)
    rasterizer.synchronize()
    start = rasterizer.tuning()
    cfgs = list(configs(space))
    best_t = {i: float("inf") for i in range(len(cfgs))}
    try:
        rasterizer.setTuning()
        default_ms = time_draws(draw, reps, warmup)
        for rnd in range(rounds):
            order = range(len(cfgs)) if rnd % 2 == 0 else reversed(range(len(cfgs)))
            for i in order:
                rasterizer.setTuning(**cfgs[i])
                best_t[i] = min(best_t[i], time_draws(draw, reps, warmup))
    except BaseException:
        rasterizer.synchronize()
        rasterizer.setTuning(**start)
        raise
    rasterizer.synchronize()
    i_best = min(best_t, key=best_t.get)
    rasterizer.setTuning(**cfgs[i_best])
    table = sorted(((cfgs[i], t) for i, t in best_t.items()), key=lambda ct: ct[1])
    return {"best": cfgs[i_best], "ms": best_t[i_best], "default_ms": default_ms, "table": table}
