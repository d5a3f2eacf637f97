"""Differentiable rendering on the gfx950 draw path: a thin ``torch.autograd.Function`` over
the C ABI.

Forward, for V views of one triangle mesh:
  1. one batched visibility draw of every view (``swr_batch_draw`` with the visibility
     program: nearest source triangle + 1 and its depth per pixel);
  2. ``swr_diff_forward``: perspective-correct barycentrics, the interpolated attributes and
     the analytic edge-antialiasing blend across silhouette edges.
Backward: ``swr_diff_backward`` recomputes everything it needs from the visibility buffer,
the barycentrics and the colour, and returns d(loss)/d(clip positions) per view and
d(loss)/d(attributes).  Gradients w.r.t. world-space positions follow through torch's
own autograd of the projection that produced the clip positions.

The renderer owns static buffers, so a step can be captured in a HIP graph (``capture``).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import _abi, api


def adjacency(tris: np.ndarray) -> np.ndarray:
    """[T, 3] int32: the triangle across the edge opposite vertex e of each triangle (the
    edge (v[e+1], v[e+2])), -1 where no single other triangle shares that edge."""
    tris = np.asarray(tris, np.int64).reshape(-1, 3)
    T = tris.shape[0]
    a = tris[:, [1, 2, 0]].reshape(-1)
    b = tris[:, [2, 0, 1]].reshape(-1)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * (int(tris.max(initial=0)) + 1) + hi
    order = np.argsort(key, kind="stable")
    ks = key[order]
    adj = np.full(3 * T, -1, np.int64)
    start = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
    count = np.diff(np.r_[start, ks.size])
    pairs = start[count == 2]
    e1, e2 = order[pairs], order[pairs + 1]
    adj[e1] = e2 // 3
    adj[e2] = e1 // 3
    return adj.reshape(T, 3).astype(np.int32)


class DiffRenderer:
    """Renders V views (each width x height) of a fixed triangle list with C attribute
    channels per vertex; see :func:`render`."""

    def __init__(self, triangles, num_vertices: int, views: int, width: int, height: int,
                 channels: int = 3, cull_mode: int = api.CullMode.CW,
                 raster_mode: int = api.RasterMode.Span, device: int = 0):
        tris = np.ascontiguousarray(np.asarray(triangles, np.int32).reshape(-1, 3))
        self.V, self.W, self.H, self.N, self.C = views, width, height, num_vertices, channels
        self.T = tris.shape[0]
        dev = torch.device("cuda", device)
        self.device = dev
        self.tris = torch.from_numpy(tris).to(dev)
        self.adj = torch.from_numpy(adjacency(tris)).to(dev)
        self.clip = torch.zeros((views, num_vertices, 4), dtype=torch.float32, device=dev)
        self.ids = torch.zeros((views, height, width), dtype=torch.int32, device=dev)
        self.depth = torch.ones((views, height, width), dtype=torch.float32, device=dev)
        self.bary = torch.zeros((views, height, width, 2), dtype=torch.float32, device=dev)
        self.color = torch.zeros((views, height, width, channels), dtype=torch.float32, device=dev)
        self.image = torch.zeros_like(self.color)
        self.flags = 1 if cull_mode != api.CullMode.None_ else 0  # SWR_DIFF_CULLED
        self.r = api.Rasterizer(device)
        self.r.setPixelShader(api.VisibilityShader())
        vp = api.VertexProcessor(self.r)
        vp.setVertexShader(api.RawVertexShader(0, 0))
        self._stream = None
        scenes = []
        for v in range(views):
            scenes.append({"indices": self.tris.view(-1), "vertices": self.clip[v], "stride": 16,
                           "viewport": (0, 0, width, height), "scissor": (0, 0, width, height),
                           "cull_mode": cull_mode, "raster_mode": raster_mode,
                           "targets": (self.ids[v], self.depth[v]), "size": (height, width)})
        self.batch = api.Batch(self.r, scenes)
        self._generation = 0
        self._attr_keep: Optional[torch.Tensor] = None
        # one synchronous draw sizes the context's arenas, so later draws never allocate
        self.batch.draw()

    def _use_current_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        if s != self._stream:
            self.r.setStream(s)
            self._stream = s

    def _desc(self, attr: torch.Tensor) -> _abi.DiffDesc:
        d = _abi.DiffDesc()
        d.views, d.width, d.height = self.V, self.W, self.H
        d.vertex_count, d.triangle_count, d.channels = self.N, self.T, self.C
        d.flags = self.flags
        d.clip, d.triangles, d.adjacency = (self.clip.data_ptr(), self.tris.data_ptr(),
                                            self.adj.data_ptr())
        d.attributes = attr.data_ptr()
        d.ids, d.depth = self.ids.data_ptr(), self.depth.data_ptr()
        d.bary, d.color, d.image = (self.bary.data_ptr(), self.color.data_ptr(),
                                    self.image.data_ptr())
        return d

    def forward(self, clip: torch.Tensor, attr: torch.Tensor, sync: bool = False) -> torch.Tensor:
        """Fills the renderer's image buffer; returns it (no copy)."""
        if clip.shape != (self.V, self.N, 4) or attr.shape != (self.N, self.C):
            raise ValueError("clip must be [views, vertices, 4] and attr [vertices, channels]")
        self._use_current_stream()
        self.clip.copy_(clip.detach())
        self.ids.zero_()
        self.depth.fill_(1.0)
        self.batch.draw(async_=not sync)
        attr = attr.detach().contiguous().float()
        self._attr_keep = attr
        d = self._desc(attr)
        self.r._check(self.r._lib.swr_diff_forward(self.r._ctx, C.byref(d)), "swr_diff_forward")
        self._generation += 1
        return self.image

    def backward(self, grad_image: torch.Tensor):
        """(grad_clip [V, N, 4], grad_attr [N, C]) of the last forward."""
        self._use_current_stream()
        g = grad_image.detach().contiguous().float()
        gclip = torch.zeros((self.V, self.N, 4), dtype=torch.float32, device=self.device)
        gattr = torch.zeros((self.N, self.C), dtype=torch.float32, device=self.device)
        d = self._desc(self._attr_keep)
        self.r._check(self.r._lib.swr_diff_backward(self.r._ctx, C.byref(d), g.data_ptr(),
                                                    gclip.data_ptr(), gattr.data_ptr()),
                      "swr_diff_backward")
        return gclip, gattr

    def check(self):
        """Waits for the queued work and raises if an asynchronous draw failed."""
        self.r.synchronize()

    def close(self):
        self.batch.close()
        self.r.close()


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, clip, attr, renderer: DiffRenderer):
        image = renderer.forward(clip, attr).clone()
        ctx.renderer = renderer
        ctx.generation = renderer._generation
        return image

    @staticmethod
    def backward(ctx, grad_image):
        r: DiffRenderer = ctx.renderer
        if ctx.generation != r._generation:
            raise RuntimeError("DiffRenderer: another forward ran before this backward")
        gclip, gattr = r.backward(grad_image)
        return gclip, gattr, None


def render(renderer: DiffRenderer, clip: torch.Tensor, attr: torch.Tensor) -> torch.Tensor:
    """Differentiable image [V, H, W, C] of the mesh under per-view clip positions
    clip [V, N, 4] with per-vertex attributes attr [N, C]."""
    return _Render.apply(clip, attr, renderer)
