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
        self._build([(triangles, num_vertices, views)], width, height, channels, cull_mode,
                     raster_mode, device)
        self._clip_shape = (views, num_vertices, 4)

    def _build(self, scenes, width, height, channels, cull_mode, raster_mode, device):
        """scenes: [(triangles [T_s, 3] in the scene's own vertex numbering, N_s, V_s)]."""
        tl, ns, vs = [], [], []
        for tris, n, v in scenes:
            t = np.ascontiguousarray(np.asarray(tris, np.int32).reshape(-1, 3))
            if t.size and (t.min() < 0 or t.max() >= n):
                raise ValueError("triangle vertex numbers must lie in 0 .. vertices - 1")
            if n < 0 or v < 0:
                raise ValueError("negative vertex or view count")
            tl.append(t)
            ns.append(int(n))
            vs.append(int(v))
        S = len(tl)
        pv = np.concatenate([[0], np.cumsum(vs)]).astype(np.int32)
        pn = np.concatenate([[0], np.cumsum(ns)]).astype(np.int32)
        pt = np.concatenate([[0], np.cumsum([t.shape[0] for t in tl])]).astype(np.int32)
        self.V, self.W, self.H, self.C = int(pv[-1]), width, height, channels
        self.N, self.T = int(pn[-1]), int(pt[-1])
        self.S = S
        self.rows = int(sum(v * n for v, n in zip(vs, ns)))
        self._tables = [np.ascontiguousarray(a) for a in (pv, pn, pt)]
        gtris = np.concatenate([t + pn[k] for k, t in enumerate(tl)]) if S else \
            np.zeros((0, 3), np.int32)
        gadj = np.concatenate([np.where(adjacency(t) >= 0, adjacency(t) + pt[k], -1)
                               for k, t in enumerate(tl)]) if S else np.zeros((0, 3), np.int32)
        dev = torch.device("cuda", device)
        self.device = dev
        self.tris = torch.from_numpy(np.ascontiguousarray(gtris, np.int32)).to(dev)
        self.adj = torch.from_numpy(np.ascontiguousarray(gadj, np.int32)).to(dev)
        self._local = [torch.from_numpy(t.reshape(-1)).to(dev) for t in tl]
        V = self.V
        self.clip = torch.zeros((self.rows, 4), dtype=torch.float32, device=dev)
        self.ids = torch.zeros((V, height, width), dtype=torch.int32, device=dev)
        self.depth = torch.ones((V, height, width), dtype=torch.float32, device=dev)
        self.bary = torch.zeros((V, height, width, 2), dtype=torch.float32, device=dev)
        self.image = torch.zeros((V, height, width, channels), dtype=torch.float32, device=dev)
        self.flags = 1 if cull_mode != api.CullMode.None_ else 0  # SWR_DIFF_CULLED
        self.r = api.Rasterizer(device)
        self.r.setPixelShader(api.VisibilityShader())
        vp = api.VertexProcessor(self.r)
        vp.setVertexShader(api.RawVertexShader(0, 0))
        self._stream = None
        batch = []
        row = 0
        for k in range(S):
            for j in range(vs[k]):
                v = int(pv[k]) + j
                batch.append({"indices": self._local[k], "vertices": self.clip[row:row + ns[k]],
                              "stride": 16, "viewport": (0, 0, width, height),
                              "scissor": (0, 0, width, height), "cull_mode": cull_mode,
                              "raster_mode": raster_mode,
                              "targets": (self.ids[v], self.depth[v]), "size": (height, width)})
                row += ns[k]
        self.batch = api.Batch(self.r, batch)
        self._generation = 0
        self._drawn = False
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
        # no colour buffer: the kernels recompute the few colours before antialiasing they
        # need instead of the forward writing one per pixel
        d.bary, d.image = self.bary.data_ptr(), self.image.data_ptr()
        if self.S > 1:
            d.scenes = self.S
            d.scene_views, d.scene_vertices, d.scene_triangles = (
                t.ctypes.data for t in self._tables)
        return d

    def forward(self, clip: torch.Tensor, attr: torch.Tensor,
                sync: Optional[bool] = None) -> torch.Tensor:
        """Fills the renderer's image buffer; returns it (no copy).  The visibility draw is
        synchronous when `sync` is true, and by default on the first forward: a synchronous
        draw grows the arenas to what the scene needs, an asynchronous one (which a HIP graph
        can capture) reports SWR_ERR_CAPACITY at the next check() if they are too small."""
        if tuple(clip.shape) != self._clip_shape or attr.shape != (self.N, self.C):
            raise ValueError(f"clip must be {list(self._clip_shape)} and attr "
                             f"[{self.N}, {self.C}]")
        self._use_current_stream()
        self.clip.copy_(clip.detach().reshape(self.rows, 4))
        self.ids.zero_()
        self.depth.fill_(1.0)
        if sync is None:
            sync = not self._drawn
        self.batch.draw(async_=not sync)
        self._drawn = True
        attr = attr.detach().contiguous().float()
        self._attr_keep = attr
        d = self._desc(attr)
        self.r._check(self.r._lib.swr_diff_forward(self.r._ctx, C.byref(d)), "swr_diff_forward")
        self._generation += 1
        return self.image

    def backward(self, grad_image: torch.Tensor):
        """(grad_clip shaped like the forward's clip, grad_attr [N, C]) of the last
        forward."""
        return self._backward(grad_image, l2=False)

    def backward_l2(self, target: torch.Tensor):
        """The gradients of the loss 1/2 |image - target|^2 for the last forward: the same
        as backward(image - target), with the difference formed inside the backward
        kernels (SWR_DIFF_L2_TARGET) instead of as a separate image-sized tensor."""
        return self._backward(target, l2=True)

    def _backward(self, grad_image: torch.Tensor, l2: bool):
        self._use_current_stream()
        g = grad_image.detach().contiguous().float()
        gclip = torch.zeros(self._clip_shape, dtype=torch.float32, device=self.device)
        gattr = torch.zeros((self.N, self.C), dtype=torch.float32, device=self.device)
        d = self._desc(self._attr_keep)
        if l2:
            d.flags |= 2  # SWR_DIFF_L2_TARGET
        self.r._check(self.r._lib.swr_diff_backward(self.r._ctx, C.byref(d), g.data_ptr(),
                                                    gclip.data_ptr(), gattr.data_ptr()),
                      "swr_diff_backward")
        return gclip, gattr

    def reserve(self, factor: float = 1.5, primitives: Optional[int] = None,
                pairs: Optional[int] = None):
        """Grows the arenas to `factor` times what the last synchronous visibility draw
        needed, or to the given primitive and tile-pair counts (the next synchronous draw
        allocates them), so that asynchronous or captured draws of changing geometry, such
        as an optimisation moving the vertices, fit."""
        st = self.r.stats()
        if primitives is None:
            primitives = int(st.raster_primitives * factor) + 1024
        if pairs is None:
            pairs = int(st.pairs_emitted * factor) + 4096
        self.r.reserve(int(primitives), int(pairs))

    def check(self):
        """Waits for the queued work and raises if an asynchronous draw failed."""
        self.r.synchronize()

    def close(self):
        self.batch.close()
        self.r.close()


class SceneDiffRenderer(DiffRenderer):
    """Several meshes (scenes), each seen from views of its own, rendered and differentiated
    in one pass: ``scenes`` is a list of (triangles [T_s, 3] in the scene's own vertex
    numbering, vertex count N_s, view count V_s).  Views are numbered scene by scene; all
    are width x height.  Clip positions are one tensor of rows [sum V_s * N_s, 4]: scene by
    scene, each of its views' rows for its own vertices (:meth:`clip_rows` builds it from
    per-scene [V_s, N_s, 4] tensors); attributes are [sum N_s, C], scene by scene."""

    def __init__(self, scenes, width: int, height: int, channels: int = 3,
                 cull_mode: int = api.CullMode.CW, raster_mode: int = api.RasterMode.Span,
                 device: int = 0):
        self._build(list(scenes), width, height, channels, cull_mode, raster_mode, device)
        self._clip_shape = (self.rows, 4)

    @staticmethod
    def clip_rows(per_scene) -> torch.Tensor:
        """[sum V_s * N_s, 4] from per-scene clip tensors [V_s, N_s, 4] (differentiable)."""
        return torch.cat([c.reshape(-1, 4) for c in per_scene], 0)

    def scene_slices(self):
        """Per scene: (view slice, vertex slice, clip-row slice)."""
        pv, pn, _ = self._tables
        out, row = [], 0
        for k in range(self.S):
            n = int((pv[k + 1] - pv[k]) * (pn[k + 1] - pn[k]))
            out.append((slice(int(pv[k]), int(pv[k + 1])), slice(int(pn[k]), int(pn[k + 1])),
                        slice(row, row + n)))
            row += n
        return out


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
    clip [V, N, 4] (a SceneDiffRenderer: clip rows [sum V_s * N_s, 4]) with per-vertex
    attributes attr [N, C]."""
    return _Render.apply(clip, attr, renderer)
