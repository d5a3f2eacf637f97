"""Shared drivers for the parity tests: one pipeline description, run on the oracle, the
reference build, or the gfx950 path, with identical inputs."""
from __future__ import annotations

import numpy as np

import oracle as O

LOG_WORDS = O.LOG_WORDS


def run_cpu(renderer, pipe: O.Pipeline, scene=None, mode=O.DRAW_TRIANGLE, indices=None,
            vertices=None, prim="triangle", log_capacity=0):
    """Draw on a CpuRenderer; returns the Targets.  With `vertices` the IRasterizer list
    entry (window-space RasterizerVertex array) is used instead of drawElements."""
    t = pipe.targets(log_capacity=log_capacity)
    if vertices is not None:
        renderer.draw_list(pipe, t, prim, vertices, indices)
    else:
        idx = scene.indices if indices is None else indices
        renderer.draw_elements(pipe, t, [(scene.vertices, scene.stride)], mode, idx)
    return t


def run_gpu(pipe: O.Pipeline, scene=None, mode=O.DRAW_TRIANGLE, indices=None, vertices=None,
            prim="triangle", log_capacity=0, device_buffers=True, rasterizer=None,
            pixel_shader=None, vertex_shader=None, attribs=None, draw_async=False):
    """Draw through the C ABI (softwarerenderer_amd.api); returns numpy arrays shaped like
    oracle.Targets (t0, t1, log records, log_count).  With `pixel_shader` that shader draws
    instead of the pipeline's program (the caller reads its own buffers); with
    `vertex_shader` that shader replaces the pipeline's vertex program, reading the
    attribute buffers `attribs` ([(stride, array), ...]; default the scene's vertices)."""
    import torch
    from softwarerenderer_amd import api

    H, W = pipe.height, pipe.width
    ref = pipe.targets(log_capacity=log_capacity)  # initial values
    dev = "cuda" if device_buffers else "cpu"

    def tens(a):
        if a is None:
            return None
        return torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).to(dev)

    t0, t1 = tens(ref.t0), tens(ref.t1)
    log = log_count = None
    if pipe.ps_program == O.PS_LOG:
        log = torch.zeros((max(log_capacity, 1), LOG_WORDS), dtype=torch.int32, device=dev)
        log_count = torch.zeros(1, dtype=torch.int64, device=dev)

    ps = api.PixelShaderBase()
    ps.Program = pipe.ps_program
    ps.InterpolateZ, ps.InterpolateW = pipe.ps_z, pipe.ps_w
    ps.AVarCount, ps.PVarCount = pipe.ps_a, pipe.ps_p
    ps.checker_cells = pipe.checker_cells
    ps.targets = lambda: (t0, t1, None, W, H, W)
    if log is not None:
        ps.log, ps.log_count = log, log_count

    r = rasterizer or api.Rasterizer(0)
    tex = None
    if pipe.ps_program == O.PS_TEXTURE:
        tex = api.Texture(r, pipe.texture.texels, pipe.texture.max_anisotropy)
        ps.texture = tex
    r.setRasterMode(pipe.raster_mode)
    r.setScissorRect(*pipe.scissor)
    r.setPixelShader(ps if pixel_shader is None else pixel_shader)
    if vertices is not None:
        vt = torch.from_numpy(np.ascontiguousarray(vertices).view(np.uint8)).to(dev)
        it = torch.from_numpy(np.ascontiguousarray(indices, np.int32)).to(dev)
        fn = {"triangle": r.drawTriangleList, "line": r.drawLineList, "point": r.drawPointList}
        fn[prim](vt, it)
    else:
        vp = api.VertexProcessor(r)
        vp.setViewport(*pipe.viewport)
        vp.setDepthRange(*pipe.depth_range)
        vp.setCullMode(pipe.cull_mode)
        if vertex_shader is None:
            vs = api.VertexShaderBase()
            vs.Program = pipe.vs_program
            vs.AttribCount, vs.AVarCount, vs.PVarCount = pipe.vs_counts
            vs.mvp = pipe.mvp
            vp.setVertexShader(vs)
        else:
            vp.setVertexShader(vertex_shader)
        for i, (stride, arr) in enumerate(attribs or [(scene.stride, scene.vertices)]):
            vp.setVertexAttribPointer(i, stride, torch.from_numpy(np.ascontiguousarray(arr))
                                      .to(dev))
        idx = scene.indices if indices is None else indices
        it = torch.from_numpy(np.ascontiguousarray(idx, np.int32)).to(dev)
        if draw_async:
            vp.drawElementsAsync(mode, it)
            r.synchronize()
        else:
            vp.drawElements(mode, it)
    out = O.Targets(pipe.ps_program, W, H, log_capacity)
    if t0 is not None:
        out.t0 = t0.cpu().numpy().view(ref.t0.dtype).copy()
    if t1 is not None:
        out.t1 = t1.cpu().numpy().view(ref.t1.dtype).copy()
    if log is not None:
        out.log = log.cpu().numpy().view(np.uint32).copy()
        out.log_count = log_count.cpu().numpy().view(np.uint64).copy()
    if tex is not None:
        tex.close()
    if rasterizer is None:
        r.close()
    return out


def sorted_log(t: O.Targets) -> np.ndarray:
    """Log records ordered by (y, x, seq): each pixel's invocations in draw order."""
    rec = t.log_records()
    order = np.lexsort((rec[:, 3], rec[:, 0], rec[:, 1]))
    return rec[order]


def assert_same(a: O.Targets, b: O.Targets, what=""):
    if a.t0 is not None:
        bad = np.argwhere(a.t0.view(np.uint32) != b.t0.view(np.uint32))
        assert bad.size == 0, f"{what}: target0 differs at {len(bad)} pixels, first (y,x) {bad[:5].tolist()}"
    if a.t1 is not None:
        bad = np.argwhere(a.t1.view(np.uint32) != b.t1.view(np.uint32))
        assert bad.size == 0, f"{what}: target1 differs at {len(bad)} pixels, first (y,x) {bad[:5].tolist()}"
    if a.log is not None:
        assert int(a.log_count[0]) == int(b.log_count[0]), f"{what}: invocation counts differ"
        la, lb = sorted_log(a), sorted_log(b)
        assert la.shape == lb.shape
        bad = np.nonzero(np.any(la != lb, axis=1))[0]
        assert bad.size == 0, f"{what}: {bad.size} invocation records differ, first {la[bad[:1]].tolist()} vs {lb[bad[:1]].tolist()}"
