"""Asynchronous and captured draws: failures are reported, never dropped.

* every asynchronous draw folds its overflow / error words into a sticky device word that
  the next draw does not clear; swr_synchronize (or the next synchronous draw) reports it
  once and clears it;
* an asynchronous draw takes device buffers only (nothing is staged or copied back);
* a draw that would have to grow a device arena while its stream is being captured fails
  before it enqueues anything.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SWR_ERR_INVALID_ARG = -1
SWR_ERR_CAPACITY = -5
W, H = 1920, 1080


def _quad(z=0.5):
    """Two triangles covering the whole viewport (identity MVP), stride 32: pos3 rgb3 uv2."""
    v = np.zeros((4, 8), np.float32)
    v[:, 0:3] = [[-1, -1, z], [1, -1, z], [1, 1, z], [-1, 1, z]]
    v[:, 3:6] = 0.75
    v[:, 6:8] = [[0, 0], [1, 0], [1, 1], [0, 1]]
    return v, np.array([0, 1, 2, 0, 2, 3], np.int32)


def _small(z=0.25):
    """One small triangle near the centre."""
    v = np.zeros((3, 8), np.float32)
    v[:, 0:3] = [[-0.01, -0.01, z], [0.01, -0.01, z], [0.0, 0.01, z]]
    v[:, 3:6] = 0.5
    return v, np.array([0, 1, 2], np.int32)


def _setup(r, api, color, depth, verts):
    vp = api.VertexProcessor(r)
    r.setRasterMode(api.RasterMode.Block)
    r.setScissorRect(0, 0, W, H)
    r.setPixelShader(api.DepthCheckerShader(color, depth))
    vp.setViewport(0, 0, W, H)
    vp.setCullMode(api.CullMode.None_)
    vp.setVertexShader(api.MvpVertexShader(np.eye(4, dtype=np.float32)))
    vp.setVertexAttribPointer(0, 32, verts)
    return vp


def _targets(torch, dev="cuda"):
    color = torch.zeros((H, W), dtype=torch.int32, device=dev)
    depth = torch.ones((H, W), dtype=torch.float32, device=dev)
    return color, depth


def test_overflowing_async_draw_is_reported_after_a_fitting_one(gpu):
    """A fresh context's arenas hold 2T + 4096 tile pairs; a full-screen quad needs ~65k, so
    its asynchronous draw overflows and draws nothing.  A fitting draw queued after it does
    not hide that: swr_synchronize reports SWR_ERR_CAPACITY once, then the context is clean."""
    import torch
    from softwarerenderer_amd import api
    r = api.Rasterizer(0)
    color, depth = _targets(torch)
    qv, qi = _quad()
    sv, si = _small()
    qv_d, qi_d = torch.from_numpy(qv).cuda(), torch.from_numpy(qi).cuda()
    sv_d, si_d = torch.from_numpy(sv).cuda(), torch.from_numpy(si).cuda()
    torch.cuda.synchronize()
    vp = _setup(r, api, color, depth, qv_d)
    vp.drawElementsAsync(api.DrawMode.Triangle, qi_d)       # overflows: skipped
    vp.setVertexAttribPointer(0, 32, sv_d)
    vp.drawElementsAsync(api.DrawMode.Triangle, si_d)       # fits
    with pytest.raises(api._abi.SwrError) as ei:
        r.synchronize()
    assert ei.value.code == SWR_ERR_CAPACITY
    r.synchronize()  # reported once
    torch.cuda.synchronize()
    # window depth: 0.625 for the small triangle, 0.75 for the quad (to plane rounding)
    d = depth.cpu().numpy()
    assert np.count_nonzero(np.abs(d - 0.625) < 1e-4) > 0   # the small triangle drew
    assert np.count_nonzero(np.abs(d - 0.75) < 1e-4) == 0   # the quad did not
    # a synchronous draw after a failed asynchronous one reports the earlier failure too,
    # and the overflowing quad itself grows the arenas and draws
    vp.setVertexAttribPointer(0, 32, qv_d)
    vp.drawElementsAsync(api.DrawMode.Triangle, qi_d)       # still too small: skipped
    with pytest.raises(api._abi.SwrError) as ei:
        vp.drawElements(api.DrawMode.Triangle, qi_d)
    assert ei.value.code == SWR_ERR_CAPACITY
    torch.cuda.synchronize()
    d = depth.cpu().numpy()
    assert np.count_nonzero(np.abs(d - 0.75) < 1e-4) > 0.9 * W * H
    vp.drawElementsAsync(api.DrawMode.Triangle, qi_d)       # now sized: fits
    r.synchronize()
    r.close()


def test_async_draw_rejects_host_buffers(gpu):
    """Host targets or host vertex data are not copied back / staged by an asynchronous
    draw, so it refuses them (SWR_ERR_INVALID_ARG) instead of returning SWR_OK."""
    import torch
    from softwarerenderer_amd import api
    r = api.Rasterizer(0)
    sv, si = _small()
    si_d = torch.from_numpy(si).cuda()
    color_h, depth_h = _targets(torch, "cpu")
    vp = _setup(r, api, color_h, depth_h, torch.from_numpy(sv).cuda())
    with pytest.raises(api._abi.SwrError) as ei:
        vp.drawElementsAsync(api.DrawMode.Triangle, si_d)
    assert ei.value.code == SWR_ERR_INVALID_ARG and "target0" in str(ei.value)
    color, depth = _targets(torch)
    vp = _setup(r, api, color, depth, sv)  # host vertex data
    with pytest.raises(api._abi.SwrError) as ei:
        vp.drawElementsAsync(api.DrawMode.Triangle, si_d)
    assert ei.value.code == SWR_ERR_INVALID_ARG and "attribute" in str(ei.value)
    # the same draw with device buffers goes through
    vp.setVertexAttribPointer(0, 32, torch.from_numpy(sv).cuda())
    vp.drawElementsAsync(api.DrawMode.Triangle, si_d)
    r.synchronize()
    assert torch.count_nonzero(depth != 1.0).item() > 0
    r.close()


def test_captured_draw_that_must_grow_fails_cleanly(gpu):
    """Under stream capture a draw whose arenas are too small returns SWR_ERR_CAPACITY before
    it enqueues anything (no synchronize, free, malloc or memset enters the graph); the
    capture itself stays valid, and after one uncaptured draw the same capture works."""
    import torch
    from softwarerenderer_amd import api
    r = api.Rasterizer(0)
    color, depth = _targets(torch)
    qv, qi = _quad()
    qv_d, qi_d = torch.from_numpy(qv).cuda(), torch.from_numpy(qi).cuda()
    vp = _setup(r, api, color, depth, qv_d)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    r.setStream(s.cuda_stream)
    x = torch.zeros(16, device="cuda")
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        x.add_(1.0)
        with pytest.raises(api._abi.SwrError) as ei:
            vp.drawElementsAsync(api.DrawMode.Triangle, qi_d)
    assert ei.value.code == SWR_ERR_CAPACITY
    assert "captured" in str(ei.value)
    graph.replay()
    torch.cuda.synchronize()
    assert float(x[0]) == 1.0  # the rest of the capture is intact
    assert torch.all(depth == 1.0)
    # sized by one uncaptured synchronous draw, the captured draw replays correctly
    with torch.cuda.stream(s):
        vp.drawElements(api.DrawMode.Triangle, qi_d)
        expect = color.clone(), depth.clone()
        depth.fill_(1.0)
        color.zero_()
    torch.cuda.synchronize()
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=s):
        vp.drawElementsAsync(api.DrawMode.Triangle, qi_d)
    for _ in range(3):
        g2.replay()
        torch.cuda.synchronize()
        r.synchronize()
        assert torch.equal(color, expect[0]) and torch.equal(depth, expect[1])
        depth.fill_(1.0)
        color.zero_()
        torch.cuda.synchronize()
    r.close()
