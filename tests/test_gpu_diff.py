"""The differentiable path on the GPU: the batched visibility draw against the oracle, the
forward image and both gradients against the FP64 reference (1e-4 relative), the torch
autograd wrapper against central finite differences, and a HIP-graph replay against eager
execution."""
import numpy as np
import pytest

import oracle as O
from softwarerenderer_amd import scenes as S

pytestmark = pytest.mark.gpu

from test_diff_reference import make_case, visibility  # noqa: E402


def gpu_renderer(F, nverts, views, size, channels=3):
    from softwarerenderer_amd.diff import DiffRenderer
    return DiffRenderer(F, nverts, views, size, size, channels=channels)


def backward_unstored(r, g):
    """swr_diff_backward over copies of the visibility buffers: the decisions the forward
    stored do not name them, so the backward walks the silhouettes itself (one kernel)."""
    import ctypes as C
    import torch
    ids2, depth2 = r.ids.clone(), r.depth.clone()
    d = r._desc(r._attr_keep)
    d.ids, d.depth = ids2.data_ptr(), depth2.data_ptr()
    gclip = torch.zeros(r._clip_shape, dtype=torch.float32, device=r.device)
    gattr = torch.zeros((r.N, r.C), dtype=torch.float32, device=r.device)
    r.r._check(r.r._lib.swr_diff_backward(r.r._ctx, C.byref(d), g.data_ptr(), gclip.data_ptr(),
                                          gattr.data_ptr()), "swr_diff_backward")
    r.check()
    return gclip.cpu().numpy(), gattr.cpu().numpy()


def rel_err(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def test_visibility_matches_oracle(gpu, cpu_oracle):
    import torch
    size = 64
    V, F, clip, attr = make_case(views=3, size=size, subdiv=3, seed=11)
    ids_ref, depth_ref = visibility(cpu_oracle, F, clip, size)
    r = gpu_renderer(F, V.shape[0], 3, size)
    r.forward(torch.from_numpy(clip).cuda(), torch.from_numpy(attr).cuda(), sync=True)
    assert np.array_equal(r.ids.cpu().numpy().view(np.uint32), ids_ref)
    assert np.array_equal(r.depth.cpu().numpy().view(np.uint32), depth_ref.view(np.uint32))
    r.close()


def _fp64_check(r, clip, F, attr, seed):
    """Forward image and both gradients of `r` against the FP64 reference on the visibility
    buffer the GPU drew: image 1e-5 absolute; gradients 1e-4 relative as a norm and, per
    element, 1e-4 of the element or of 1% of the largest one on 99.9% of the elements."""
    import torch
    img = r.forward(torch.from_numpy(clip).cuda(), torch.from_numpy(attr).cuda(), sync=True)
    img = img.cpu().numpy()
    ids = r.ids.cpu().numpy().view(np.uint32)
    depth = r.depth.cpu().numpy()
    rng = np.random.default_rng(seed)
    g = rng.normal(size=img.shape).astype(np.float32)
    gt = torch.from_numpy(g).cuda()
    gclip, gattr = r.backward(gt)
    gclip, gattr = gclip.cpu().numpy(), gattr.cpu().numpy()
    adj = r.adj.cpu().numpy()
    ref_img, ref_gclip, ref_gattr = O.diff_reference(clip, F, adj, attr, ids, depth, g)
    assert np.abs(img - ref_img).max() <= 1e-5
    # the backward reusing the forward's decisions (two kernels) and the one walking them
    # again (one kernel) both match the reference
    for gc_, ga_ in ((gclip, gattr), backward_unstored(r, gt)):
        assert rel_err(ga_, ref_gattr) <= 1e-4
        assert rel_err(gc_, ref_gclip) <= 1e-4
        for got, ref in ((gc_, ref_gclip), (ga_, ref_gattr)):
            tol = 1e-4 * np.maximum(np.abs(ref), 1e-2 * np.abs(ref).max())
            assert (np.abs(got - ref) <= tol).mean() >= 0.999
    return ids


def test_fp64_at_the_differentiation_workload_size(gpu):
    """The differentiation workload's mesh and image size: the 40,962-vertex icosphere
    (subdivision 6) at 512x512, four of its orbit views, against the FP64 reference."""
    V, F, clip, attr = make_case(views=4, size=512, subdiv=6, seed=6)
    r = gpu_renderer(F, V.shape[0], 4, 512)
    ids = _fp64_check(r, clip, F, attr, seed=8)
    assert (ids > 0).sum() > 250_000
    r.close()


def test_fp64_on_a_large_heightfield(gpu):
    """One heightfield of the batched-differentiation workload's upper range: 256x256 cells
    (131,072 triangles) at 1024x1024, against the FP64 reference."""
    pos, F, clip, col = _terrain_case(256, 1, 1024, seed=9)
    r = gpu_renderer(F, pos.shape[0], 1, 1024)
    ids = _fp64_check(r, clip, F, col, seed=10)
    assert (ids > 0).sum() > 300_000
    r.close()


@pytest.mark.parametrize("subdiv,size,views", [(2, 48, 2), (4, 128, 3)])
def test_forward_and_gradients_match_fp64(gpu, subdiv, size, views):
    import torch
    V, F, clip, attr = make_case(views=views, size=size, subdiv=subdiv, seed=subdiv)
    r = gpu_renderer(F, V.shape[0], views, size)
    img = r.forward(torch.from_numpy(clip).cuda(), torch.from_numpy(attr).cuda(), sync=True)
    img = img.cpu().numpy()
    ids = r.ids.cpu().numpy().view(np.uint32)
    depth = r.depth.cpu().numpy()
    rng = np.random.default_rng(7)
    g = rng.normal(size=img.shape).astype(np.float32)
    gt = torch.from_numpy(g).cuda()
    gclip, gattr = r.backward(gt)
    gclip, gattr = gclip.cpu().numpy(), gattr.cpu().numpy()
    adj = r.adj.cpu().numpy()
    ref_img, ref_gclip, ref_gattr = O.diff_reference(clip, F, adj, attr, ids, depth, g)
    assert np.abs(img - ref_img).max() <= 1e-5
    # the backward reusing the forward's decisions (two kernels) and the one walking them
    # again (one kernel) both match the reference
    for gc_, ga_ in ((gclip, gattr), backward_unstored(r, gt)):
        assert rel_err(ga_, ref_gattr) <= 1e-4
        assert rel_err(gc_, ref_gclip) <= 1e-4
        # per element, beside the norm: relative 1e-4 of the element or of the largest one
        tol = 1e-4 * np.maximum(np.abs(ref_gclip), 1e-2 * np.abs(ref_gclip).max())
        assert (np.abs(gc_ - ref_gclip) <= tol).mean() >= 0.999
    r.close()


def _loss_setup(subdiv, size, views, seed, const_color=False):
    import torch
    from softwarerenderer_amd.diff import render
    Vt, F = S.icosphere(subdiv)
    mvps = torch.from_numpy(S.orbit_cameras(views, size, size, seed=seed)).float().cuda()
    rng = np.random.default_rng(seed + 1)
    pos = torch.from_numpy(Vt).float().cuda()
    col = (torch.full((Vt.shape[0], 3), 0.7) if const_color
           else torch.from_numpy(rng.random((Vt.shape[0], 3))).float()).cuda()
    r = gpu_renderer(F, Vt.shape[0], views, size)

    def clip_of(p):
        hom = torch.cat([p, torch.ones_like(p[:, :1])], 1)
        return torch.einsum("vij,nj->vni", mvps, hom).contiguous()

    with torch.no_grad():
        tp = pos + torch.from_numpy(rng.normal(0, 0.02, Vt.shape)).float().cuda()
        tc = col + torch.from_numpy(rng.normal(0, 0.1, (Vt.shape[0], 3))).float().cuda()
        target = render(r, clip_of(tp), tc).clone()

    def loss(p, c):
        img = render(r, clip_of(p), c)
        return ((img.double() - target.double()) ** 2).sum()

    p = pos.clone().requires_grad_(True)
    c = col.clone().requires_grad_(True)
    loss(p, c).backward()
    return r, Vt, F, mvps, pos, col, loss, p.grad.clone(), c.grad.clone(), rng


def test_autograd_colour_gradient_matches_finite_differences(gpu):
    """The image is linear in the colours: FD along random directions is exact up to
    rounding."""
    import torch
    r, Vt, F, mvps, pos, col, loss, gp, gc, rng = _loss_setup(4, 128, 4, 3)
    h = 1e-2
    for _ in range(3):
        dc = torch.from_numpy(rng.normal(size=(Vt.shape[0], 3))).float().cuda()
        with torch.no_grad():
            fd = (loss(pos, col + h * dc) - loss(pos, col - h * dc)).item() / (2 * h)
        an = float((gc * dc).sum())
        assert abs(fd - an) <= 1e-3 * abs(an) + 1e-6, (fd, an)
    r.close()


def test_autograd_interior_position_gradient_matches_finite_differences(gpu):
    """Vertices well inside the visible disc (all adjacent faces turned towards the camera)
    move the image only through the barycentrics and interior visibility changes, both
    continuous: central differences with h = 1e-3 agree per coordinate."""
    import torch
    r, Vt, F, mvps, pos, col, loss, gp, gc, rng = _loss_setup(3, 96, 1, 5)
    # direction to the camera: from the far-plane centre to the near-plane centre
    inv = np.linalg.inv(mvps[0].double().cpu().numpy())
    c0 = inv @ np.array([0.0, 0.0, -1.0, 1.0])
    c1 = inv @ np.array([0.0, 0.0, 1.0, 1.0])
    to_cam = c0[:3] / c0[3] - c1[:3] / c1[3]
    to_cam /= np.linalg.norm(to_cam)
    cand = np.flatnonzero(Vt @ to_cam > 0.8)  # unit sphere: the position is the normal
    ok = checked = 0
    for n in cand[:: max(1, len(cand) // 12)]:
        for k in range(3):
            an = float(gp[n, k])
            d = torch.zeros_like(pos)
            d[n, k] = 1.0
            h = 1e-3
            with torch.no_grad():
                fd = (loss(pos + h * d, col) - loss(pos - h * d, col)).item() / (2 * h)
            checked += 1
            if abs(fd - an) <= 0.03 * abs(an) + 0.02 * float(gp.abs().max()):
                ok += 1
    assert checked >= 24 and ok >= 0.9 * checked, (ok, checked)
    r.close()


def test_colour_and_interior_gradients_fd_at_workload_size(gpu):
    """Per coordinate, central differences (h = 1e-3) at the differentiation workload's size
    (subdivision 6, 512x512): colours of visible vertices over four views (the image is
    linear in them), and positions of vertices well inside the visible disc of one view."""
    import torch
    r, Vt, F, mvps, pos, col, loss, gp, gc, rng = _loss_setup(6, 512, 4, 11)
    h = 1e-3
    vis = np.flatnonzero(gc.abs().sum(1).cpu().numpy() > 0)
    assert len(vis) > 1000
    checked = ok = 0
    for n in rng.choice(vis, 16, replace=False):
        for k in range(3):
            d = torch.zeros_like(col)
            d[n, k] = 1.0
            with torch.no_grad():
                fd = (loss(pos, col + h * d) - loss(pos, col - h * d)).item() / (2 * h)
            an = float(gc[n, k])
            checked += 1
            ok += abs(fd - an) <= 1e-3 * abs(an) + 1e-4 * float(gc.abs().max())
    assert ok == checked, (ok, checked)
    r.close()
    r, Vt, F, mvps, pos, col, loss, gp, gc, rng = _loss_setup(6, 512, 1, 12)
    inv = np.linalg.inv(mvps[0].double().cpu().numpy())
    c0 = inv @ np.array([0.0, 0.0, -1.0, 1.0])
    c1 = inv @ np.array([0.0, 0.0, 1.0, 1.0])
    to_cam = c0[:3] / c0[3] - c1[:3] / c1[3]
    to_cam /= np.linalg.norm(to_cam)
    cand = np.flatnonzero(Vt @ to_cam > 0.8)
    ok = checked = 0
    for n in rng.choice(cand, 12, replace=False):
        for k in range(3):
            an = float(gp[n, k])
            d = torch.zeros_like(pos)
            d[n, k] = 1.0
            with torch.no_grad():
                fd = (loss(pos + h * d, col) - loss(pos - h * d, col)).item() / (2 * h)
            checked += 1
            ok += abs(fd - an) <= 0.03 * abs(an) + 0.02 * float(gp.abs().max())
    assert checked == 36 and ok >= 0.9 * checked, (ok, checked)
    r.close()


def test_autograd_silhouette_gradient_matches_finite_differences(gpu):
    """Silhouette vertices of a coarse mesh with one colour: the image changes only through
    the antialiased silhouette, and its gradient agrees with central differences."""
    import torch
    r, Vt, F, mvps, pos, col, loss, gp, gc, rng = _loss_setup(1, 64, 1, 7, const_color=True)
    nz = np.flatnonzero(gp.abs().sum(1).cpu().numpy() > 1e-3 * float(gp.abs().max()))
    assert len(nz) >= 6
    ok = checked = 0
    for n in nz[:12]:
        for k in range(3):
            an = float(gp[n, k])
            d = torch.zeros_like(pos)
            d[n, k] = 1.0
            h = 1e-4  # a wider step lets a pixel pair change its silhouette edge
            with torch.no_grad():
                fd = (loss(pos + h * d, col) - loss(pos - h * d, col)).item() / (2 * h)
            checked += 1
            if abs(fd - an) <= 0.05 * abs(an) + 0.01 * float(gp.abs().max()):
                ok += 1
    assert ok >= 0.9 * checked, (ok, checked)
    r.close()


def test_autograd_full_loss_direction(gpu):
    """The whole loss (colours and positions, random directions) at the differentiation
    workload's proportions: finite differences and the gradient point the same way.  Edge
    antialiasing leaves small discontinuities where the silhouette changes topology (limb
    triangles turning away), so this check is loose by design; the FP64 reference above is
    the tight one."""
    import torch
    r, Vt, F, mvps, pos, col, loss, gp, gc, rng = _loss_setup(3, 96, 2, 3)
    fds, ans = [], []
    h = 1e-3
    for _ in range(6):
        dp = torch.from_numpy(rng.normal(size=Vt.shape)).float().cuda()
        dc = torch.from_numpy(rng.normal(size=(Vt.shape[0], 3))).float().cuda()
        with torch.no_grad():
            fd = (loss(pos + h * dp, col + h * dc) - loss(pos - h * dp, col - h * dc)).item() / (2 * h)
        fds.append(fd)
        ans.append(float((gp * dp).sum() + (gc * dc).sum()))
    fds, ans = np.array(fds), np.array(ans)
    cos = float(fds @ ans / (np.linalg.norm(fds) * np.linalg.norm(ans)))
    assert cos >= 0.7, (cos, fds, ans)
    r.close()


def test_fused_l2_backward_matches_explicit_gradient(gpu):
    """backward_l2(target) (SWR_DIFF_L2_TARGET: dL/dimage = image - target formed where the
    kernels read it) equals backward(image - target), up to the order of float atomics."""
    import torch
    size, views = 96, 3
    V, F, clip, attr = make_case(views=views, size=size, subdiv=3, seed=31)
    r = gpu_renderer(F, V.shape[0], views, size)
    clip_t, attr_t = torch.from_numpy(clip).cuda(), torch.from_numpy(attr).cuda()
    target = torch.from_numpy(np.random.default_rng(2).random((views, size, size, 3))).float().cuda()
    img = r.forward(clip_t, attr_t, sync=True)
    gc1, ga1 = r.backward(img - target)
    r.forward(clip_t, attr_t)
    gc2, ga2 = r.backward_l2(target)
    r.check()
    assert torch.allclose(gc2, gc1, rtol=1e-5, atol=1e-6 * gc1.abs().max().item())
    assert torch.allclose(ga2, ga1, rtol=1e-5, atol=1e-6 * ga1.abs().max().item())
    assert gc1.abs().max().item() > 0
    r.close()


def test_graph_replay_matches_eager(gpu):
    """A forward + backward step captured in a HIP graph replays to the eager results."""
    import torch
    size, views = 64, 4
    V, F, clip, attr = make_case(views=views, size=size, subdiv=3, seed=21)
    r = gpu_renderer(F, V.shape[0], views, size)
    clip_t = torch.from_numpy(clip).cuda()
    attr_t = torch.from_numpy(attr).cuda()
    g = torch.from_numpy(np.random.default_rng(1).normal(size=(views, size, size, 3))).float().cuda()
    img = r.forward(clip_t, attr_t, sync=True).clone()
    gclip, gattr = r.backward(g)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        r.forward(clip_t, attr_t)  # warm up on the capture stream
        out = {}
        with torch.cuda.graph(graph, stream=s):
            out["img"] = r.forward(clip_t, attr_t)
            out["gclip"], out["gattr"] = r.backward(g)
    torch.cuda.current_stream().wait_stream(s)
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        r.check()
        assert torch.equal(out["img"], img)
        # float atomics may add in another order: equal to rounding
        assert torch.allclose(out["gclip"], gclip, rtol=1e-4, atol=1e-6 * gclip.abs().max().item())
        assert torch.allclose(out["gattr"], gattr, rtol=1e-4, atol=1e-6 * gattr.abs().max().item())
        out["img"].zero_()  # eager work between replays
    r.close()


def _terrain_case(cells, views, size, seed):
    """A heightfield seen by `views` terrain cameras: (positions, triangles, clip, colours)."""
    pos, F, col = S.heightfield(cells, seed=seed)
    mvps = S.terrain_cameras(views, size, size, seed=seed)
    hom = np.concatenate([pos, np.ones((pos.shape[0], 1))], axis=1)
    clip = np.einsum("vij,nj->vni", mvps, hom).astype(np.float32)
    return pos, F, clip, col.astype(np.float32)


def test_scene_renderer_matches_per_scene(gpu):
    """Several meshes in one pass (SceneDiffRenderer: one visibility batch, one forward and
    one backward call for all) against each mesh on a renderer of its own (same images and
    visibility, bit for bit) and against the FP64 reference per mesh (gradients 1e-4)."""
    import torch
    from softwarerenderer_amd.diff import DiffRenderer, SceneDiffRenderer
    size = 64
    cases = [make_case(views=2, size=size, subdiv=2, seed=1)[1:],
             _terrain_case(12, 3, size, seed=3)[1:],
             make_case(views=1, size=size, subdiv=3, seed=2)[1:],
             _terrain_case(20, 2, size, seed=4)[1:]]
    sr = SceneDiffRenderer([(F, c.shape[1], c.shape[0]) for F, c, _ in cases], size, size)
    clip = SceneDiffRenderer.clip_rows([torch.from_numpy(c).cuda() for _, c, _ in cases])
    attr = torch.from_numpy(np.concatenate([a for _, _, a in cases])).cuda()
    img = sr.forward(clip, attr, sync=True).clone()
    rng = np.random.default_rng(5)
    g = torch.from_numpy(rng.normal(size=tuple(img.shape)).astype(np.float32)).cuda()
    gclip, gattr = sr.backward(g)
    unstored = backward_unstored(sr, g)
    sr.check()
    for k, ((F, c, a), (vs, ns, rs)) in enumerate(zip(cases, sr.scene_slices())):
        r = DiffRenderer(F, c.shape[1], c.shape[0], size, size)
        one = r.forward(torch.from_numpy(c).cuda(), torch.from_numpy(a).cuda(), sync=True)
        assert torch.equal(one, img[vs]), f"scene {k}: image"
        assert torch.equal(r.ids, sr.ids[vs]), f"scene {k}: visibility"
        ids = r.ids.cpu().numpy().view(np.uint32)
        depth = r.depth.cpu().numpy()
        adj = r.adj.cpu().numpy()
        gk = g[vs].cpu().numpy()
        ref_img, ref_gclip, ref_gattr = O.diff_reference(c, F, adj, a, ids, depth, gk)
        assert np.abs(img[vs].cpu().numpy() - ref_img).max() <= 1e-5
        for gc_, ga_ in ((gclip.cpu().numpy(), gattr.cpu().numpy()), unstored):
            gc_k = gc_[rs].reshape(c.shape)
            assert rel_err(ga_[ns], ref_gattr) <= 1e-4, f"scene {k}: attributes"
            assert rel_err(gc_k, ref_gclip) <= 1e-4, f"scene {k}: clip"
        r.close()
    sr.close()


def test_terrain_adam_graph(gpu):
    """A small instance of the batched-differentiation workload: four heightfields of uneven
    size, four views each, heights and colours fitted with Adam (lr 1e-2) to the images of
    perturbed terrains by a HIP-graph-captured step (render, L2 loss, backward, Adam).  The
    graph's steps track eager steps from the same start, and the loss falls."""
    import torch
    from softwarerenderer_amd.diff import SceneDiffRenderer, render
    size, views = 96, 4
    work = S.terrain_workload(4, seed=9, cells_range=(10, 40), views=views, size=size)
    sr = SceneDiffRenderer([(w["tris"], w["pos"].shape[0], views) for w in work], size, size)
    mvps = [torch.from_numpy(w["mvps"]).float().cuda() for w in work]
    xz = [torch.from_numpy(w["pos"][:, [0, 2]]).float().cuda() for w in work]
    rng = np.random.default_rng(3)
    h0 = torch.from_numpy(np.concatenate([w["pos"][:, 1] for w in work])).float().cuda()
    c0 = torch.from_numpy(np.concatenate([w["col"] for w in work])).float().cuda()
    cuts = np.cumsum([0] + [w["pos"].shape[0] for w in work])

    def clip_of(h):
        out = []
        for k in range(len(work)):
            hk = h[cuts[k]:cuts[k + 1]]
            p = torch.stack([xz[k][:, 0], hk, xz[k][:, 1], torch.ones_like(hk)], 1)
            out.append(torch.matmul(mvps[k], p.t()).transpose(1, 2))
        return SceneDiffRenderer.clip_rows(out)

    with torch.no_grad():
        th = h0 + torch.from_numpy(rng.normal(0, 0.02, h0.shape)).float().cuda()
        tc = c0 + torch.from_numpy(rng.normal(0, 0.1, c0.shape)).float().cuda()
        target = render(sr, clip_of(th), tc).clone()
        sr.reserve(6.0)
        sr.forward(clip_of(h0), c0, sync=True)

    def make():
        h = h0.clone().requires_grad_(True)
        c = c0.clone().requires_grad_(True)
        return h, c, torch.optim.Adam([h, c], lr=1e-2, capturable=True)

    def step(h, c, opt):
        opt.zero_grad(set_to_none=False)
        img = render(sr, clip_of(h), c)
        loss = 0.5 * ((img - target) ** 2).sum()
        loss.backward()
        opt.step()
        return loss.detach()

    # eager reference
    h, c, opt = make()
    eager = [float(step(h, c, opt)) for _ in range(5)]
    sr.check()
    h_e, c_e = h.detach().clone(), c.detach().clone()
    # captured: one warm-up step on the side stream, then reset and capture
    h, c, opt = make()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step(h, c, opt)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.no_grad():
        h.copy_(h0)
        c.copy_(c0)
    for st in opt.state.values():
        for v in st.values():
            v.zero_()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        out = step(h, c, opt)
    torch.cuda.synchronize()
    losses = []
    for _ in range(5):
        graph.replay()
        torch.cuda.synchronize()
        losses.append(float(out))
    sr.check()
    # float atomics add in another order, and Adam turns a gradient that is zero up to
    # rounding into a full step of either sign: losses agree to rounding, parameters
    # everywhere but at such points
    assert np.allclose(losses, eager, rtol=1e-3), (losses, eager)
    for x, y in ((h, h_e), (c, c_e)):
        assert float(((x - y).abs() <= 1e-5).float().mean()) >= 0.99
    for _ in range(95):
        graph.replay()
    torch.cuda.synchronize()
    sr.check()
    final = float(out)
    assert final < 0.5 * eager[0], (eager[0], final)
    sr.close()


def test_scene_tables_rejected(gpu):
    """Scene tables that do not run from 0 to the totals, that decrease, or that are missing
    are refused before any kernel runs (SWR_ERR_INVALID_ARG); a good call afterwards still
    gives the same image."""
    import ctypes as C
    import torch
    from softwarerenderer_amd.diff import SceneDiffRenderer
    size = 32
    cases = [make_case(views=1, size=size, subdiv=1, seed=1)[1:],
             make_case(views=2, size=size, subdiv=1, seed=2)[1:]]
    sr = SceneDiffRenderer([(F, c.shape[1], c.shape[0]) for F, c, _ in cases], size, size)
    clip = SceneDiffRenderer.clip_rows([torch.from_numpy(c).cuda() for _, c, _ in cases])
    attr = torch.from_numpy(np.concatenate([a for _, _, a in cases])).cuda()
    img = sr.forward(clip, attr, sync=True).clone()
    good = [t.copy() for t in sr._tables]
    assert [t.tolist() for t in good] == [[0, 1, 3], [0, 42, 84], [0, 80, 160]]
    bad = [(0, [1, 1, 3]), (0, [0, 1, 2]), (1, [0, 90, 84]), (2, [0, 80, 159]),
           (0, [0, 4, 3])]
    for k, t in bad:
        tabs = [x.copy() for x in good]
        tabs[k] = np.asarray(t, np.int32)
        d = sr._desc(sr._attr_keep)
        d.scene_views, d.scene_vertices, d.scene_triangles = (x.ctypes.data for x in tabs)
        assert sr.r._lib.swr_diff_forward(sr.r._ctx, C.byref(d)) == -1, (k, t)
    d = sr._desc(sr._attr_keep)
    d.scene_triangles = None
    assert sr.r._lib.swr_diff_forward(sr.r._ctx, C.byref(d)) == -1
    again = sr.forward(clip, attr, sync=True)
    sr.check()
    assert torch.equal(again, img)
    sr.close()
