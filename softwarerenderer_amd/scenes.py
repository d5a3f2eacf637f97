"""Synthetic scenes for the draw path: the workloads the tests and the benchmark draw.

Every generator is deterministic in its seed and returns host numpy arrays; the same
buffers go to the GPU path and to the CPU checkers.

* :func:`sphere_scene` — UV sphere, 64 stacks x 128 slices (16,128 triangles), per-vertex RGB
  from a seeded uniform and uv = (slice/128, stack/64), camera at z=3, 60 degree fovy,
  near 0.1, far 10.  Vertex layout ``{px,py,pz, r,g,b, u,v}`` (32 bytes).
* :func:`soup_scene` — triangle soup: NDC centroids uniform in [-1.2,1.2]^2, view depth
  uniform in [0.05,12], screen area log-normal (median 20 px, sigma 1.5), random
  orientation and facing, 30% of index slots reusing an earlier vertex.
* :func:`benchmark_scene` — the scene of the reference's Benchmark.cpp:66-105 (40,960
  triangles with x,y,z,r,g,b from .NET ``System.Random(0)``, Random.cpp).
* :func:`clip_scene` — clip-space triangles that straddle every frustum plane.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

VERTEX_FLOATS = 8  # px,py,pz, r,g,b, u,v


@dataclass
class Scene:
    vertices: np.ndarray          # float32 [V, k]
    indices: np.ndarray           # int32 [3T]
    mvp: np.ndarray               # float32 [4,4], row-major: clip = M @ (p, 1)
    width: int
    height: int
    info: dict = field(default_factory=dict)

    @property
    def triangles(self) -> int:
        return self.indices.size // 3

    @property
    def stride(self) -> int:
        return self.vertices.shape[1] * 4


def perspective(fovy_deg: float, aspect: float, near: float, far: float) -> np.ndarray:
    """OpenGL-style projection (the matrix of vector_math.h:737-759)."""
    rad = math.radians(fovy_deg) / 2.0
    cot = math.cos(rad) / math.sin(rad)
    m = np.zeros((4, 4))
    m[0, 0] = cot / aspect
    m[1, 1] = cot
    m[2, 2] = -(far + near) / (far - near)
    m[2, 3] = -2.0 * near * far / (far - near)
    m[3, 2] = -1.0
    return m


def lookat(eye, center, up) -> np.ndarray:
    """View matrix (the matrix of vector_math.h:772-794)."""
    eye, center, up = (np.asarray(v, float) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    t = np.eye(4)
    t[:3, 3] = -eye
    return m @ t


def camera(width: int, height: int, eye=(0.0, 0.0, 3.0)):
    view = lookat(eye, (0, 0, 0), (0, 1, 0))
    proj = perspective(60.0, width / height, 0.1, 10.0)
    return proj @ view, view


def sphere_scene(stacks: int = 64, slices: int = 128, seed: int = 1, width: int = 640,
                 height: int = 480, radius: float = 1.0) -> Scene:
    rng = np.random.default_rng(seed)
    s = np.arange(stacks + 1)[:, None]
    k = np.arange(slices + 1)[None, :]
    theta = np.pi * s / stacks
    phi = 2.0 * np.pi * k / slices
    pos = np.stack([radius * np.sin(theta) * np.cos(phi),
                    radius * np.cos(theta) * np.ones_like(phi),
                    -radius * np.sin(theta) * np.sin(phi)], axis=-1).reshape(-1, 3)
    uv = np.stack([np.broadcast_to(k / slices, (stacks + 1, slices + 1)),
                   np.broadcast_to(s / stacks, (stacks + 1, slices + 1))], -1).reshape(-1, 2)
    rgb = rng.random((pos.shape[0], 3))
    verts = np.concatenate([pos, rgb, uv], axis=1).astype(np.float32)

    # per quad (stack a, slice b) in row-major order: (p00, p10, p01) unless a is the top
    # row, then (p10, p11, p01) unless a is the bottom row
    a = np.arange(stacks)[:, None]
    b = np.arange(slices)[None, :]
    p00, p10 = a * (slices + 1) + b, (a + 1) * (slices + 1) + b
    p01, p11 = p00 + 1, p10 + 1
    quad = np.stack([np.stack(np.broadcast_arrays(p00, p10, p01), -1),
                     np.stack(np.broadcast_arrays(p10, p11, p01), -1)], 2)
    keep = np.stack(np.broadcast_arrays(a != 0, a != stacks - 1), -1)
    keep = np.broadcast_to(keep, (stacks, slices, 2))
    idx = quad[keep].astype(np.int32).reshape(-1, 3)
    # outward-facing triangles are counter-clockwise seen from outside
    p = verts[:, :3].astype(np.float64)
    nrm = np.cross(p[idx[:, 1]] - p[idx[:, 0]], p[idx[:, 2]] - p[idx[:, 0]])
    outward = np.einsum("ij,ij->i", nrm, p[idx[:, 0]] + p[idx[:, 1]] + p[idx[:, 2]])
    assert np.all(outward > 0), "sphere winding"
    mvp, _ = camera(width, height)
    return Scene(verts, idx.reshape(-1), mvp.astype(np.float32), width, height,
                 {"name": "sphere", "stacks": stacks, "slices": slices, "seed": seed})


BATCH_SIZES = ((256, 256), (512, 512), (1024, 1024), (1920, 1080))


def batch_workload(n: int = 256, seed: int = 3, tri_range=(1_000, 200_000)) -> list:
    """The many-scene batching workload: ``n`` UV spheres built as :func:`sphere_scene`
    (slices = 2 x stacks, so T = 4 stacks (stacks - 1)) with triangle counts log-uniform in
    ``tri_range``, each at a resolution drawn uniformly from 256x256, 512x512, 1024x1024 and
    1920x1080 (camera aspect to match) and drawn in Span or Block mode, drawn uniformly
    (``info["raster_mode"]``: 0 Span, 1 Block).  About 9.6 M triangles and 221 M pixels in
    all for n = 256."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    out = []
    for i in range(n):
        t = math.exp(rng.uniform(lo, hi))
        stacks = max(2, int(round((1.0 + math.sqrt(1.0 + t)) / 2.0)))
        w, h = BATCH_SIZES[int(rng.integers(len(BATCH_SIZES)))]
        mode = int(rng.integers(2))
        sc = sphere_scene(stacks, 2 * stacks, seed=seed * 1000 + i, width=w, height=h)
        sc.info["raster_mode"] = mode
        out.append(sc)
    return out


def soup_scene(ntri: int = 1_000_000, seed: int = 2, width: int = 1920, height: int = 1080,
               reuse: float = 0.30, median_area: float = 20.0, sigma: float = 1.5) -> Scene:
    """Triangle soup: ~30% of triangles cross a side or the near plane, ~50% face away.

    Vertex reuse: a triangle is, with probability q, the parallelogram partner of the
    previous (fresh) triangle, sharing its edge (v1, v2); a fraction f = q / (1 + q) of
    triangles are partners and 2f/3 of the index slots reuse a vertex, so f = 1.5 * reuse.
    """
    rng = np.random.default_rng(seed)
    aspect = width / height
    mvp, view = camera(width, height)
    inv_view = np.linalg.inv(view)
    tan_half = math.tan(math.radians(30.0))

    f_partner = 1.5 * reuse
    q = f_partner / (1.0 - f_partner)
    partner = rng.random(ntri) < q
    partner[0] = False
    # a partner's predecessor must be fresh: in each run of candidates keep every other one
    ar = np.arange(ntri)
    starts = partner & ~np.r_[False, partner[:-1]]
    run_start = np.maximum.accumulate(np.where(starts, ar, 0))
    partner &= (ar - run_start) % 2 == 0
    fresh = ~partner

    nf = int(fresh.sum())
    c = rng.uniform(-1.2, 1.2, (nf, 2))
    depth = rng.uniform(0.05, 12.0, nf)
    area = np.exp(math.log(median_area) + sigma * rng.standard_normal(nf))
    shape = rng.standard_normal((nf, 3, 2))
    shape -= shape.mean(axis=1, keepdims=True)
    a0 = 0.5 * np.abs((shape[:, 1, 0] - shape[:, 0, 0]) * (shape[:, 2, 1] - shape[:, 0, 1])
                      - (shape[:, 2, 0] - shape[:, 0, 0]) * (shape[:, 1, 1] - shape[:, 0, 1]))
    shape *= np.sqrt(area / np.maximum(a0, 1e-12))[:, None, None]
    ndc = c[:, None, :] + shape * np.array([2.0 / width, 2.0 / height])
    dv = depth[:, None] * (1.0 + rng.uniform(-0.1, 0.1, (nf, 3)))
    view_pts = np.stack([ndc[..., 0] * dv * tan_half * aspect, ndc[..., 1] * dv * tan_half,
                         -dv, np.ones_like(dv)], axis=-1)
    world = view_pts @ inv_view.T  # [nf, 3, 4]

    # assemble vertices and indices in triangle order: a fresh triangle adds 3 vertices,
    # a partner (v2', v1', v1' + v2' - v0' of its fresh predecessor) adds 1
    nverts = 3 * nf + int(partner.sum())
    nnew = np.where(fresh, 3, 1)
    off = np.cumsum(nnew) - nnew
    rank = np.cumsum(fresh) - 1  # index into `world` of each fresh triangle
    pos = np.empty((nverts, 3))
    idx = np.empty((ntri, 3), np.int64)
    tf = np.nonzero(fresh)[0]
    wf = world[rank[tf], :, :3]
    for k in range(3):
        pos[off[tf] + k] = wf[:, k]
        idx[tf, k] = off[tf] + k
    tp = np.nonzero(partner)[0]
    wp = world[rank[tp - 1], :, :3]
    pos[off[tp]] = wp[:, 1] + wp[:, 2] - wp[:, 0]
    idx[tp, 0] = off[tp - 1] + 2
    idx[tp, 1] = off[tp - 1] + 1
    idx[tp, 2] = off[tp]
    rgb = rng.random((nverts, 3))
    uv = rng.random((nverts, 2))
    verts = np.concatenate([pos, rgb, uv], axis=1).astype(np.float32)
    return Scene(verts, idx.reshape(-1).astype(np.int32), mvp.astype(np.float32), width, height,
                 {"name": "soup", "ntri": ntri, "seed": seed, "reuse": reuse})


class DotNetRandom:
    """System.Random (Knuth subtractive generator) as published by .NET, the generator of
    the reference's Random.cpp:7-84; NextDouble = InternalSample() * (1 / int.MaxValue)."""

    MBIG = 2147483647
    MSEED = 161803398

    def __init__(self, seed: int = 0):
        sa = [0] * 56
        sub = self.MBIG if seed == -2147483648 else abs(seed)
        mj = self.MSEED - sub
        sa[55] = mj
        mk = 1
        for i in range(1, 55):
            ii = (21 * i) % 55
            sa[ii] = mk
            mk = mj - mk
            if mk < 0:
                mk += self.MBIG
            mj = sa[ii]
        for _ in range(1, 5):
            for i in range(1, 56):
                sa[i] -= sa[1 + (i + 30) % 55]
                if sa[i] < 0:
                    sa[i] += self.MBIG
        self.sa, self.inext, self.inextp = sa, 0, 21

    def _sample(self) -> int:
        a = self.inext + 1
        if a >= 56:
            a = 1
        b = self.inextp + 1
        if b >= 56:
            b = 1
        r = self.sa[a] - self.sa[b]
        if r == self.MBIG:
            r -= 1
        if r < 0:
            r += self.MBIG
        self.sa[a] = r
        self.inext, self.inextp = a, b
        return r

    def next_double(self) -> float:
        return self._sample() * (1.0 / self.MBIG)


def benchmark_scene(ntri: int = 4096 * 10, width: int = 640, height: int = 480) -> Scene:
    """The reference Benchmark (Benchmark.cpp:66-105): {x,y,z,r,g,b} per vertex, w = 1."""
    rnd = DotNetRandom(0)
    vals = np.array([rnd.next_double() for _ in range(ntri * 3 * 6)], np.float64)
    verts = vals.astype(np.float32).reshape(ntri * 3, 6)
    idx = np.arange(ntri * 3, dtype=np.int32)
    return Scene(verts, idx, np.eye(4, dtype=np.float32), width, height,
                 {"name": "benchmark", "ntri": ntri})


def clip_scene(ntri: int = 2000, seed: int = 3, width: int = 256, height: int = 192,
               navar: int = 3, npvar: int = 2, spread: float = 1.6,
               shrink=(0.05, 1.0)) -> Scene:
    """Clip-space triangles {x,y,z,w, avar[navar], pvar[npvar]} for the raw vertex program.
    Vertices are spread up to `spread` times the frustum so most triangles cross planes;
    each triangle is scaled about its first vertex by a factor drawn from `shrink`."""
    rng = np.random.default_rng(seed)
    n = 3 * ntri
    w = rng.uniform(0.2, 2.0, n)
    xyz = rng.uniform(-spread, spread, (n, 3)) * w[:, None]
    # keep triangles reasonably sized: pull each triangle's vertices toward its first vertex
    xyz = xyz.reshape(ntri, 3, 3)
    f = rng.uniform(shrink[0], shrink[1], (ntri, 1, 1))
    xyz = xyz[:, :1, :] + (xyz - xyz[:, :1, :]) * f
    xyz = xyz.reshape(n, 3)
    var = rng.random((n, navar + npvar))
    verts = np.concatenate([xyz, w[:, None], var], axis=1).astype(np.float32)
    idx = np.arange(n, dtype=np.int32)
    return Scene(verts, idx, np.eye(4, dtype=np.float32), width, height,
                 {"name": "clip", "ntri": ntri, "seed": seed})


def vertex_processor_test_scene() -> Scene:
    """The triangle of the reference's VertexProcessorTest (VertexProcessorTest.cpp:86-118):
    {x, y, z, r, g, b} vertices drawn with w = 1 into 640x480 through the viewport
    (100, 100, 440, 280), cull mode None."""
    v = np.array([[0.0, 0.5, 0.0, 1.0, 0.0, 0.0],
                  [-1.5, -0.5, 0.0, 0.0, 1.0, 0.0],
                  [1.5, -0.5, 0.0, 0.0, 0.0, 1.0]], np.float32)
    return Scene(v, np.arange(3, dtype=np.int32), np.eye(4, dtype=np.float32), 640, 480,
                 {"name": "vertex_processor_test"})


def box_camera(angle: float, width: int = 640, height: int = 480) -> np.ndarray:
    """The reference Box demo's camera (Box.cpp:182-187): eye (5 cos a, 2, 5 sin a) looking
    at the origin, 60 degree fovy, near 0.1, far 10; MVP as float32."""
    view = lookat((5.0 * math.cos(angle), 2.0, 5.0 * math.sin(angle)), (0, 0, 0), (0, 1, 0))
    proj = perspective(60.0, width / height, 0.1, 10.0)
    return (proj @ view).astype(np.float32)


def obj_scene(path: str, angle: float = 0.6, width: int = 640, height: int = 480) -> Scene:
    """An OBJ mesh (softwarerenderer_amd/obj.py) under the Box camera: vertices
    {x, y, z, nx, ny, nz, u, v} (32 bytes, the SWR_VS_MVP_POS_UV layout)."""
    from .obj import load_obj
    v, idx = load_obj(path)
    return Scene(v, idx, box_camera(angle, width, height), width, height,
                 {"name": "obj", "path": path, "angle": angle})


def floor_scene(width: int = 640, height: int = 480, extent: float = 20.0,
                repeat: float = 16.0) -> Scene:
    """A textured floor quad (y = -1, +-extent, uv 0..repeat) seen from just above it: grazing
    footprints (anisotropic sampling, minification down the mip chain) and a near-plane clip.
    Vertices in the {x, y, z, nx, ny, nz, u, v} layout of SWR_VS_MVP_POS_UV."""
    e, r = extent, repeat
    v = np.array([[-e, -1.0, -e, 0, 1, 0, 0.0, 0.0], [e, -1.0, -e, 0, 1, 0, r, 0.0],
                  [e, -1.0, e, 0, 1, 0, r, r], [-e, -1.0, e, 0, 1, 0, 0.0, r]], np.float32)
    idx = np.array([0, 2, 1, 0, 3, 2], np.int32)  # counter-clockwise seen from above
    view = lookat((0.3, -0.4, 6.0), (0.0, -1.0, -10.0), (0, 1, 0))
    mvp = (perspective(60.0, width / height, 0.1, 50.0) @ view).astype(np.float32)
    return Scene(v, idx, mvp, width, height, {"name": "floor"})


def procedural_texture(kind: str, width: int, height: int, seed: int = 0) -> np.ndarray:
    """uint32 [height, width] texels 0x00RRGGBB: 'noise' (uniform random bytes), 'checker'
    (8-texel cells with a colour gradient)."""
    rng = np.random.default_rng(seed)
    if kind == "noise":
        return rng.integers(0, 1 << 24, (height, width), dtype=np.uint32)
    yy, xx = np.mgrid[0:height, 0:width]
    cell = ((xx // 8 + yy // 8) % 2).astype(np.uint32)
    r = (xx * 255 // max(width - 1, 1)).astype(np.uint32)
    g = (yy * 255 // max(height - 1, 1)).astype(np.uint32)
    b = np.where(cell == 1, 230, 30).astype(np.uint32)
    return (r << 16) | (g << 8) | b


# RasterizerVertex, IRasterizer.h:42-53 (x, y, z, w, avar[16], pvar[16])
RASTERIZER_VERTEX = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("w", "<f4"),
                              ("avar", "<f4", (16,)), ("pvar", "<f4", (16,))])


def edge_indices(tri_indices: np.ndarray) -> np.ndarray:
    """Line-list indices of every triangle's three edges (v0 v1, v1 v2, v2 v0), in triangle
    order: the wireframe of a triangle list."""
    t = np.asarray(tri_indices, np.int32).reshape(-1, 3)
    return np.ascontiguousarray(np.stack([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]], 1).reshape(-1))


def window_lines(n: int, seed: int, width: int, height: int, margin: float = 12.0,
                 long_frac: float = 0.2, grid: bool = False) -> np.ndarray:
    """2n window-space RasterizerVertex endpoints (x, y, z, w, 16 avar, 16 pvar): short lines
    around random points plus a fraction crossing the whole target, endpoints up to `margin`
    pixels outside it; `grid` puts endpoints on half-pixel positions (truncation ties)."""
    rng = np.random.default_rng(seed)
    v = np.zeros(2 * n, RASTERIZER_VERTEX)
    c = rng.uniform([-margin, -margin], [width + margin, height + margin], (n, 2))
    d = rng.normal(0.0, 6.0, (n, 2))
    far = rng.random(n) < long_frac
    d[far] = rng.uniform([-width, -height], [width, height], (int(far.sum()), 2))
    p = np.stack([c, c + d], 1).reshape(-1, 2)
    if grid:
        p = np.round(p * 2) / 2
    v["x"], v["y"] = p[:, 0], p[:, 1]
    v["z"] = rng.random(2 * n)
    v["w"] = rng.uniform(0.5, 2.0, 2 * n)
    v["avar"] = rng.random((2 * n, 16))
    v["pvar"] = rng.random((2 * n, 16))
    return v


def icosphere(subdivisions: int = 6):
    """Unit icosphere: (vertices [N, 3] float64, triangles [T, 3] int32), outward faces
    counter-clockwise.  Subdivision s gives 10 * 4^s + 2 vertices and 20 * 4^s triangles
    (s = 6: 40,962 and 81,920)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.asarray(p, float) / np.linalg.norm(p) for p in v]
    faces = np.asarray(f, np.int64)
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = np.asarray(nf, np.int64)
    V = np.asarray(verts)
    nrm = np.cross(V[faces[:, 1]] - V[faces[:, 0]], V[faces[:, 2]] - V[faces[:, 0]])
    assert np.all(np.einsum("ij,ij->i", nrm, V[faces].sum(1)) > 0), "icosphere winding"
    return V, faces.astype(np.int32)


def orbit_cameras(views: int, width: int, height: int, radius: float = 3.0, seed: int = 0):
    """MVPs [views, 4, 4] (float64) of cameras uniform on a sphere of `radius` looking at
    the origin: 60-degree fovy, near 0.1, far 10."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(views, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    proj = perspective(60.0, width / height, 0.1, 10.0)
    out = np.zeros((views, 4, 4))
    for i in range(views):
        eye = radius * d[i]
        up = (0.0, 1.0, 0.0) if abs(d[i][1]) < 0.99 else (1.0, 0.0, 0.0)
        out[i] = proj @ lookat(eye, (0, 0, 0), up)
    return out


def fbm(nx: int, ny: int, seed: int, octaves: int = 6, persistence: float = 0.5,
        base: int = 4) -> np.ndarray:
    """Fractal value noise [ny, nx] over the unit square, zero-mean per octave: octave o
    interpolates (smoothstep weights) a seeded lattice of (base 2^o + 1)^2 uniform values
    and is weighted persistence^o."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, 1.0, nx)
    y = np.linspace(0.0, 1.0, ny)
    h = np.zeros((ny, nx))
    amp = 1.0
    for o in range(octaves):
        f = base << o
        lat = rng.random((f + 1, f + 1)) - 0.5
        gx, gy = x * f, y * f
        ix = np.minimum(gx.astype(np.int64), f - 1)
        iy = np.minimum(gy.astype(np.int64), f - 1)
        tx, ty = gx - ix, gy - iy
        tx, ty = tx * tx * (3 - 2 * tx), ty * ty * (3 - 2 * ty)
        top = lat[iy][:, ix] * (1 - tx) + lat[iy][:, ix + 1] * tx
        bot = lat[iy + 1][:, ix] * (1 - tx) + lat[iy + 1][:, ix + 1] * tx
        h += amp * (top * (1 - ty)[:, None] + bot * ty[:, None])
        amp *= persistence
    return h


def heightfield(cells: int, seed: int, octaves: int = 6, persistence: float = 0.5,
                relief: float = 1.2):
    """fBm terrain over [-1, 1]^2 in x-z with y up: (positions [N, 3] float64, triangles
    [T, 3] int32, colours [N, 3] float64).  cells x cells quads, (cells + 1)^2 vertices,
    2 cells^2 triangles, counter-clockwise seen from above; colour follows height."""
    n = cells + 1
    h = relief * fbm(n, n, seed, octaves, persistence)
    zs, xs = np.meshgrid(np.linspace(-1.0, 1.0, n), np.linspace(-1.0, 1.0, n), indexing="ij")
    pos = np.stack([xs, h, zs], -1).reshape(-1, 3)
    i = np.arange(cells)[:, None]
    j = np.arange(cells)[None, :]
    v00 = i * n + j
    v01, v10 = v00 + 1, v00 + n
    v11 = v10 + 1
    tris = np.stack([np.stack(np.broadcast_arrays(v00, v10, v11), -1),
                     np.stack(np.broadcast_arrays(v00, v11, v01), -1)], 2).reshape(-1, 3)
    t = (h.reshape(-1) - h.min()) / max(float(np.ptp(h)), 1e-12)
    rng = np.random.default_rng(seed + 7)
    col = np.stack([0.2 + 0.6 * t, 0.5 + 0.3 * (1 - t), 0.3 * np.ones_like(t)], -1)
    col = np.clip(col + rng.normal(0, 0.05, col.shape), 0.0, 1.0)
    return pos, tris.astype(np.int32), col


def terrain_cameras(views: int, width: int, height: int, seed: int = 0, radius: float = 2.6,
                    elevation=(20.0, 35.0)):
    """MVPs [views, 4, 4] (float64) of cameras around a terrain at the origin: azimuths
    evenly spaced from a seeded start, elevations uniform in `elevation` degrees, looking at
    the origin; 60-degree fovy, near 0.1, far 10."""
    rng = np.random.default_rng(seed)
    a0 = rng.uniform(0, 2 * np.pi)
    proj = perspective(60.0, width / height, 0.1, 10.0)
    out = np.zeros((views, 4, 4))
    for k in range(views):
        az = a0 + 2 * np.pi * k / views
        el = np.radians(rng.uniform(*elevation))
        eye = radius * np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
        out[k] = proj @ lookat(eye, (0, 0, 0), (0, 1, 0))
    return out


def terrain_workload(n: int = 32, seed: int = 5, cells_range=(64, 1024), views: int = 4,
                     size: int = 1024) -> list:
    """The batched-differentiation workload: n fBm heightfields (6 octaves, persistence 0.5)
    with cells per side log-uniform in `cells_range` (8k to 2.1M triangles), each seen by
    `views` cameras at size x size.  Per scene: {"cells", "pos", "tris", "col", "mvps"}."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(cells_range[0]), math.log(cells_range[1])
    out = []
    for k in range(n):
        cells = int(round(math.exp(rng.uniform(lo, hi))))
        pos, tris, col = heightfield(cells, seed=seed * 1000 + k)
        out.append({"cells": cells, "pos": pos, "tris": tris, "col": col,
                    "mvps": terrain_cameras(views, size, size, seed=seed * 1000 + k)})
    return out
