"""The geometry stage skips the clipper for triangles it can prove fully clipped
(swr_geometry.hip, fully_clipped_early).  Restated here in float32 numpy and checked
against the oracle's Sutherland-Hodgman clip (VertexProcessor.cpp:217-263 /
PolyClipper.cpp:45-81) on triangles hugging the planes: whenever the early-out fires, the
full clip must leave fewer than three vertices."""
import numpy as np

import oracle as O

PLANES = np.array([[-1, 0, 0, 1], [1, 0, 0, 1], [0, -1, 0, 1], [0, 1, 0, 1], [0, 0, -1, 1],
                   [0, 0, 1, 1]], np.float32)


def clip_mask(v):
    x, y, z, w = v
    m = 0
    m |= 1 if (w - x) < 0 else 0
    m |= 2 if (x + w) < 0 else 0
    m |= 4 if (w - y) < 0 else 0
    m |= 8 if (y + w) < 0 else 0
    m |= 16 if (w - z) < 0 else 0
    m |= 32 if (z + w) < 0 else 0
    return m


def plane_dot(p, v):
    a, b, c, d = PLANES[p]
    x, y, z, w = v
    return np.float32(np.float32(np.float32(a * x) + np.float32(b * y)) + np.float32(c * z)) + np.float32(d * w)


def fully_clipped_early(P):
    if not np.all(np.isfinite(P)):
        return False
    masks = [clip_mask(v) for v in P]
    cm = masks[0] | masks[1] | masks[2]
    if cm == 0:
        return False
    all_out = masks[0] & masks[1] & masks[2]
    if all_out & (cm & -cm):
        return True
    if not all_out:
        return False
    m = np.float32(max(np.float32(np.abs(v).sum(dtype=np.float32)) for v in P))
    margin = np.float32(-1e-5) * m
    for p in range(6):
        if all_out & (1 << p) and all(plane_dot(p, v) < margin for v in P):
            return True
    return False


def test_early_out_never_drops_a_visible_triangle(cpu_oracle):
    rng = np.random.default_rng(17)
    fired = 0
    for i in range(6000):
        w = rng.uniform(0.01, 4.0, 3).astype(np.float32)
        P = np.zeros((3, 4), np.float32)
        P[:, 3] = w
        P[:, :3] = (rng.uniform(-1.6, 1.6, (3, 3)) * w[:, None]).astype(np.float32)
        # push all vertices just outside one plane, others straddling
        p = rng.integers(0, 6)
        axis, sign = p // 2, (-1.0 if p % 2 == 0 else 1.0)
        eps = rng.choice([0.0, 1e-7, 1e-6, 1e-4, 1e-2, 0.3])
        P[:, axis] = (-sign * w * (1.0 + eps * rng.random(3))).astype(np.float32)
        if rng.random() < 0.5:  # and another plane crossed, so the first plane is not p
            q = rng.integers(0, 6)
            ax2 = q // 2
            if ax2 != axis:
                P[0, ax2] = np.float32(2.0 * w[0] * (1 if q % 2 == 0 else -1))
        v = np.zeros(3, O.RASTERIZER_VERTEX)
        v["x"], v["y"], v["z"], v["w"] = P[:, 0], P[:, 1], P[:, 2], P[:, 3]
        if fully_clipped_early(P):
            fired += 1
            poly, n = cpu_oracle.clip_triangle(v[0], v[1], v[2], 3, 2)
            assert n < 3, (i, P)
    assert fired > 1000
