// swr_geometry.hip — the geometry stage of the draw path on gfx950.
//
// One thread per input triangle.  A count pass and an emit pass run the same device code
// (process_triangle<NV, EMIT>): vertex program, clip masks and Sutherland-Hodgman clipping
// against the six frustum planes, perspective divide + viewport, facing cull, triangle
// setup and the tile rectangle each surviving triangle touches.  Between the passes two
// exclusive scans turn per-triangle counts into draw-order positions, so the emit pass
// writes primitive records and (tile, primitive) pairs in the reference's draw order with
// no sorting and no atomics:
//   key order = (batch, phase, j, k): per 1024-triangle batch, first triangle j's fan
//   triangle 0 (written in place, VertexProcessor.cpp:254-256), then the fan triangles
//   k >= 1 of every triangle in j order (appended, VertexProcessor.cpp:257-261).
//
// The vertex cache of the reference (VertexCache.h:29-63, used at VertexProcessor.cpp:86-117)
// maps an index to an already shaded output vertex within a batch.  Vertex programs are
// pure, so the results do not depend on which reference to a vertex shades it.  Here the
// vertex pass (k_vertex) shades every vertex of the attribute buffer once, in index order,
// into post-transform arrays (window-space position and clip mask), which both geometry
// passes read by index; the emit pass fetches only the varyings the vertex program copies.
// Triangles with a vertex outside a clip plane (clip mask != 0) re-run the program for the
// clip-space position the clipper needs.  Without the vertex pass (a batch of scenes, an
// attribute buffer of unknown size, or the tuning's choice) every corner runs the program.
#include "swr_common.h"
#include "swr_internal.h"
#include "swr_vertex.h"

namespace swr {

// ---- primitive setup ------------------------------------------------------------------
struct Edge {
    float a, b, c;
    bool tie;
};

// EdgeEquation::init, EdgeEquation.h:37-43
__device__ inline Edge edge_init(float x0, float y0, float x1, float y1)
{
    Edge e;
    e.a = y0 - y1;
    e.b = x1 - x0;
    e.c = -(e.a * (x0 + x1) + e.b * (y0 + y1)) / 2.0f;
    e.tie = e.a != 0.0f ? e.a > 0.0f : e.b > 0.0f;
    return e;
}

struct SpanPart {
    float ox, oy, is1, is2;
    int lo, hi;
};

struct PrimGeo {
    Edge e0, e1, e2;
    float area2;
    uint32_t flags;
    int tx0, ty0, tx1, ty1;  // screen bins
    int kx0, ky0, kx1, ky1;  // Block: the 8x8 blocks of the block grid
    SpanPart bot, top;
};

__device__ inline bool finitef(float x) { return __builtin_isfinite(x); }

// x extent [xl, xr) of the span on scanline sy of one flat part (Rasterizer.h:370-380)
__device__ inline void span_row(const GeomArgs &a, const SpanPart &p, int sy, int &xl, int &xr)
{
    float dy = ((float)sy - p.oy) + 0.5f;
    float cx1 = p.ox + p.is1 * dy + 0.5f;
    float cx2 = p.ox + p.is2 * dy + 0.5f;
    int f1 = f2i(cx1), f2 = f2i(cx2);
    xl = imax(a.sc_minx, f1);
    xr = imin(a.sc_maxx, f2);
    if (f1 == INT_MIN || f2 == INT_MIN) {  // non-monotonic sentinel: be conservative
        xl = a.sc_minx;
        xr = a.sc_maxx;
    }
}

// Adds the tiles a flat part may cover to the running rectangle.
__device__ inline void span_part_rect(const GeomArgs &a, const SpanPart &p, long long r0,
                                      long long r1, int &tx0, int &ty0, int &tx1, int &ty1)
{
    if (r0 < 0) r0 = 0;
    if (r1 > a.height - 1) r1 = a.height - 1;
    if (r0 > r1)
        return;
    int xl0, xr0, xl1, xr1;
    span_row(a, p, (int)r0, xl0, xr0);
    span_row(a, p, (int)r1, xl1, xr1);
    long long x0 = imin(xl0, xl1), x1 = (long long)imax(xr0, xr1) - 1;
    if (x0 < 0) x0 = 0;
    if (x1 > a.width - 1) x1 = a.width - 1;
    if (x0 > x1)
        return;
    tx0 = imin(tx0, (int)(x0 >> a.bin_shift));
    tx1 = imax(tx1, (int)(x1 >> a.bin_shift));
    ty0 = imin(ty0, (int)(r0 >> a.bin_shift));
    ty1 = imax(ty1, (int)(r1 >> a.bin_shift));
}

// Rasterizer::drawTriangleSpanTemplate, Rasterizer.h:318-357: vertex order and the two
// flat parts (drawBottomFlatTriangle :360-385, drawTopFlatTriangle :387-411).
__device__ inline void span_setup(const GeomArgs &a, const WinV &v0, const WinV &v1,
                                  const WinV &v2, PrimGeo &g)
{
    float tx = v0.x, ty = v0.y, mx = v1.x, my = v1.y, bx = v2.x, by = v2.y, sx, sy;
#define SWAP_V(P, Q) { sx = P##x; sy = P##y; P##x = Q##x; P##y = Q##y; Q##x = sx; Q##y = sy; }
    if (ty > my) SWAP_V(t, m);
    if (my > by) SWAP_V(m, b);
    if (ty > my) SWAP_V(t, m);
#undef SWAP_V
    float dy = by - ty;
    float iy = my - ty;
    g.flags &= ~(FLAG_BOTTOM | FLAG_TOP);
    auto bottom = [&](float x0, float y0, float x1, float y1, float x2, float y2) {
        g.bot.is1 = (x1 - x0) / (y1 - y0);
        g.bot.is2 = (x2 - x0) / (y2 - y0);
        g.bot.ox = x0;
        g.bot.oy = y0;
        g.bot.lo = f2i(y0 + 0.5f);
        g.bot.hi = f2i(y1 + 0.5f);
        g.flags |= FLAG_BOTTOM;
    };
    auto top = [&](float x0, float y0, float x1, float y1, float x2, float y2) {
        g.top.is1 = (x2 - x0) / (y2 - y0);
        g.top.is2 = (x2 - x1) / (y2 - y1);
        g.top.ox = x2;
        g.top.oy = y2;
        g.top.hi = f2i(y2 - 0.5f);
        g.top.lo = f2i(y0 - 0.5f);
        g.flags |= FLAG_TOP;
    };
    if (my == ty) {
        float lx = mx, ly = my, rx = tx, ry = ty;
        if (lx > rx) { float s0 = lx, s1 = ly; lx = rx; ly = ry; rx = s0; ry = s1; }
        top(lx, ly, rx, ry, bx, by);
    } else if (my == by) {
        float lx = mx, ly = my, rx = bx, ry = by;
        if (lx > rx) { float s0 = lx, s1 = ly; lx = rx; ly = ry; rx = s0; ry = s1; }
        bottom(tx, ty, lx, ly, rx, ry);
    } else {
        float v4x = tx + ((bx - tx) / dy) * iy, v4y = my;  // split vertex (Rasterizer.h:344-346)
        float lx = mx, ly = my, rx = v4x, ry = v4y;
        if (lx > rx) { float s0 = lx, s1 = ly; lx = rx; ly = ry; rx = s0; ry = s1; }
        bottom(tx, ty, lx, ly, rx, ry);
        top(lx, ly, rx, ry, bx, by);
    }
    int tx0 = INT_MAX, ty0 = INT_MAX, tx1 = INT_MIN, ty1 = INT_MIN;
    if (g.flags & FLAG_BOTTOM)
        span_part_rect(a, g.bot, (long long)g.bot.lo, (long long)g.bot.hi - 1, tx0, ty0, tx1, ty1);
    if (g.flags & FLAG_TOP)
        span_part_rect(a, g.top, (long long)g.top.lo + 1, (long long)g.top.hi, tx0, ty0, tx1, ty1);
    g.tx0 = tx0;
    g.ty0 = ty0;
    g.tx1 = tx1;
    g.ty1 = ty1;
}

// Rasterizer::drawTriangleBlockTemplate, Rasterizer.h:234-258: the block grid as a rect of
// screen bins.
__device__ inline void block_setup(const GeomArgs &a, const WinV &v0, const WinV &v1,
                                   const WinV &v2, PrimGeo &g)
{
    long long minX = f2i(ref_min(ref_min(v0.x, v1.x), v2.x));
    long long maxX = f2i(ref_max(ref_max(v0.x, v1.x), v2.x));
    long long minY = f2i(ref_min(ref_min(v0.y, v1.y), v2.y));
    long long maxY = f2i(ref_max(ref_max(v0.y, v1.y), v2.y));
    minX = minX > a.sc_minx ? minX : a.sc_minx;
    maxX = maxX < a.sc_maxx ? maxX : a.sc_maxx;
    minY = minY > a.sc_miny ? minY : a.sc_miny;
    maxY = maxY < a.sc_maxy ? maxY : a.sc_maxy;
    minX &= ~7LL;
    maxX &= ~7LL;
    minY &= ~7LL;
    maxY &= ~7LL;
    long long sx = (maxX - minX) / 8 + 1, sy = (maxY - minY) / 8 + 1;
    long long bx0, bx1, by0, by1;
    g.tx0 = INT_MAX;
    g.tx1 = INT_MIN;
    g.ty0 = INT_MAX;
    g.ty1 = INT_MIN;
    if (sx > 0 && sy > 0) {
        bx0 = minX; bx1 = minX + 8 * (sx - 1);
        by0 = minY; by1 = minY + 8 * (sy - 1);
    } else if (sx < 0 && sy < 0) {  // both negative: positive block count (Rasterizer.h:254-258)
        bx0 = minX; bx1 = minX + 8 * (-sx - 1);
        by0 = minY - 8 * (-sy - 1); by1 = minY;
    } else {
        return;
    }
    // the block grid, clamped to the target's blocks (pixels outside are never written)
    long long k0 = bx0 >> 3, k1 = bx1 >> 3, j0 = by0 >> 3, j1 = by1 >> 3;
    const long long kmax = ((long long)a.width + TILE - 1) / TILE - 1;
    const long long jmax = ((long long)a.height + TILE - 1) / TILE - 1;
    if (k0 < 0) k0 = 0;
    if (j0 < 0) j0 = 0;
    if (k1 > kmax) k1 = kmax;
    if (j1 > jmax) j1 = jmax;
    if (k0 > k1 || j0 > j1)
        return;
    g.kx0 = (int)k0;
    g.kx1 = (int)k1;
    g.ky0 = (int)j0;
    g.ky1 = (int)j1;
    // the bins holding those blocks
    const int sh = a.bin_shift - 3;
    g.tx0 = (int)(k0 >> sh);
    g.tx1 = (int)(k1 >> sh);
    g.ty0 = (int)(j0 >> sh);
    g.ty1 = (int)(j1 >> sh);
}

// Setup shared by both passes.  Returns false when the triangle draws nothing.
__device__ inline bool prim_geometry(const GeomArgs &a, const WinV &v0, const WinV &v1,
                                     const WinV &v2, PrimGeo &g)
{
    if (!(finitef(v0.x) && finitef(v0.y) && finitef(v1.x) && finitef(v1.y) && finitef(v2.x) &&
          finitef(v2.y)))
        return false;
    g.e0 = edge_init(v1.x, v1.y, v2.x, v2.y);
    g.e1 = edge_init(v2.x, v2.y, v0.x, v0.y);
    g.e2 = edge_init(v0.x, v0.y, v1.x, v1.y);
    g.area2 = g.e0.c + g.e1.c + g.e2.c;
    if (!(g.area2 > 0.0f) || !finitef(g.area2))  // TriangleEquations.h:53-56, Rasterizer.h:231
        return false;
    g.flags = (g.e0.tie ? FLAG_TIE0 : 0u) | (g.e1.tie ? FLAG_TIE1 : 0u) | (g.e2.tie ? FLAG_TIE2 : 0u);
    bool block = a.raster_mode == SWR_RASTER_BLOCK;
    if (a.raster_mode == SWR_RASTER_ADAPTIVE) {  // Rasterizer.h:413-428
        float minX = ref_min(ref_min(v0.x, v1.x), v2.x);
        float maxX = ref_max(ref_max(v0.x, v1.x), v2.x);
        float minY = ref_min(ref_min(v0.y, v1.y), v2.y);
        float maxY = ref_max(ref_max(v0.y, v1.y), v2.y);
        float orient = (maxX - minX) / (maxY - minY);
        block = (double)orient > 0.4 && (double)orient < 1.6;
    }
    if (block) {
        g.flags |= FLAG_BLOCK;
        block_setup(a, v0, v1, v2, g);
    } else {
        span_setup(a, v0, v1, v2, g);
    }
    return g.tx0 <= g.tx1 && g.ty0 <= g.ty1;
}

// The (key, value) pair of one candidate bin.  Bins in which the primitive cannot shade a
// pixel get the sentinel key and are dropped before the sort:
//   Block, 8x8 bins: the reference's corner classification (0 = skip), then
//     block_may_cover; the class rides in the pair value;
//   16x16 bins with cls4: the class of each of the bin's four blocks (in the block grid)
//     rides in the pair value, Span primitives a bit per block with pixels;
//   Block, other larger bins: dropped when no pixel of the bin can pass all three edge
//     tests (block_may_cover over the bin); the raster kernel classifies each 8x8 block;
//   Span: exact, the bin's scanline ranges are those the rasterizer computes.
__device__ inline void pair_for_tile(const GeomArgs &a, const float *rec, uint32_t flags,
                                     uint32_t pos, int tx, int ty, uint32_t &key, uint32_t &val)
{
    int tiles_x = a.tiles_x, sc_minx = a.sc_minx, sc_maxx = a.sc_maxx;
    uint32_t tile0 = 0;
    if (a.nscenes) {
        const SceneDev &S = a.scenes[__float_as_uint(rec[REC_SCENE])];
        tiles_x = S.bins_x[a.bin_shift - 3];
        sc_minx = S.sc_minx;
        sc_maxx = S.sc_maxx;
        tile0 = S.bin0[a.bin_shift - 3];
    }
    key = tile0 + (uint32_t)(ty * tiles_x + tx);
    val = pos;
    const int size = 1 << a.bin_shift;
    const int bx = tx << a.bin_shift, by = ty << a.bin_shift;
    if (flags & FLAG_BLOCK) {
        if (size == TILE) {
            int cls = classify_block(rec + REC_EDGE, flags, bx, by);
            if (cls == 1 && !block_may_cover(rec + REC_EDGE, bx, by))
                cls = 0;
            key = cls ? key : a.sentinel;
            val |= (uint32_t)cls << 30;
        } else if (a.cls4) {
            const uint32_t *ru = reinterpret_cast<const uint32_t *>(rec);
            uint32_t packed = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int kx = (bx >> 3) + (b & 1), ky = (by >> 3) + (b >> 1);
                int cls = 0;
                if (kx >= (int)ru[REC_BLOCKS] && kx <= (int)ru[REC_BLOCKS + 2] &&
                    ky >= (int)ru[REC_BLOCKS + 1] && ky <= (int)ru[REC_BLOCKS + 3]) {
                    cls = classify_block(rec + REC_EDGE, flags, kx << 3, ky << 3);
                    if (cls == 1 && !block_may_cover(rec + REC_EDGE, kx << 3, ky << 3))
                        cls = 0;
                }
                packed |= (uint32_t)cls << (2 * b);
            }
            key = packed ? key : a.sentinel;
            val |= packed << PAIR_CLS4_SHIFT;
        } else {
            key = block_may_cover(rec + REC_EDGE, bx, by, size - 1) ? key : a.sentinel;
        }
    } else if (a.cls4) {
        uint32_t packed = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            packed |= span_block_any(rec, flags, bx + (b & 1) * TILE, by + (b >> 1) * TILE,
                                     sc_minx, sc_maxx) ? 1u << (2 * b) : 0u;
        key = packed ? key : a.sentinel;
        val |= packed << PAIR_CLS4_SHIFT;
    } else {
        bool any = false;
        for (int r = 0; r < size && !any; ++r) {
            int xl, xr;
            bool rowv = span_row_extent(rec, flags, by + r, sc_minx, sc_maxx, xl, xr);
            any = rowv && imax(xl, bx) < imin(xr, bx + size);
        }
        key = any ? key : a.sentinel;
    }
}

// ParameterEquation::init, ParameterEquation.h:36-48
__device__ inline void param_write(float *rec, int slot, float p0, float p1, float p2,
                                   const PrimGeo &g, float f)
{
    rec[REC_PARAM + 3 * slot + 0] = f * (p0 * g.e0.a + p1 * g.e1.a + p2 * g.e2.a);
    rec[REC_PARAM + 3 * slot + 1] = f * (p0 * g.e0.b + p1 * g.e1.b + p2 * g.e2.b);
    rec[REC_PARAM + 3 * slot + 2] = f * (p0 * g.e0.c + p1 * g.e1.c + p2 * g.e2.c);
}

template <int NV>
__device__ __forceinline__ void write_prim(const GeomArgs &a, uint64_t pos, uint32_t key,
                                           uint32_t src, uint32_t sid, const PrimGeo &g,
                           const WinV &v0, const WinV &v1, const WinV &v2, const float *var0,
                           const float *var1, const float *var2, uint64_t pair_base)
{
    float *rec = a.prims + pos * (uint64_t)a.lay.stride;
    uint32_t *ru = reinterpret_cast<uint32_t *>(rec);
    ru[REC_KEY] = key;
    ru[REC_SRC] = src;
    ru[REC_SCENE] = sid;
    ru[REC_FLAGS] = g.flags;
    ru[REC_TILES + 0] = (uint32_t)g.tx0;
    ru[REC_TILES + 1] = (uint32_t)g.ty0;
    ru[REC_TILES + 2] = (uint32_t)g.tx1;
    ru[REC_TILES + 3] = (uint32_t)g.ty1;
    if (g.flags & FLAG_BLOCK) {  // Span coverage comes from the flat parts alone
        rec[REC_EDGE + 0] = g.e0.a;
        rec[REC_EDGE + 1] = g.e0.b;
        rec[REC_EDGE + 2] = g.e0.c;
        rec[REC_EDGE + 3] = g.e1.a;
        rec[REC_EDGE + 4] = g.e1.b;
        rec[REC_EDGE + 5] = g.e1.c;
        rec[REC_EDGE + 6] = g.e2.a;
        rec[REC_EDGE + 7] = g.e2.b;
        rec[REC_EDGE + 8] = g.e2.c;
    }
    if (g.flags & FLAG_BLOCK) {
        ru[REC_BLOCKS + 0] = (uint32_t)g.kx0;
        ru[REC_BLOCKS + 1] = (uint32_t)g.ky0;
        ru[REC_BLOCKS + 2] = (uint32_t)g.kx1;
        ru[REC_BLOCKS + 3] = (uint32_t)g.ky1;
    } else {
        const SpanPart *parts[2] = {&g.bot, &g.top};
        const uint32_t pf[2] = {FLAG_BOTTOM, FLAG_TOP};
        const int base[2] = {REC_BOT, REC_TOP};
        for (int q = 0; q < 2; ++q) {
            if (!(g.flags & pf[q]))
                continue;
            rec[base[q] + 0] = parts[q]->ox;
            rec[base[q] + 1] = parts[q]->oy;
            rec[base[q] + 2] = parts[q]->is1;
            rec[base[q] + 3] = parts[q]->is2;
            ru[base[q] + 4] = (uint32_t)parts[q]->lo;
            ru[base[q] + 5] = (uint32_t)parts[q]->hi;
        }
    }
    // TriangleEquations.h:58-70
    const ParamLayout &L = a.lay;
    float f = 1.0f / g.area2;
    int slot = 0;
    if (L.nz)
        param_write(rec, slot++, v0.z, v1.z, v2.z, g, f);
    float iw0 = 1.0f / v0.w, iw1 = 1.0f / v1.w, iw2 = 1.0f / v2.w;
    if (L.nw)
        param_write(rec, slot++, iw0, iw1, iw2, g, f);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        if (k < L.na) {
            if (k < L.set_na)
                param_write(rec, slot + k, var0[k], var1[k], var2[k], g, f);
            else
                param_write(rec, slot + k, 0.0f, 0.0f, 0.0f, g, 0.0f);
        } else if (k < L.na + L.np) {
            if (k - L.na < L.set_np)
                param_write(rec, slot + k, var0[k] * iw0, var1[k] * iw1, var2[k] * iw2, g, f);
            else
                param_write(rec, slot + k, 0.0f, 0.0f, 0.0f, g, 0.0f);
        }
    }
    ru[REC_PAIRBASE] = (uint32_t)pair_base;
    a.prim_pb[pos] = (uint32_t)pair_base;  // the expansion kernel searches these
}

// (tile, primitive) pairs, load-balanced: each workgroup owns EXPAND_CHUNK consecutive pair
// slots (256 to 4096; tunable).  Pair bases are increasing in draw order, so the primitives
// owning a chunk are a contiguous run found by one binary search; each thread then searches
// that run (in LDS) for the primitive of its slot and classifies the slot's tile (row-major
// rect order).  The tiles the primitive cannot shade are dropped here: the kept pairs are
// written in slot order to the start of the chunk and their count to chunk_kept[chunk], so
// the sort reads only kept pairs (without chunk_kept the chunk is written whole, rejected
// pairs keyed with the sentinel).
template <int EXPAND_CHUNK>
__global__ void __launch_bounds__(256) k_expand(GeomArgs a)
{
    __shared__ uint32_t s_pb[EXPAND_CHUNK + 1];
    __shared__ uint32_t s_p0, s_p1;
    __shared__ uint32_t s_w[4];
    if (a.status[ST_OVERFLOW] | a.status[ST_ERROR])
        return;
    const uint64_t M = a.status[ST_PAIRS], P = a.status[ST_PRIMS];
    const uint64_t c0 = (uint64_t)blockIdx.x * EXPAND_CHUNK;
    if (c0 >= M)
        return;
    const uint64_t c1 = c0 + EXPAND_CHUNK < M ? c0 + EXPAND_CHUNK : M;
    if (threadIdx.x < 64) {
        // the last primitive whose first pair is <= target, for the chunk's first slot (lanes
        // 0-31) and its last (lanes 32-63): a 32-way search, one round of parallel loads per
        // step (log32 P steps instead of log2 P dependent loads)
        const uint32_t lane = threadIdx.x, sub = lane & 31u;
        const uint64_t target = lane < 32 ? c0 : c1 - 1;
        uint64_t lo = 0, hi = P;  // invariant: pb[lo] <= target (pb[0] = 0), answer in [lo, hi)
        while (__ballot(hi - lo > 1) != 0ull) {
            const uint64_t span = hi - lo, step = (span + 31) / 32;
            const uint64_t idx = lo + (uint64_t)sub * step;
            const bool le = span > 1 && idx < hi && a.prim_pb[idx] <= target;
            const unsigned long long m = __ballot(le);
            const uint32_t half = lane < 32 ? (uint32_t)m : (uint32_t)(m >> 32);
            if (span > 1) {  // sub 0 (idx = lo) always qualifies, so half != 0
                lo += (uint64_t)(31 - __builtin_clz(half)) * step;
                hi = lo + step < hi ? lo + step : hi;
            }
        }
        if (lane == 0)
            s_p0 = (uint32_t)lo;
        if (lane == 32)
            s_p1 = (uint32_t)lo;
    }
    __syncthreads();
    const uint32_t p0 = s_p0, np = s_p1 - s_p0 + 1;
    for (uint32_t i = threadIdx.x; i <= np; i += blockDim.x)
        s_pb[i] = (p0 + i < P) ? a.prim_pb[p0 + i] : (uint32_t)M;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t kept = 0;  // pairs kept by the earlier rounds (uniform)
    for (uint64_t r0 = c0; r0 < c1; r0 += 256) {  // blockDim.x == 256
        const uint64_t q = r0 + threadIdx.x;
        uint32_t key = a.sentinel, val = 0;
        if (q < c1) {
            uint32_t lo = 0, hi = np;
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (s_pb[mid] <= (uint32_t)q)
                    lo = mid;
                else
                    hi = mid;
            }
            const uint32_t pos = p0 + lo;
            const uint32_t local = (uint32_t)q - s_pb[lo];
            const float *rec = a.prims + (uint64_t)pos * (uint64_t)a.lay.stride;
            const uint32_t *ru = reinterpret_cast<const uint32_t *>(rec);
            const int tx0 = (int)ru[REC_TILES + 0], ty0 = (int)ru[REC_TILES + 1];
            const uint32_t w0 = ru[REC_TILES + 2] - ru[REC_TILES + 0] + 1u;
            const int tx = tx0 + (int)(local % w0), ty = ty0 + (int)(local / w0);
            pair_for_tile(a, rec, ru[REC_FLAGS], pos, tx, ty, key, val);
        }
        if (!a.chunk_kept) {
            if (q < c1) {
                a.pair_keys[q] = key;
                a.pair_vals[q] = val;
            }
            continue;
        }
        // block-wide rank of the kept pairs of this round, in slot order
        const bool keep = q < c1 && key != a.sentinel;
        const unsigned long long m = __ballot(keep);
        if (lane == 0)
            s_w[w] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            pre += k < (int)w ? s_w[k] : 0u;
            tot += s_w[k];
        }
        __syncthreads();  // s_w is written again by the next round
        if (keep) {
            const uint64_t o = c0 + kept + pre + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            a.pair_keys[o] = key;
            a.pair_vals[o] = val;
        }
        kept += tot;
    }
    if (a.chunk_kept && threadIdx.x == 0)
        a.chunk_kept[blockIdx.x] = kept;
}

__device__ inline uint64_t rect_tiles(const PrimGeo &g)
{
    return (uint64_t)(g.tx1 - g.tx0 + 1) * (uint64_t)(g.ty1 - g.ty0 + 1);
}

// Per-input-triangle bookkeeping shared by the fan triangles it produces.
struct FanCtx {
    uint32_t t, b, j;               // virtual triangle, its batch, its slot in the batch
    uint32_t src, kb, sid;          // scene-local triangle and batch, scene
    uint32_t slot1;                 // batch-local slot of fan triangle 1
    uint64_t pos0, pairs0, pos1, pairs1;
    uint32_t cnt0, cntp0, cnt1, cntp1;
    uint32_t app_n;                 // count pass: fan triangles appended (k >= 1)
};

// Facing cull and setup of one fan triangle (VertexProcessor::cullTriangles,
// VertexProcessor.cpp:330-343).  A triangle the cull mode keeps with facing >= 0 has its
// vertices 0 and 2 swapped (VertexProcessor.cpp:342): w0 / w2 come back swapped and `swap`
// tells the caller to swap the varyings too.  Returns false when it draws nothing.
__device__ __forceinline__ bool fan_geometry(const GeomArgs &a, WinV &w0, const WinV &w1,
                                             WinV &w2, bool cull, PrimGeo &g, bool &swap)
{
    swap = false;
    if (cull) {
        float facing = (w0.x - w1.x) * (w2.y - w1.y) - (w2.x - w1.x) * (w0.y - w1.y);
        if (facing < 0) {
            if (a.cull_mode == SWR_CULL_CW)
                return false;
        } else {
            if (a.cull_mode == SWR_CULL_CCW)
                return false;
            swap = true;
        }
    }
    // component by component: a select between the two structs would address them in memory
    const WinV s0 = w0, s2 = w2;
    w0.x = swap ? s2.x : s0.x;
    w0.y = swap ? s2.y : s0.y;
    w0.z = swap ? s2.z : s0.z;
    w0.w = swap ? s2.w : s0.w;
    w2.x = swap ? s0.x : s2.x;
    w2.y = swap ? s0.y : s2.y;
    w2.z = swap ? s0.z : s2.z;
    w2.w = swap ? s0.w : s2.w;
    return prim_geometry(a, w0, w1, w2, g);
}

// The emit pass's write of one fan triangle at record `pos` with its pairs from `pb`.
template <int NV>
__device__ __forceinline__ void fan_write(const GeomArgs &a, uint64_t pos, uint64_t pb,
                                          uint32_t nt, uint32_t key, uint32_t src, uint32_t sid,
                                          const PrimGeo &g, const WinV &w0, const WinV &w1,
                                          const WinV &w2, const float *v0, const float *v1,
                                          const float *v2, bool swap)
{
    if (pos >= a.prim_cap || pb + nt > a.pair_cap) {  // totals are checked; this guards wrap
        atomicOr(&a.status[ST_OVERFLOW], 4ull);
        return;
    }
    // swap by value: selecting between the two arrays by pointer would force them to memory
    constexpr int NVS = NV > 0 ? NV : 1;
    float x0[NVS], x2[NVS];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        x0[q] = swap ? v2[q] : v0[q];
        x2[q] = swap ? v0[q] : v2[q];
    }
    write_prim<NV>(a, pos, key, src, sid, g, w0, w1, w2, x0, v1, x2, pb);
}

// One fan triangle (p0, p[k+1], p[k+2]): cull, setup, count or emit.
template <int NV, bool EMIT>
__device__ __forceinline__ void fan_triangle(const GeomArgs &a, FanCtx &c, int k, WinV w0, WinV w1,
                                    WinV w2, const float *v0, const float *v1, const float *v2,
                                    bool cull)
{
    bool swap;
    PrimGeo g;
    if (!fan_geometry(a, w0, w1, w2, cull, g, swap))
        return;
    uint32_t nt = (uint32_t)rect_tiles(g);
    if (!EMIT) {
        if (k == 0) {
            c.cnt0 += 1;
            c.cntp0 += nt;
        } else {
            c.cnt1 += 1;
            c.cntp1 += nt;
        }
        return;
    }
    uint32_t key = a.window_input ? c.src
                                  : ((c.kb << 15) + (k == 0 ? c.j : c.slot1 + (uint32_t)(k - 1)));
    uint64_t pos, pb;
    if (k == 0) {
        pos = c.pos0;
        pb = c.pairs0;
    } else {
        pos = c.pos1 + c.cnt1;
        pb = c.pairs1 + c.cntp1;
        c.cnt1 += 1;
        c.cntp1 += nt;
    }
    fan_write<NV>(a, pos, pb, nt, key, c.src, c.sid, g, w0, w1, w2, v0, v1, v2, swap);
}

// All three vertices are outside the first plane the clipper applies (the lowest set bit of
// the OR of the clip masks, VertexProcessor.cpp:237-242): that pass keeps no vertex and
// creates none (PolyClipper.cpp:64-74), so the triangle is fully clipped.  For finite
// coordinates the plane value a*x + b*y + c*z + d*w equals the clip-mask difference
// (w - x, x + w, ...) exactly, so the test is exact.
//
// Beyond the first plane: every vertex the clipper can create is a lerp v0*(1-t) + v1*t
// (t in [0,1]) of earlier polygon vertices, so a convex combination of the three input
// vertices up to rounding: at most 5 lerp levels of ~4 units in the last place plus the
// plane evaluation, under 1.5e-6 of M = max_i (|x|+|y|+|z|+|w|).  If all three vertices lie
// outside a plane by more than 1e-5 M, every polygon reaching that plane lies outside it,
// that pass empties the polygon, and the triangle is fully clipped.
__device__ inline bool fully_clipped_early(const Pos4 *P, int cm)
{
    for (int q = 0; q < 3; ++q)
        if (!(finitef(P[q].x) && finitef(P[q].y) && finitef(P[q].z) && finitef(P[q].w)))
            return false;
    const int all_out = clip_mask(P[0]) & clip_mask(P[1]) & clip_mask(P[2]);
    if (all_out & (cm & -cm))
        return true;
    if (!all_out)
        return false;
    float m = 0.0f;
    for (int q = 0; q < 3; ++q)
        m = fmaxf(m, fabsf(P[q].x) + fabsf(P[q].y) + fabsf(P[q].z) + fabsf(P[q].w));
    const float margin = -1e-5f * m;
    for (int p = 0; p < 6; ++p) {
        if (!(all_out & (1 << p)))
            continue;
        bool far_out = true;
        for (int q = 0; q < 3; ++q)
            far_out = far_out && plane_dot(p, P[q]) < margin;
        if (far_out)
            return true;
    }
    return false;
}

// Exactly one plane selected, finite coordinates, no vertex on that plane (dp != 0): the
// register-only clip_one_plane applies.  Fills dp with the plane values.
__device__ __forceinline__ bool one_plane_case(const Pos4 *P, int cm, float *dp)
{
    if (cm & (cm - 1))
        return false;
    const int p = __builtin_ctz(cm);
    bool ok = true;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        ok = ok && finitef(P[q].x) && finitef(P[q].y) && finitef(P[q].z) && finitef(P[q].w);
        dp[q] = plane_dot(p, P[q]);
        ok = ok && dp[q] != 0.0f;
    }
    return ok;
}

// The clip queue is split over CLIPQ_REGIONS counters in separate cache lines (block b uses
// region b % CLIPQ_REGIONS) so that enqueueing does not serialise on one address.
constexpr uint32_t CLIPQ_REGIONS = 64;
constexpr uint32_t CLIPQ_STRIDE = 16;  // u32 words between counters
static_assert(CLIPQ_REGIONS * CLIPQ_STRIDE == CLIPQ_COUNTERS, "the counter buffer layout");

__device__ inline uint32_t clipq_region_cap(uint32_t tris)
{
    uint32_t blocks = (tris + 255) / 256;
    return ((blocks + CLIPQ_REGIONS - 1) / CLIPQ_REGIONS) * 256;
}

// One clip plane, no vertex on it: the Sutherland-Hodgman pass (PolyClipper.cpp:52-78) over
// (v0,v1), (v1,v2), (v2,v0) keeps each inside vertex and inserts I_ab = lerp(a, b, t) where
// the sign changes, so with s_i = vertex i inside the polygon is one of six fixed sequences:
//   (1,1,0) v0 v1 I12 I20   (1,0,1) v0 I01 I12 v2   (0,1,1) I01 v1 v2 I20
//   (1,0,0) v0 I01 I20      (0,1,0) I01 v1 I12      (0,0,1) I12 v2 I20
// Fan triangles (p0, p1, p2) and (p0, p2, p3) follow (VertexProcessor.cpp:254-261).
template <int NV>
struct CVert {
    Pos4 p;
    float v[NV];
};

template <int NV, bool EMIT>
__device__ __forceinline__ CVert<NV> intersect(const GeomArgs &a, const Pos4 *P, const float (*var)[NV],
                                      const float *dp, int ia, int ib)
{
    const float dprev = dp[ia], dcur = dp[ib];
    const float t = dcur < 0 ? dprev / (dprev - dcur) : -dprev / (dcur - dprev);
    CVert<NV> r;
    r.p.x = lerp1(P[ia].x, P[ib].x, t);
    r.p.y = lerp1(P[ia].y, P[ib].y, t);
    r.p.z = lerp1(P[ia].z, P[ib].z, t);
    r.p.w = lerp1(P[ia].w, P[ib].w, t);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        bool lerped = k < a.lay.na ? (k < a.vs_na) : ((k - a.lay.na) < a.vs_np);
        r.v[k] = (EMIT && lerped) ? lerp1(var[ia][k], var[ib][k], t) : 0.0f;
    }
    return r;
}

template <int NV>
__device__ __forceinline__ CVert<NV> pick(bool c1, const CVert<NV> &x1, bool c2, const CVert<NV> &x2,
                                 const CVert<NV> &x3)
{
    CVert<NV> r;
    r.p.x = c1 ? x1.p.x : (c2 ? x2.p.x : x3.p.x);
    r.p.y = c1 ? x1.p.y : (c2 ? x2.p.y : x3.p.y);
    r.p.z = c1 ? x1.p.z : (c2 ? x2.p.z : x3.p.z);
    r.p.w = c1 ? x1.p.w : (c2 ? x2.p.w : x3.p.w);
#pragma unroll
    for (int k = 0; k < NV; ++k)
        r.v[k] = c1 ? x1.v[k] : (c2 ? x2.v[k] : x3.v[k]);
    return r;
}

template <int NV>
__device__ __forceinline__ void put(WinV &w, float *v, const GeomArgs &a, const CVert<NV> &x)
{
    w = to_window(a, x.p);
#pragma unroll
    for (int k = 0; k < NV; ++k)
        v[k] = x.v[k];
}

// Writes the clipped polygon to W / V (3 or 4 vertices) and returns its fan count.
template <int NV, bool EMIT>
__device__ __forceinline__ int clip_one_plane(const GeomArgs &a, const Pos4 *P,
                                              const float (*var)[NV], const float *dp, WinV *W,
                                              float (*V)[NV])
{
    CVert<NV> v0, v1, v2;
    v0.p = P[0];
    v1.p = P[1];
    v2.p = P[2];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        v0.v[k] = var[0][k];
        v1.v[k] = var[1][k];
        v2.v[k] = var[2][k];
    }
    const CVert<NV> i01 = intersect<NV, EMIT>(a, P, var, dp, 0, 1);
    const CVert<NV> i12 = intersect<NV, EMIT>(a, P, var, dp, 1, 2);
    const CVert<NV> i20 = intersect<NV, EMIT>(a, P, var, dp, 2, 0);
    const int cs = (dp[0] > 0 ? 1 : 0) | (dp[1] > 0 ? 2 : 0) | (dp[2] > 0 ? 4 : 0);
    const bool quad = cs == 3 || cs == 5 || cs == 6;
    // o0: v0 (1xx), I01 (010, 011), I12 (001)
    const CVert<NV> o0 = pick(cs & 1, v0, cs == 2 || cs == 6, i01, i12);
    // o1: v1 (110, 011, 010), I01 (101, 100), v2 (001)
    const CVert<NV> o1 = pick(cs == 3 || cs == 6 || cs == 2, v1, cs == 5 || cs == 1, i01, v2);
    // o2: I12 (110, 101, 010), v2 (011), I20 (100, 001)
    const CVert<NV> o2 = pick(cs == 3 || cs == 5 || cs == 2, i12, cs == 6, v2, i20);
    // o3: I20 (110, 011), v2 (101)
    const CVert<NV> o3 = pick(cs == 5, v2, false, v2, i20);
    put(W[0], V[0], a, o0);
    put(W[1], V[1], a, o1);
    put(W[2], V[2], a, o2);
    if (quad)
        put(W[3], V[3], a, o3);
    return quad ? 2 : 1;
}

// The main pass over all input triangles.  Triangles inside every plane, fully clipped or
// crossing one plane are finished here; the others go to the clip queue (count pass) or are
// left to the queue kernel (emit pass), which clips them a wave at a time (clip_queued).
// LISTED: the emit pass over the compacted list of triangles the count pass found live (a
// workgroup's triangles may then come from several batches and scenes).
//
// The count pass writes app[t] = (fan triangles appended, live): live when the main pass
// itself emits a primitive for t (a triangle the clip queue takes is not live).  Their scan
// gives the appended fans' draw-order slots and, with the live flags, the compacted list.
template <int NV, bool EMIT, bool LISTED = false>
__device__ void process_triangle(const GeomArgs &a0, uint32_t t, uint32_t usid = 0)
{
    GeomArgs a = a0;
    FanCtx c;
    c.t = t;
    c.b = t >> 10;
    c.j = t & 1023u;
    c.src = t;
    c.kb = c.b;
    c.sid = 0;
    c.slot1 = 0;
    c.pos0 = c.pairs0 = c.pos1 = c.pairs1 = 0;
    c.cnt0 = c.cntp0 = c.cnt1 = c.cntp1 = 0;
    c.app_n = 0;
    uint32_t ntris = a.tris, nbatches = a.batches;  // extent of this triangle's draw
    bool live = true;
    if (a0.nscenes) {
        // a 256-thread block of the main pass lies inside one 1024-triangle batch; the listed
        // emit hands in the wave-uniform scene of its waterfall step
        uint32_t sid = LISTED ? usid : a0.batch_scene[c.b];
        if (!LISTED)
            sid = __builtin_amdgcn_readfirstlane(sid);
        const SceneDev &S = a0.scenes[sid];
        apply_scene(a, S);
        c.sid = sid;
        c.kb = c.b - S.batch0;
        c.src = t - S.batch0 * (uint32_t)BATCH;
        ntris = S.tris;
        nbatches = (S.tris + BATCH - 1) / BATCH;
        live = c.src < S.tris;  // else padding up to the scene's last batch boundary
    }
    const uint32_t nbt = (c.kb + 1 < nbatches) ? (uint32_t)BATCH : ntris - c.kb * (uint32_t)BATCH;
    if (!live) {
        if (!EMIT) {
            a.app[t] = make_uint2(0, 0);
            a.grp[(size_t)c.b * 2048 + c.j] = make_uint2(0, 0);
            a.grp[(size_t)c.b * 2048 + 1024 + c.j] = make_uint2(0, 0);
        }
        return;
    }
    if (EMIT) {
        if (!a.window_input)
            c.slot1 = nbt + (a.app_ex[t].x - a.app_ex[c.b * BATCH].x);
        uint2 g0 = a.grp_ex[(size_t)c.b * 2048 + c.j];
        uint2 g1 = a.grp_ex[(size_t)c.b * 2048 + 1024 + c.j];
        c.pos0 = g0.x;
        c.pairs0 = g0.y;
        c.pos1 = g1.x;
        c.pairs1 = g1.y;
    }
    constexpr int NVS = NV > 0 ? NV : 1;
    float var3[3][NVS];
    // the polygon handed to the fan loop below: 3 or 4 window-space vertices
    WinV W[4];
    float V[4][NVS];
    int nfan = 0;
    bool cull = true;

    if (a.window_input) {
        // IRasterizer seam: window-space vertices, no clip, transform or cull
        const uint64_t ib = 3ull * c.src;
        int32_t i0 = a.indices[ib], i1 = a.indices[ib + 1], i2 = a.indices[ib + 2];
        if (i0 != -1) {
            if (i0 < 0 || i1 < 0 || i2 < 0 || (uint64_t)i0 >= a.wvert_count ||
                (uint64_t)i1 >= a.wvert_count || (uint64_t)i2 >= a.wvert_count) {
                atomicOr(&a.status[ST_ERROR], 1ull);
            } else {
                const int32_t ids[3] = {i0, i1, i2};
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const swr_rasterizer_vertex &v = a.wverts[ids[q]];
                    W[q].x = v.x;
                    W[q].y = v.y;
                    W[q].z = v.z;
                    W[q].w = v.w;
                    if (EMIT) {
#pragma unroll
                        for (int k = 0; k < NV; ++k) {
                            int q2 = k - a.lay.na;
                            V[q][k] = k < a.lay.na ? v.avar[k & 15]
                                                   : (q2 < a.lay.np ? v.pvar[q2 & 15] : 0.0f);
                        }
                    }
                }
                nfan = 1;
                cull = false;
            }
        }
    } else {
        const uint64_t ib = 3ull * c.src;
        int32_t ids[3] = {a.indices[ib], a.indices[ib + 1], a.indices[ib + 2]};
        bool bad = false;
        for (int q = 0; q < 3; ++q)
            if (ids[q] < 0 ||
                (a.attr0_bytes && (uint64_t)ids[q] * (uint64_t)a.stride0 + a.attr0_need > a.attr0_bytes))
                bad = true;
        if (bad) {
            atomicOr(&a.status[ST_ERROR], 1ull);
        } else {
            int cmv = -1;  // clip mask from the post-transform vertices
            if (a.vwin)
                cmv = a.vmask[ids[0]] | a.vmask[ids[1]] | a.vmask[ids[2]];
            Pos4 P[3];
            if (cmv == 0) {
                // inside every plane: the shaded window-space vertices as they are
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const float4 w = a.vwin[ids[q]];
                    W[q].x = w.x;
                    W[q].y = w.y;
                    W[q].z = w.z;
                    W[q].w = w.w;
                    if (EMIT) {
                        fetch_vertex<NV, false>(a, ids[q], P[q], var3[q]);
#pragma unroll
                        for (int k = 0; k < NVS; ++k)
                            V[q][k] = var3[q][k];
                    }
                }
                nfan = 1;
                goto fans;
            }
#pragma unroll
            for (int q = 0; q < 3; ++q)
                fetch_vertex<EMIT ? NV : 0>(a, ids[q], P[q], var3[q]);
            int cm = clip_mask(P[0]) | clip_mask(P[1]) | clip_mask(P[2]);
            float dp1[3];
            if (cm == 0) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    W[q] = to_window(a, P[q]);
#pragma unroll
                    for (int k = 0; k < NVS; ++k)
                        V[q][k] = var3[q][k];
                }
                nfan = 1;
            } else if (fully_clipped_early(P, cm)) {
                nfan = 0;
            } else if (one_plane_case(P, cm, dp1)) {
                nfan = clip_one_plane<NVS, EMIT>(a, P, var3, dp1, W, V);
            } else {
                if (!EMIT) {  // the queue kernels clip it and write its counts
                    const uint32_t r = blockIdx.x % CLIPQ_REGIONS;
                    uint32_t q = atomicAdd(&a.clipq_cnt[r * CLIPQ_STRIDE], 1u);
                    a.clipq[r * clipq_region_cap(a.tris) + q] = t;
                }
                return;
            }
            c.app_n = nfan == 2 ? 1u : 0u;
        }
    }
fans:
    // fan triangles (p0, p[k+1], p[k+2]), VertexProcessor.cpp:254-261; one call site so the
    // vertex arrays stay in registers
#pragma unroll 1
    for (int k = 0; k < nfan; ++k) {
        const WinV wb = k == 0 ? W[1] : W[2];
        const WinV wc = k == 0 ? W[2] : W[3];
        float vb[NVS], vc[NVS];
#pragma unroll
        for (int q = 0; q < NVS; ++q) {
            vb[q] = k == 0 ? V[1][q] : V[2][q];
            vc[q] = k == 0 ? V[2][q] : V[3][q];
        }
        fan_triangle<NV, EMIT>(a, c, k, W[0], wb, wc, V[0], vb, vc, cull);
    }
    if (!EMIT) {
        a.grp[(size_t)c.b * 2048 + c.j] = make_uint2(c.cnt0, c.cntp0);
        a.grp[(size_t)c.b * 2048 + 1024 + c.j] = make_uint2(c.cnt1, c.cntp1);
        if (!a.window_input)
            a.app[t] = make_uint2(c.app_n, c.cnt0 + c.cnt1 > 0 ? 1u : 0u);
    }
}

__global__ void __launch_bounds__(256) k_begin(GeomArgs a)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ST_WORDS && i != ST_STICKY)
        a.status[i] = 0;
    if (i < CLIPQ_REGIONS)
        a.clipq_cnt[i * CLIPQ_STRIDE] = 0;
    if (i < RASTER_QUEUES)
        a.clipq_cnt[CLIPQ_COUNTERS + i * RASTER_QUEUE_STRIDE] = 0;
    for (uint32_t k = i; k < 2 * a.tiles; k += gridDim.x * blockDim.x)
        a.ranges[k] = 0;
    // the last batch's unused (batch, phase, j) slots count nothing
    uint32_t tail = a.batches * 1024u - a.tris;
    for (uint32_t k = i; k < tail; k += gridDim.x * blockDim.x) {
        uint32_t tt = a.tris + k;
        uint32_t bb = tt >> 10, jj = tt & 1023u;
        a.grp[(size_t)bb * 2048 + jj] = make_uint2(0, 0);
        a.grp[(size_t)bb * 2048 + 1024 + jj] = make_uint2(0, 0);
    }
}

// The vertex pass: every vertex of attribute 0 shaded once (VertexProcessor.cpp:134-150 with
// the built-in program), its clip mask (VertexProcessor.cpp:122-132) and its window-space
// position (VertexProcessor.cpp:364-373; meaningful where the mask is 0, which is the only
// case that reads it).
__global__ void __launch_bounds__(256) k_vertex(GeomArgs a)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.nverts)
        return;
    Pos4 P;
    float none[1];
    fetch_vertex<0>(a, (int32_t)v, P, none);
    const WinV w = to_window(a, P);
    a.vwin[v] = make_float4(w.x, w.y, w.z, w.w);
    a.vmask[v] = (uint8_t)clip_mask(P);
}

__global__ void __launch_bounds__(256) k_geom_count(GeomArgs a)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.tris)
        process_triangle<0, false>(a, t);
}

template <int NV>
__global__ void __launch_bounds__(256) k_geom_emit(GeomArgs a)
{
    if (a.status[ST_OVERFLOW] | a.status[ST_ERROR])
        return;
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.tris)
        process_triangle<NV, true>(a, t);
}

// The emit pass over the live triangles only, in draw order (the list the scan compacted).
// In a batch of scenes a wave's triangles may come from several scenes: the wave takes them
// one scene at a time (a waterfall over the scene of its first pending lane), so the scene's
// state stays wave-uniform (scalar loads and registers).  One thread per live triangle, no
// grid-stride loop: a loop keeps every argument it uses live in scalar registers across
// iterations, and their spills to vector lanes cost this latency-bound kernel time.
template <int NV, bool SCENES>
__device__ __forceinline__ void emit_listed_one(const GeomArgs &a, uint32_t n, uint32_t i)
{
    if (i >= n)
        return;
    const uint32_t t = a.live[i];
    if (!SCENES) {
        process_triangle<NV, true, true>(a, t);
        return;
    }
    const uint32_t sid = a.batch_scene[t >> 10];
    bool pending = true;
    while (pending) {
        const uint32_t s0 = __builtin_amdgcn_readfirstlane(sid);
        if (sid == s0) {
            process_triangle<NV, true, true>(a, t, s0);
            pending = false;
        }
    }
}

template <int NV, bool SCENES>
__global__ void __launch_bounds__(256) k_geom_emit_listed(GeomArgs a)
{
    if (a.status[ST_OVERFLOW] | a.status[ST_ERROR])
        return;
    emit_listed_one<NV, SCENES>(a, (uint32_t)a.status[ST_LIVE],
                                blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- the clip queue: a wave per queued triangle -------------------------------------------
// Triangles that cross two or more planes (about 0.3% of a dense frame) are clipped by whole
// waves, one triangle per wave at a time, lane k holding polygon vertex k.  A pass of
// PolyClipper::clipToPlane (PolyClipper.cpp:45-81) walks the edges (v_k, v_k+1 mod n) in
// order and outputs v_k if it is inside (dpPrev >= 0), then the edge's intersection if the
// plane value changes sign along it.  Here every lane takes its own edge at once: vertex k's
// output slot is the number of inside lanes below k plus the number of crossing lanes below
// k (two ballots), its intersection goes right after it, and the new polygon is staged in
// LDS and read back a vertex per lane.  The fan triangles (VertexProcessor.cpp:254-261) are
// set up a fan per lane; the emit pass places fan k >= 1 after the drawn fans below it (a
// ballot count) and its pairs after theirs (a wave prefix sum), as a walk in fan order
// would.  A polygon that outgrows MAX_CLIP vertices fails the draw, as the oracle's clipper
// does.
constexpr int CLIPQ_THREADS = 64;  // one wave per workgroup
static_assert(MAX_CLIP <= CLIPQ_THREADS, "a polygon vertex per lane");

template <int NV>
struct ClipStage {
    Pos4 p[MAX_CLIP];
    float v[MAX_CLIP][NV];
};

// wave-wide inclusive prefix sum
__device__ inline uint32_t wave_incl_sum(uint32_t x, uint32_t lane)
{
#pragma unroll
    for (int sh = 1; sh < 64; sh <<= 1) {
        const uint32_t o = __shfl_up(x, sh, 64);
        x += lane >= (uint32_t)sh ? o : 0u;
    }
    return x;
}

__device__ inline WinV shfl_win(const WinV &w, int src)
{
    WinV r;
    r.x = __shfl(w.x, src);
    r.y = __shfl(w.y, src);
    r.z = __shfl(w.z, src);
    r.w = __shfl(w.w, src);
    return r;
}

// One queued triangle t, clipped by the calling wave (all 64 lanes, t uniform).  The count
// pass writes its (batch, phase, j) counts and appended-fan count; the emit pass its records.
template <int NV, bool EMIT>
__device__ void clip_queued(const GeomArgs &a0, uint32_t t, ClipStage<NV> &st)
{
    const uint32_t lane = threadIdx.x & 63u;
    t = __builtin_amdgcn_readfirstlane(t);
    GeomArgs a = a0;
    const uint32_t b = t >> 10, j = t & 1023u;
    uint32_t src = t, kb = b, sid = 0, ntris = a.tris, nbatches = a.batches;
    if (a0.nscenes) {
        sid = a0.batch_scene[b];
        const SceneDev &S = a0.scenes[sid];
        apply_scene(a, S);
        kb = b - S.batch0;
        src = t - S.batch0 * (uint32_t)BATCH;
        ntris = S.tris;
        nbatches = (S.tris + BATCH - 1) / BATCH;
    }
    // the draw-order slots the scans gave this triangle (it is live: the count pass checked
    // its indices before queueing it)
    uint32_t slot1 = 0;
    uint64_t pos0 = 0, pairs0 = 0, pos1 = 0, pairs1 = 0;
    if (EMIT) {
        const uint32_t nbt = (kb + 1 < nbatches) ? (uint32_t)BATCH : ntris - kb * (uint32_t)BATCH;
        slot1 = nbt + (a.app_ex[t].x - a.app_ex[b * BATCH].x);
        const uint2 g0 = a.grp_ex[(size_t)b * 2048 + j];
        const uint2 g1 = a.grp_ex[(size_t)b * 2048 + 1024 + j];
        pos0 = g0.x;
        pairs0 = g0.y;
        pos1 = g1.x;
        pairs1 = g1.y;
    }
    // lanes 0-2: the input vertices
    Pos4 P = {0.0f, 0.0f, 0.0f, 0.0f};
    float var[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k)
        var[k] = 0.0f;
    int cml = 0;
    if (lane < 3) {
        fetch_vertex<EMIT ? NV : 0>(a, a.indices[3ull * src + lane], P, var);
        cml = clip_mask(P);
    }
    const int cm = __builtin_amdgcn_readfirstlane(__shfl(cml, 0) | __shfl(cml, 1) | __shfl(cml, 2));
    // VertexProcessor.cpp:233-250: the planes of the mask in order, while 3 or more remain
    int n = 3;
    for (int p = 0; p < 6; ++p) {
        if (!(cm & (1 << p)) || n < 3)
            continue;
        const bool act = (int)lane < n;
        const int nx = (int)lane + 1 < n ? (int)lane + 1 : 0;
        const float dp = plane_dot(p, P);
        const float dn = __shfl(dp, nx);
        const bool vin = act && dp >= 0;
        const bool cross = act && sgnf(dn) != sgnf(dp);
        const unsigned long long m_in = __ballot(vin), m_cr = __ballot(cross);
        const unsigned long long below = (1ull << lane) - 1ull;
        const int m = __popcll(m_in) + __popcll(m_cr);
        if (m > MAX_CLIP) {
            if (lane == 0)
                atomicOr(&a.status[ST_ERROR], 2ull);
            n = 0;
            break;
        }
        const int slot = __popcll(m_in & below) + __popcll(m_cr & below);
        Pos4 Pn;
        Pn.x = __shfl(P.x, nx);
        Pn.y = __shfl(P.y, nx);
        Pn.z = __shfl(P.z, nx);
        Pn.w = __shfl(P.w, nx);
        float vn[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k)
            vn[k] = EMIT ? __shfl(var[k], nx) : 0.0f;
        lds_sync();  // the previous pass has read the stage (the stage is the wave's own)
        if (vin) {
            st.p[slot] = P;
#pragma unroll
            for (int k = 0; k < NV; ++k)
                if (EMIT)
                    st.v[slot][k] = var[k];
        }
        if (cross) {
            const float tt = dn < 0 ? dp / (dp - dn) : -dp / (dn - dp);
            Pos4 R;
            R.x = lerp1(P.x, Pn.x, tt);
            R.y = lerp1(P.y, Pn.y, tt);
            R.z = lerp1(P.z, Pn.z, tt);
            R.w = lerp1(P.w, Pn.w, tt);
            const int o = slot + (vin ? 1 : 0);
            st.p[o] = R;
            // Helper::interpolateVertex lerps the vertex shader's counts only
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const bool lerped = k < a.lay.na ? (k < a.vs_na) : ((k - a.lay.na) < a.vs_np);
                if (EMIT)
                    st.v[o][k] = lerped ? lerp1(var[k], vn[k], tt) : 0.0f;
            }
        }
        lds_sync();
        n = m;
        if ((int)lane < n) {
            P = st.p[lane];
#pragma unroll
            for (int k = 0; k < NV; ++k)
                if (EMIT)
                    var[k] = st.v[lane][k];
        }
    }
    // fan k = (p0, p[k+1], p[k+2]) on lane k
    const int nf = n >= 3 ? n - 2 : 0;
    const WinV wl = to_window(a, P);
    const int i1 = (int)lane + 1 < 64 ? (int)lane + 1 : 63;
    const int i2 = (int)lane + 2 < 64 ? (int)lane + 2 : 63;
    WinV w0 = shfl_win(wl, 0), w2 = shfl_win(wl, i2);
    const WinV w1 = shfl_win(wl, i1);
    float v0[NV], v1[NV], v2[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        v0[k] = EMIT ? __shfl(var[k], 0) : 0.0f;
        v1[k] = EMIT ? __shfl(var[k], i1) : 0.0f;
        v2[k] = EMIT ? __shfl(var[k], i2) : 0.0f;
    }
    bool swap = false, drawn = false;
    PrimGeo g;
    uint32_t nt = 0;
    if ((int)lane < nf) {
        drawn = fan_geometry(a, w0, w1, w2, true, g, swap);
        if (drawn)
            nt = (uint32_t)rect_tiles(g);
    }
    const unsigned long long m_drawn = __ballot(drawn);
    const uint32_t nt1 = lane >= 1 ? nt : 0u;  // pairs of the appended fans
    const uint32_t incl = wave_incl_sum(nt1, lane);
    if (!EMIT) {
        const uint32_t tot1 = __shfl(incl, 63);
        if (lane == 0) {
            a.grp[(size_t)b * 2048 + j] = make_uint2(drawn ? 1u : 0u, nt);
            a.grp[(size_t)b * 2048 + 1024 + j] =
                make_uint2((uint32_t)__popcll(m_drawn & ~1ull), tot1);
            a.app[t] = make_uint2(n > 3 ? (uint32_t)(n - 3) : 0u, 0u);
        }
        return;
    }
    if (!drawn)
        return;
    uint64_t pos = pos0, pb = pairs0;
    uint32_t key = (kb << 15) + j;
    if (lane >= 1) {
        pos = pos1 + (uint64_t)__popcll(m_drawn & ((1ull << lane) - 1ull) & ~1ull);
        pb = pairs1 + (incl - nt1);
        key = (kb << 15) + slot1 + (lane - 1);
    }
    fan_write<NV>(a, pos, pb, nt, key, src, sid, g, w0, w1, w2, v0, v1, v2, swap);
}

// The queued triangles, densely packed (order does not matter: each writes its own (batch,
// phase, j) counts and draw-order positions).  Wave `wid` of `nwaves` prefix-sums the region
// counts into its own s_off, then takes dense indices wid, wid + nwaves, ... and maps each
// to its (region, slot); s_off and st are the wave's own LDS.
template <int NV, bool EMIT>
__device__ void run_clip_queue(const GeomArgs &a, uint32_t wid, uint32_t nwaves,
                               uint32_t *s_off, ClipStage<NV> &st)
{
    static_assert(CLIPQ_REGIONS == CLIPQ_THREADS, "one region per lane");
    {
        // the 64 counters load in parallel, one per lane; an inclusive scan across the wave
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t cnt = a.clipq_cnt[lane * CLIPQ_STRIDE];
        const uint32_t incl = wave_incl_sum(cnt, lane);
        s_off[lane] = incl - cnt;
        if (lane == 63)
            s_off[CLIPQ_REGIONS] = incl;
    }
    lds_sync();
    const uint32_t n = s_off[CLIPQ_REGIONS], cap = clipq_region_cap(a.tris);
    for (uint32_t i = wid; i < n; i += nwaves) {
        uint32_t lo = 0, hi = CLIPQ_REGIONS;  // last region with s_off <= i
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (s_off[mid] <= i)
                lo = mid;
            else
                hi = mid;
        }
        clip_queued<NV, EMIT>(a, a.clipq[lo * cap + (i - s_off[lo])], st);
    }
}

__global__ void __launch_bounds__(CLIPQ_THREADS) k_clip_count(GeomArgs a)
{
    __shared__ uint32_t s_off[CLIPQ_REGIONS + 1];
    __shared__ ClipStage<1> st;
    run_clip_queue<1, false>(a, blockIdx.x, gridDim.x, s_off, st);
}

template <int NV>
__global__ void __launch_bounds__(CLIPQ_THREADS) k_clip_emit(GeomArgs a)
{
    if (a.status[ST_OVERFLOW] | a.status[ST_ERROR])
        return;
    __shared__ uint32_t s_off[CLIPQ_REGIONS + 1];
    __shared__ ClipStage<NV> st;
    run_clip_queue<NV, true>(a, blockIdx.x, gridDim.x, s_off, st);
}

// The emit pass over the live triangles with the clip queue's emit folded in: workgroups
// [0, nq) clip queued triangles a wave each (4 nq waves take dense queue indices), the ones
// after them emit the listed triangles as k_geom_emit_listed does.  The queue's waves are
// dispatched first, so its long per-triangle chains of dependent loads run while the listed
// workgroups fill the rest of the GPU, instead of in a launch of their own before them.
// The two parts write disjoint records and read nothing the other writes.
template <int NV, bool SCENES>
__global__ void __launch_bounds__(256) k_geom_emit_clip(GeomArgs a, uint32_t nq)
{
    static_assert(NV == 3 || NV == 5, "instantiated for the classes emit_kernels joins");
    if (a.status[ST_OVERFLOW] | a.status[ST_ERROR])
        return;
    if (blockIdx.x >= nq) {
        emit_listed_one<NV, SCENES>(a, (uint32_t)a.status[ST_LIVE],
                                    (blockIdx.x - nq) * 256u + threadIdx.x);
        return;
    }
    __shared__ uint32_t s_off[4][CLIPQ_REGIONS + 1];
    __shared__ ClipStage<NV> st[4];
    const uint32_t w = threadIdx.x >> 6;
    run_clip_queue<NV, true>(a, blockIdx.x * 4u + w, nq * 4u, s_off[w], st[w]);
}

// ---- host launchers -----------------------------------------------------------------------
// queue kernels: a workgroup (one wave) per queued triangle, looping beyond the grid; up to
// 4096 workgroups (16 waves per CU), so that a 1M-triangle draw whose queue holds up to 0.4%
// of its triangles gives each workgroup one
static uint32_t clip_blocks(uint64_t tris)
{
    const uint64_t b = tris / 64 + 64;
    return (uint32_t)(b < 4096 ? b : 4096);
}
hipError_t launch_begin(const GeomArgs &a, hipStream_t s)
{
    uint32_t work = 2 * a.tiles > a.batches * 1024u ? 2 * a.tiles : a.batches * 1024u;
    uint32_t blocks = (work + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_begin, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_geom_count(const GeomArgs &a, hipStream_t s)
{
    if (!a.tris)
        return hipSuccess;
    if (a.vwin && a.nverts)
        hipLaunchKernelGGL(k_vertex, dim3((a.nverts + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_geom_count, dim3((a.tris + 255) / 256), dim3(256), 0, s, a);
    if (!a.window_input)
        hipLaunchKernelGGL(k_clip_count, dim3(clip_blocks(a.tris)), dim3(CLIPQ_THREADS), 0, s, a);
    return hipGetLastError();
}

template <int CH>
static hipError_t launch_expand_ch(const GeomArgs &a, hipStream_t s)
{
    uint64_t blocks = (a.pair_cap + CH - 1) / CH;
    if (blocks < 1)
        blocks = 1;
    hipLaunchKernelGGL(k_expand<CH>, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_expand(const GeomArgs &a, hipStream_t s, int chunk)
{
    switch (chunk) {  // the chunk sizes swr_api.cpp's expand_chunk_of names
    case 256: return launch_expand_ch<256>(a, s);
    case 512: return launch_expand_ch<512>(a, s);
    case 2048: return launch_expand_ch<2048>(a, s);
    case 4096: return launch_expand_ch<4096>(a, s);
    default: return launch_expand_ch<1024>(a, s);
    }
}

int emit_class(int nvary)
{
    if (nvary <= 0) return 0;
    if (nvary <= 3) return 3;
    if (nvary <= 5) return 5;
    if (nvary <= 8) return 8;
    return 32;
}

// The clipped triangles' emit (k_clip_emit) and the others' write disjoint records, one after
// the other on the draw stream.  (The queue emit is short: a second stream beside the main
// emit measured slower, its fork and join costing more than the overlap saved.)
template <int NV>
static void emit_kernels(const GeomArgs &a, hipStream_t s, bool q)
{
    dim3 g((a.tris + 255) / 256), blk(256), gq(clip_blocks(a.tris)), bq(CLIPQ_THREADS);
    // the listed emit and the queue's emit in one launch, for the varying classes where the
    // joint kernel keeps the listed emit's registers (with 8 or 32 varyings it spills; without
    // varyings it holds 14 registers more, a wave per SIMD less)
    if (q && a.live && (NV == 3 || NV == 5)) {
        const uint32_t nq = (gq.x + 3) / 4;  // as many queue waves as the queue kernel's
        const dim3 gc(nq + g.x);
        constexpr int NJ = NV == 3 ? 3 : 5;  // only 3 and 5 reach here
        if (a.nscenes)
            hipLaunchKernelGGL((k_geom_emit_clip<NJ, true>), gc, blk, 0, s, a, nq);
        else
            hipLaunchKernelGGL((k_geom_emit_clip<NJ, false>), gc, blk, 0, s, a, nq);
        return;
    }
    if (q)
        hipLaunchKernelGGL(k_clip_emit<NV == 0 ? 1 : NV>, gq, bq, 0, s, a);
    if (a.live) {  // a thread per live triangle
        const dim3 gl(g.x);
        if (a.nscenes)
            hipLaunchKernelGGL((k_geom_emit_listed<NV, true>), gl, blk, 0, s, a);
        else
            hipLaunchKernelGGL((k_geom_emit_listed<NV, false>), gl, blk, 0, s, a);
    } else {
        hipLaunchKernelGGL(k_geom_emit<NV>, g, blk, 0, s, a);
    }
}

hipError_t launch_geom_emit(const GeomArgs &a, hipStream_t s)
{
    if (!a.tris)
        return hipSuccess;
    const bool q = !a.window_input;
    switch (emit_class(a.lay.na + a.lay.np)) {
    case 0: emit_kernels<0>(a, s, q); break;
    case 3: emit_kernels<3>(a, s, q); break;
    case 5: emit_kernels<5>(a, s, q); break;
    case 8: emit_kernels<8>(a, s, q); break;
    default: emit_kernels<32>(a, s, q); break;
    }
    return hipGetLastError();
}

}  // namespace swr
