// swr_common.h — definitions shared by the gfx950 kernels and the host orchestration.
//
// Float arithmetic in every kernel follows the reference's evaluation order operation by
// operation (one rounding each, no contraction: the library is built with
// -ffp-contract=off and the pragma below), so the device reproduces the reference's bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits.h>

#include "../../include/swr_hip.h"

#pragma clang fp contract(off)

namespace swr {

// ---- primitive record: one per primitive that reaches the rasterizer, in draw order ----
// words (u32 / f32):
//   0 key   1 flags   2..5 tile rect tx0 ty0 tx1 ty1 (inclusive)
//   8..16  edge equations e0.abc e1.abc e2.abc (EdgeEquation.h:37-43)
//   17..22 bottom-flat part: ox oy is1 is2 row_lo row_hi   (rows lo <= y < hi)
//   23..28 top-flat part:    ox oy is1 is2 row_lo row_hi   (rows lo < y <= hi)
//   17..20 Block primitives instead: the 8x8 blocks of the reference's block grid,
//          bx0 by0 bx1 by1 (inclusive, in blocks; clamped to the target)
//   32..   parameter planes a,b,c: [z][invw][avar 0..na)[pvar 0..np)
constexpr int REC_KEY = 0;
constexpr int REC_FLAGS = 1;
constexpr int REC_TILES = 2;
constexpr int REC_EDGE = 8;
constexpr int REC_BOT = 17;
constexpr int REC_BLOCKS = 17;
constexpr int REC_TOP = 23;
constexpr int REC_PARAM = 32;

constexpr uint32_t FLAG_BLOCK = 1u;
constexpr uint32_t FLAG_TIE0 = 2u;
constexpr uint32_t FLAG_TIE1 = 4u;
constexpr uint32_t FLAG_TIE2 = 8u;
constexpr uint32_t FLAG_BOTTOM = 16u;
constexpr uint32_t FLAG_TOP = 32u;

constexpr int TILE = SWR_BLOCK_SIZE;   // one 8x8 tile = one reference block = one wave64
constexpr int BATCH = SWR_BATCH_PRIMITIVES;
constexpr int MAX_CLIP = SWR_MAX_CLIP_VERTS;

// status word indices (device u64 array, copied to the host after a draw)
constexpr int ST_PRIMS = 0;        // primitives emitted
constexpr int ST_PAIRS = 1;        // (primitive, tile) pairs emitted
constexpr int ST_OVERFLOW = 2;     // bit 0: prim arena, bit 1: pair arena
constexpr int ST_ERROR = 3;        // bit 0: index range, bit 1: clip overflow
constexpr int ST_BIG = 4;          // primitives queued for cooperative pair expansion
constexpr int ST_KEPT = 5;         // pairs left after dropping rejected blocks / tiles
constexpr int ST_CLIPQ = 6;        // triangles queued for the clipper
constexpr int ST_LIVE = 8;         // triangles the emit pass walks (the live list)
constexpr int ST_STICKY = 7;       // failures of asynchronous draws not yet reported: bit 0
                                   // overflow, bits 1.. the ST_ERROR bits; never cleared by
                                   // a draw, only by the host once it has reported them
constexpr int ST_WORDS = 16;       // a draw clears every word but ST_STICKY

// Counters a triangle draw's begin kernel clears, in one small buffer: the clip queue's 64
// regions (16 words apart), then the persistent raster's RASTER_QUEUES bin queues (32 words,
// 128 bytes, apart: one cache line each), RASTER_QUEUES / 8 per XCD.
constexpr uint32_t CLIPQ_COUNTERS = 64 * 16;
constexpr uint32_t RASTER_QUEUES = 64;
constexpr uint32_t RASTER_QUEUE_STRIDE = 32;
constexpr uint32_t DRAW_COUNTER_WORDS = CLIPQ_COUNTERS + RASTER_QUEUES * RASTER_QUEUE_STRIDE;

// record word 6: index of the primitive's first (tile, primitive) pair
constexpr int REC_PAIRBASE = 6;
// record word 7: the input triangle the primitive comes from (scene-local index)
constexpr int REC_SRC = 7;
// record word 29: the scene of a batched draw (0 otherwise)
constexpr int REC_SCENE = 29;

// Which plane equations the rasterizer interpolates (chosen by the pixel program).
struct ParamLayout {
    int nz;        // 1: z plane
    int nw;        // 1: invw plane (InterpolateW or PVarCount > 0)
    int na, np;    // avar / pvar planes stored (program class)
    int set_na, set_np;  // planes actually set up (declared counts, <= na / np)
    int count;     // nz + nw + na + np
    int stride;    // record stride in words
};

// ---- float -> int as x86-64 cvttss2si (out of range or NaN -> INT_MIN) ----------------
__host__ __device__ inline int f2i(float f)
{
    return (f > -2147483904.0f && f < 2147483648.0f) ? (int)f : INT_MIN;
}

// std::min / std::max argument order (matters for NaN)
__host__ __device__ inline float ref_min(float a, float b) { return (b < a) ? b : a; }
__host__ __device__ inline float ref_max(float a, float b) { return (a < b) ? b : a; }
__host__ __device__ inline int imin(int a, int b) { return (b < a) ? b : a; }
__host__ __device__ inline int imax(int a, int b) { return (a < b) ? b : a; }

// Orders one wave's LDS traffic: LDS operations of a wave complete in issue order, so
// keeping the compiler from moving them across this point and waiting for the wave's own
// LDS traffic makes its earlier LDS writes visible to its later reads (a barrier for data
// that only the calling wave shares).
__device__ inline void lds_sync()
{
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __asm__ volatile("" ::: "memory");
}

// EdgeEquation::test (EdgeEquation.h:53-61): v > 0 || (v == 0 && tie), as one compare:
// v >= 0 with the tie, v >= the smallest denormal without it (the kernels keep fp32
// denormals, so nothing lies strictly between 0 and it; NaN fails both forms).
__device__ inline bool edge_pass(float v, bool tie)
{
    return v >= (tie ? 0.0f : __uint_as_float(1u));
}

// Block classification from the four corners, Rasterizer.h:270-304: 0 skip,
// 1 drawBlock<true> (per-pixel edge tests), 2 drawBlock<false> (all 64 pixels).
// e[9] = e0.abc e1.abc e2.abc; corners use v + a*7 / v + b*7 (EdgeData::stepX/Y(eqn, s)).
__device__ inline int classify_block(const float *e, uint32_t flags, int x, int y)
{
    const float xf = (float)x + 0.5f, yf = (float)y + 0.5f, s = 7.0f;
    bool p[4][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float a = e[3 * k], b = e[3 * k + 1], c = e[3 * k + 2];
        const bool tie = (flags >> (1 + k)) & 1u;
        float v00 = a * xf + b * yf + c;
        float v01 = v00 + b * s;
        float v10 = v00 + a * s;
        float v11 = v01 + a * s;
        p[0][k] = edge_pass(v00, tie);
        p[1][k] = edge_pass(v01, tie);
        p[2][k] = edge_pass(v10, tie);
        p[3][k] = edge_pass(v11, tie);
    }
    int result = 0;
    bool same = true;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        result += (p[q][0] && p[q][1] && p[q][2]) ? 1 : 0;
        // `a == b == c` parses as (a == b) == c: true iff an odd number of edges pass
        same = same && ((p[q][0] == p[q][1]) == p[q][2]);
    }
    if (result == 0)
        return same ? 0 : 1;
    return result == 4 ? 2 : 1;
}

// True unless some edge is negative at every pixel centre of the block by more than the
// rounding the per-pixel chain can accumulate, i.e. unless no pixel of drawBlock<true>
// (PixelShaderBase.h:55-94) can pass all three edge tests.  The chained value of pixel
// (i, j) is init = (a*xf + b*yf) + c followed by j additions of b and i of a: at most 18
// roundings, each at most 2^-24 of a partial sum bounded by M = |a xf| + |b yf| + |c| +
// 7|a| + 7|b|, so below 2^-19.8 M in all.  The edge's maximum over the block (a corner: the
// function is linear) is evaluated in fp32, with at most 9 roundings of terms bounded by M
// (below 2^-20.8 M), and the block is dropped only when that evaluation is below -2^-18 M:
// the exact maximum is then below -2^-18 M + 2^-20.8 M, more negative than any chain's
// rounding can lift.  A coefficient or bound that is not finite keeps the block.  ext = 7 is
// one 8x8 block; a larger ext tests a square of (ext + 1)^2 pixels (a bin of blocks: the
// bound then covers every block's chain, which starts inside the square).
__device__ inline bool block_may_cover(const float *e, int x, int y, int ext = 7)
{
    const float xf = (float)x + 0.5f, yf = (float)y + 0.5f, s = (float)ext;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float a = e[3 * k], b = e[3 * k + 1], c = e[3 * k + 2];
        const float m = fabsf(a * xf) + fabsf(b * yf) + fabsf(c) + s * (fabsf(a) + fabsf(b));
        const float mx = a * xf + b * yf + c + fmaxf(0.0f, s * a) + fmaxf(0.0f, s * b);
        if (mx < -m * 3.814697265625e-06f)  // 2^-18; false for NaN and infinite bounds
            return false;
    }
    return true;
}

// Span: the flat part owning scanline y and its pixel range [xl, xr)
// (Rasterizer.h:370-380 bottom-flat, :397-407 top-flat).  Returns false when no part owns y.
__device__ inline bool span_row_extent(const float *rec, uint32_t flags, int y, int sc_minx,
                                       int sc_maxx, int &xl, int &xr)
{
    const int blo = (int)__float_as_uint(rec[REC_BOT + 4]);
    const int bhi = (int)__float_as_uint(rec[REC_BOT + 5]);
    const int tlo = (int)__float_as_uint(rec[REC_TOP + 4]);
    const int thi = (int)__float_as_uint(rec[REC_TOP + 5]);
    const bool inb = (flags & FLAG_BOTTOM) && blo <= y && y < bhi;
    const bool intop = (flags & FLAG_TOP) && tlo < y && y <= thi;
    const int base = inb ? REC_BOT : REC_TOP;
    const float ox = rec[base + 0], oy = rec[base + 1];
    const float is1 = rec[base + 2], is2 = rec[base + 3];
    const float dy = ((float)y - oy) + 0.5f;
    const float cx1 = ox + is1 * dy + 0.5f;
    const float cx2 = ox + is2 * dy + 0.5f;
    xl = imax(sc_minx, f2i(cx1));
    xr = imin(sc_maxx, f2i(cx2));
    return inb || intop;
}

// Span: true when some scanline of the 8x8 block at (bx, by) has pixels of the primitive
// inside the block's columns (the ranges the rasterizer computes, Rasterizer.h:370-380,
// 397-407).
__device__ inline bool span_block_any(const float *rec, uint32_t flags, int bx, int by,
                                      int sc_minx, int sc_maxx)
{
    bool any = false;
    for (int r = 0; r < TILE && !any; ++r) {
        int xl, xr;
        const bool rowv = span_row_extent(rec, flags, by + r, sc_minx, sc_maxx, xl, xr);
        any = rowv && imax(xl, bx) < imin(xr, bx + TILE);
    }
    return any;
}

// pair value: primitive index | classification << 30 (Block primitives)
constexpr uint32_t PAIR_PRIM_MASK = 0x3FFFFFFFu;
// 16x16 bins whose pair values carry the classes of the bin's four 8x8 blocks: primitive
// index (24 bits) | class of block b << (24 + 2 b), blocks in row-major order (Block: the
// corner classification of Rasterizer.h:270-304 after block_may_cover; Span: 1 when some
// scanline of the block has pixels of the primitive)
constexpr int PAIR_CLS4_SHIFT = 24;
constexpr uint32_t PAIR_CLS4_PRIM_MASK = 0x00FFFFFFu;

// ---- batched draws ------------------------------------------------------------------
// One scene of a batched draw: its inputs, its VertexProcessor / Rasterizer state and its
// render targets.  Scene s owns the virtual triangles [batch0 * 1024, batch0 * 1024 + tris)
// (each scene starts on a fresh 1024-triangle batch, so its draw order is exactly that of
// a draw of its own) and, for each bin size 8 << L (L = 0, 1, 2), the global bins
// [bin0[L], bin0[L] + bins_x[L] * bins_y[L]).
struct SceneDev {
    const int32_t *indices;
    const uint8_t *attr0;
    uint64_t attr0_bytes;
    void *t0, *t1, *t2;
    float mvp[16];
    float vp_px, vp_py, vp_ox, vp_oy, dz_scale, dz_bias;
    int stride0, cull_mode, raster_mode;
    int sc_minx, sc_maxx, sc_miny, sc_maxy;
    int width, height, pitch;
    int bins_x[3], bins_y[3];
    uint32_t bin0[3];
    uint32_t tris, batch0;
};

// ---- launch arguments ---------------------------------------------------------------
struct GeomArgs {
    // input
    const int32_t *indices;
    const uint8_t *attr0;
    int stride0;
    uint64_t attr0_bytes;       // 0: unknown
    uint64_t attr0_need;        // bytes the vertex program reads per vertex
    float mvp[16];
    int vs_program, vs_na, vs_np;
    uint32_t tris;              // input triangles (complete triples)
    uint32_t batches;
    // state
    int raster_mode, cull_mode;
    int sc_minx, sc_maxx, sc_miny, sc_maxy;   // scissor as Rasterizer.h:83-86 stores it
    float vp_px, vp_py, vp_ox, vp_oy;         // VertexProcessor.cpp:50-53
    float dz_scale, dz_bias;                  // 0.5f*(f-n), 0.5f*(n+f)
    int width, height, tiles_x, tiles_y;  // tiles_x, tiles_y, tiles: bins of 1 << bin_shift px
    int bin_shift;              // 3, 4 or 5: screen bins of 8, 16 or 32 pixels
    int cls4;                   // 16x16 bins: pair values carry the four block classes
    ParamLayout lay;
    int window_input;           // 1: vertices are window-space RasterizerVertex (IRasterizer seam)
    const swr_rasterizer_vertex *wverts;
    uint64_t wvert_count;
    // scan buffers
    uint2 *app;                 // [tris] (fan triangles appended, live) per input triangle
    const uint2 *app_ex;        // exclusive scan of app
    uint32_t *live;             // [tris] the live triangles in order (nullptr: emit walks all)
    uint2 *grp;                 // [batches*2048] (prims, pairs) per (batch, phase, j)
    const uint2 *grp_ex;        // exclusive scan of grp
    // outputs
    float *prims;               // records
    uint64_t prim_cap;
    uint32_t *pair_keys;        // tile id
    uint32_t *pair_vals;        // primitive index
    uint64_t pair_cap;
    unsigned long long *status;
    uint32_t *ranges;           // [tiles*2] cleared by the count pass
    uint32_t tiles;
    uint32_t sentinel;          // pair key of a tile the primitive cannot shade (dropped)
    uint32_t *prim_pb;          // [prim_cap] index of each primitive's first pair
    uint32_t *chunk_kept;       // [pair_cap / expansion chunk] kept pairs per chunk (or null)
    uint32_t *clipq;            // [tris + 64*256] triangles that straddle a clip plane
    uint32_t *clipq_cnt;        // [DRAW_COUNTER_WORDS]: clip queue regions, raster queues
    // post-transform vertices (the vertex pass; vwin == nullptr: off): window-space
    // (x, y, z, clip w) and the clip mask of each vertex [0, nverts) of attribute 0
    float4 *vwin;
    uint8_t *vmask;
    uint32_t nverts;
    // batched draw: per-scene state replaces the fields above (nscenes == 0: single draw)
    const SceneDev *scenes;
    const uint32_t *batch_scene;  // [batches] scene of each virtual batch
    uint32_t nscenes;
};

// A texture and its mip chain in device memory (swr_texture_create): level k is
// w[k] x h[k] texels 0x00RRGGBB at texels + off[k].
constexpr int TEX_MAX_LEVELS = 32;
struct TexDev {
    const uint32_t *texels;
    int levels, max_aniso;
    int w[TEX_MAX_LEVELS], h[TEX_MAX_LEVELS];
    uint64_t off[TEX_MAX_LEVELS];
};

struct RasterArgs {
    const float *prims;
    int stride;
    const uint32_t *ranges;     // [tiles*2] begin, end into vals
    const uint32_t *vals;       // primitive index per sorted pair
    const unsigned long long *status;
    int tiles, tiles_x;         // bins of 1 << bin_shift pixels
    int bin_shift;
    int cls4;                   // 16x16 bins: pair values carry the four block classes
    int waves_per_bin;          // bins larger than 8x8: waves sharing one bin's blocks
    int width, height, pitch;
    int sc_minx, sc_maxx;       // Span x clipping (Rasterizer.h:377-378, 404-405)
    int na_set, np_set;         // declared counts (hash / log programs)
    int iz, iw;
    void *t0, *t1, *t2;
    uint32_t *log;
    uint64_t log_cap;
    unsigned long long *log_count;
    float checker_cells;
    const SceneDev *scenes;     // batched draw (nscenes == 0: single draw)
    uint32_t nscenes;
    int xcd;                    // 1: tile blocks grouped per XCD
    const TexDev *tex;          // SWR_PS_TEXTURE
    int fold;                   // 1: an asynchronous draw; its failures go to ST_STICKY
    const void *user;           // SWR_PS_MODULE: the program's uniforms (device memory)
    int spans;                  // 0: every primitive is a Block primitive (raster mode Block)
    uint32_t *queue;            // persistent raster: RASTER_QUEUES bin-queue heads, zeroed by the
                                // draw's begin kernel (nullptr: one workgroup per bin)
};

// An asynchronous draw's failures, kept for the next swr_synchronize: once every earlier
// kernel of the draw has finished (stream order), its overflow and error words are final.
__device__ inline void fold_status(unsigned long long *st)
{
    const unsigned long long f = (st[ST_OVERFLOW] ? 1ull : 0ull) | (st[ST_ERROR] << 1);
    if (f)
        atomicOr(&st[ST_STICKY], f);
}

// Scene-local state of a batched draw over the single-draw arguments.
__device__ inline void apply_scene(GeomArgs &a, const SceneDev &s)
{
    a.indices = s.indices;
    a.attr0 = s.attr0;
    a.attr0_bytes = s.attr0_bytes;
    a.stride0 = s.stride0;
#pragma unroll
    for (int i = 0; i < 16; ++i)
        a.mvp[i] = s.mvp[i];
    a.vp_px = s.vp_px;
    a.vp_py = s.vp_py;
    a.vp_ox = s.vp_ox;
    a.vp_oy = s.vp_oy;
    a.dz_scale = s.dz_scale;
    a.dz_bias = s.dz_bias;
    a.cull_mode = s.cull_mode;
    a.raster_mode = s.raster_mode;
    a.sc_minx = s.sc_minx;
    a.sc_maxx = s.sc_maxx;
    a.sc_miny = s.sc_miny;
    a.sc_maxy = s.sc_maxy;
    a.width = s.width;
    a.height = s.height;
    a.tiles_x = s.bins_x[a.bin_shift - 3];
    a.tiles_y = s.bins_y[a.bin_shift - 3];
}

// The scene owning global bin `tile` of bin level lvl (scenes sorted by bin0[lvl]).
__device__ inline uint32_t scene_of_tile(const SceneDev *sc, uint32_t n, uint32_t tile, int lvl)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (sc[mid].bin0[lvl] <= tile)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

}  // namespace swr
