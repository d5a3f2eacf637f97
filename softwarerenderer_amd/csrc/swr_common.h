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
//   32..   parameter planes a,b,c: [z][invw][avar 0..na)[pvar 0..np)
constexpr int REC_KEY = 0;
constexpr int REC_FLAGS = 1;
constexpr int REC_TILES = 2;
constexpr int REC_EDGE = 8;
constexpr int REC_BOT = 17;
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
constexpr int ST_WORDS = 8;

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
    int width, height, tiles_x, tiles_y;
    ParamLayout lay;
    int window_input;           // 1: vertices are window-space RasterizerVertex (IRasterizer seam)
    const swr_rasterizer_vertex *wverts;
    uint64_t wvert_count;
    // scan buffers
    uint32_t *app;              // [tris] appended fan triangles per input triangle
    const uint32_t *app_ex;     // exclusive scan of app
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
};

struct RasterArgs {
    const float *prims;
    int stride;
    const uint32_t *ranges;     // [tiles*2] begin, end into vals
    const uint32_t *vals;       // primitive index per sorted pair
    const unsigned long long *status;
    int tiles, tiles_x;
    int width, height, pitch;
    int sc_minx, sc_maxx;       // Span x clipping (Rasterizer.h:377-378, 404-405)
    int na_set, np_set;         // declared counts (hash / log programs)
    int iz, iw;
    void *t0, *t1, *t2;
    uint32_t *log;
    uint64_t log_cap;
    unsigned long long *log_count;
    float checker_cells;
};

}  // namespace swr
