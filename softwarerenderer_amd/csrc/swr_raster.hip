// swr_raster.hip — the raster stage: one wave64 per 8x8 tile, one lane per pixel.
//
// A tile is exactly one of the reference's 8x8 blocks (Rasterizer.h:247-250 aligns blocks
// to multiples of 8 in window space).  The wave walks the tile's primitives in draw order;
// each lane owns one pixel, so every pixel sees its pixel-shader invocations in the
// reference's order.  The pixel's target values (colour/depth, hash/count, ...) stay in
// registers for the whole walk and are written back once, coalesced.
//
// Interpolants are reproduced bit for bit by replaying the reference's incremental chains
// instead of evaluating plane equations directly:
//   Block (PixelShaderBase.h:55-94): value(i, j) = init(x+.5, y+.5), +b repeated j times
//     (PixelData::stepY per row), then +a repeated i times (stepX per pixel); edges alike;
//     the block is classified from its four corners with v + a*7 (Rasterizer.h:270-304).
//   Span (PixelShaderBase.h:96-112): value(x) = init(xl+.5, y+.5), +a repeated (x - xl) times,
//     with [xl, xr) from the flat-part edge walk (Rasterizer.h:370-405).
//   w = 1/invw and pvar = pvarTemp * w at the pixel (PixelData.h:84-103).
#include "swr_common.h"
#include "swr_internal.h"

namespace swr {

template <int NA, int NP>
struct Pix {
    float z, w, invw;
    float avar[NA > 0 ? NA : 1];
    float pvar[NP > 0 ? NP : 1];
};

__device__ inline uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ inline uint32_t mix32(uint32_t h, uint32_t x) { return (h ^ x) * 16777619u; }
__device__ inline float clamp01(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

// ---- pixel programs (include/swr_hip.h) ---------------------------------------------------
struct ProgCoverage {
    struct State { bool hit; };
    __device__ static void load(State &s, const RasterArgs &, size_t) { s.hit = false; }
    template <class P>
    __device__ static void shade(State &s, const RasterArgs &, const P &, uint32_t, int, int)
    {
        s.hit = true;
    }
    __device__ static void store(const State &s, const RasterArgs &a, size_t o)
    {
        if (s.hit)
            static_cast<uint32_t *>(a.t0)[o] = 1u;
    }
};

// RasterizerTest.cpp:38-48
struct ProgRGB {
    struct State { uint32_t c; bool hit; };
    __device__ static void load(State &s, const RasterArgs &, size_t) { s.hit = false; s.c = 0; }
    template <class P>
    __device__ static void shade(State &s, const RasterArgs &, const P &p, uint32_t, int, int)
    {
        uint32_t r = (uint32_t)f2i(p.avar[0] * 255.0f);
        uint32_t g = (uint32_t)f2i(p.avar[1] * 255.0f);
        uint32_t b = (uint32_t)f2i(p.avar[2] * 255.0f);
        s.c = (r << 16) | (g << 8) | b;
        s.hit = true;
    }
    __device__ static void store(const State &s, const RasterArgs &a, size_t o)
    {
        if (s.hit)
            static_cast<uint32_t *>(a.t0)[o] = s.c;
    }
};

struct ProgDepthChecker {
    struct State { float depth; uint32_t color; bool hit; };
    __device__ static void load(State &s, const RasterArgs &a, size_t o)
    {
        s.depth = static_cast<const float *>(a.t1)[o];
        s.color = 0;
        s.hit = false;
    }
    template <class P>
    __device__ static void shade(State &s, const RasterArgs &a, const P &p, uint32_t, int, int)
    {
        if (p.z < s.depth) {
            s.depth = p.z;
            int cu = f2i(floorf(p.pvar[0] * a.checker_cells));
            int cv = f2i(floorf(p.pvar[1] * a.checker_cells));
            float k = (((uint32_t)cu + (uint32_t)cv) & 1u) ? 1.0f : 0.5f;
            uint32_t r = (uint32_t)(clamp01(p.avar[0] * k) * 255.0f);
            uint32_t g = (uint32_t)(clamp01(p.avar[1] * k) * 255.0f);
            uint32_t b = (uint32_t)(clamp01(p.avar[2] * k) * 255.0f);
            s.color = r | (g << 8) | (b << 16) | 0xFF000000u;
            s.hit = true;
        }
    }
    __device__ static void store(const State &s, const RasterArgs &a, size_t o)
    {
        if (s.hit) {
            static_cast<float *>(a.t1)[o] = s.depth;
            static_cast<uint32_t *>(a.t0)[o] = s.color;
        }
    }
};

template <int NA, int NP>
struct ProgHash {
    struct State { uint32_t h, n, n0; };
    __device__ static void load(State &s, const RasterArgs &a, size_t o)
    {
        s.h = static_cast<const uint32_t *>(a.t0)[o];
        s.n = s.n0 = static_cast<const uint32_t *>(a.t1)[o];
    }
    template <class P>
    __device__ static void shade(State &s, const RasterArgs &a, const P &p, uint32_t key, int, int)
    {
        uint32_t h = mix32(s.h, key);
        if (a.iz)
            h = mix32(h, fbits(p.z));
        if (a.iw) {
            h = mix32(h, fbits(p.w));
            h = mix32(h, fbits(p.invw));
        }
#pragma unroll
        for (int k = 0; k < NA; ++k)
            if (k < a.na_set)
                h = mix32(h, fbits(p.avar[k]));
#pragma unroll
        for (int k = 0; k < NP; ++k)
            if (k < a.np_set)
                h = mix32(h, fbits(p.pvar[k]));
        s.h = h;
        s.n += 1;
    }
    __device__ static void store(const State &s, const RasterArgs &a, size_t o)
    {
        if (s.n != s.n0) {
            static_cast<uint32_t *>(a.t0)[o] = s.h;
            static_cast<uint32_t *>(a.t1)[o] = s.n;
        }
    }
};

template <int NA, int NP>
struct ProgLog {
    struct State { uint32_t n, n0; };
    __device__ static void load(State &s, const RasterArgs &a, size_t o)
    {
        s.n = s.n0 = static_cast<const uint32_t *>(a.t0)[o];
    }
    template <class P>
    __device__ static void shade(State &s, const RasterArgs &a, const P &p, uint32_t key, int x,
                                 int y)
    {
        uint32_t seq = s.n++;
        unsigned long long slot = atomicAdd(a.log_count, 1ull);
        if (slot < a.log_cap) {
            uint32_t *r = a.log + slot * SWR_LOG_WORDS;
            r[0] = (uint32_t)x;
            r[1] = (uint32_t)y;
            r[2] = key;
            r[3] = seq;
            r[4] = a.iz ? fbits(p.z) : 0u;
            r[5] = a.iw ? fbits(p.w) : 0u;
            r[6] = a.iw ? fbits(p.invw) : 0u;
            for (int k = 0; k < SWR_MAX_AVARS; ++k)
                r[7 + k] = (k < NA && k < a.na_set) ? fbits(p.avar[k < NA ? k : 0]) : 0u;
            for (int k = 0; k < SWR_MAX_PVARS; ++k)
                r[7 + SWR_MAX_AVARS + k] = (k < NP && k < a.np_set) ? fbits(p.pvar[k < NP ? k : 0]) : 0u;
        }
    }
    __device__ static void store(const State &s, const RasterArgs &a, size_t o)
    {
        if (s.n != s.n0)
            static_cast<uint32_t *>(a.t0)[o] = s.n;
    }
};

// ---- the tile kernel ----------------------------------------------------------------------
__device__ inline bool edge_pass(float v, bool tie) { return v > 0.0f || (v == 0.0f && tie); }

template <class Prog, int NZ, int NW, int NA, int NP>
__global__ void __launch_bounds__(256) k_raster(RasterArgs a)
{
    constexpr int N = NZ + NW + NA + NP;
    constexpr int NS = N > 0 ? N : 1;
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= a.tiles)
        return;
    if (a.status[ST_OVERFLOW] | a.status[ST_ERROR])
        return;
    const uint32_t beg = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
    if (beg >= end)
        return;
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int bx = tx * TILE, by = ty * TILE;
    const int li = lane & 7, lj = lane >> 3;
    const int px = bx + li, py = by + lj;
    const bool inside = px < a.width && py < a.height;
    const size_t o = (size_t)py * (size_t)a.pitch + (size_t)px;
    typename Prog::State st;
    if (inside)
        Prog::load(st, a, o);
    const float xf = (float)bx + 0.5f, yf = (float)by + 0.5f;
    const float pyf = (float)py;

    for (uint32_t e = beg; e < end; ++e) {
        const uint32_t pid = __builtin_amdgcn_readfirstlane(a.vals[e]);
        const float *__restrict__ rec = a.prims + (size_t)pid * (size_t)a.stride;
        const uint32_t flags = __float_as_uint(rec[REC_FLAGS]);
        const uint32_t key = __float_as_uint(rec[REC_KEY]);
        float val[NS];
        bool cov;
        if (flags & FLAG_BLOCK) {
            const float e0a = rec[REC_EDGE + 0], e0b = rec[REC_EDGE + 1], e0c = rec[REC_EDGE + 2];
            const float e1a = rec[REC_EDGE + 3], e1b = rec[REC_EDGE + 4], e1c = rec[REC_EDGE + 5];
            const float e2a = rec[REC_EDGE + 6], e2b = rec[REC_EDGE + 7], e2c = rec[REC_EDGE + 8];
            const bool t0 = flags & FLAG_TIE0, t1 = flags & FLAG_TIE1, t2 = flags & FLAG_TIE2;
            // corner classification, Rasterizer.h:270-304
            const float s7 = 7.0f;
            float c00 = e0a * xf + e0b * yf + e0c, d00 = e1a * xf + e1b * yf + e1c,
                  f00 = e2a * xf + e2b * yf + e2c;
            float c01 = c00 + e0b * s7, d01 = d00 + e1b * s7, f01 = f00 + e2b * s7;
            float c10 = c00 + e0a * s7, d10 = d00 + e1a * s7, f10 = f00 + e2a * s7;
            float c11 = c01 + e0a * s7, d11 = d01 + e1a * s7, f11 = f01 + e2a * s7;
            bool p00[3] = {edge_pass(c00, t0), edge_pass(d00, t1), edge_pass(f00, t2)};
            bool p01[3] = {edge_pass(c01, t0), edge_pass(d01, t1), edge_pass(f01, t2)};
            bool p10[3] = {edge_pass(c10, t0), edge_pass(d10, t1), edge_pass(f10, t2)};
            bool p11[3] = {edge_pass(c11, t0), edge_pass(d11, t1), edge_pass(f11, t2)};
            int result = (p00[0] && p00[1] && p00[2]) + (p01[0] && p01[1] && p01[2]) +
                         (p10[0] && p10[1] && p10[2]) + (p11[0] && p11[1] && p11[2]);
            bool test_edges = true;
            if (result == 0) {
                // `a == b == c` is (a == b) == c: true iff an odd number of edges pass
                bool same = ((p00[0] == p00[1]) == p00[2]) && ((p01[0] == p01[1]) == p01[2]) &&
                            ((p10[0] == p10[1]) == p10[2]) && ((p11[0] == p11[1]) == p11[2]);
                if (same)
                    continue;
            } else if (result == 4) {
                test_edges = false;
            }
#pragma unroll
            for (int p = 0; p < N; ++p) {
                const float pa = rec[REC_PARAM + 3 * p], pb = rec[REC_PARAM + 3 * p + 1],
                            pc = rec[REC_PARAM + 3 * p + 2];
                float v = pa * xf + pb * yf + pc;
                for (int k = 0; k < lj; ++k)
                    v = v + pb;
                for (int k = 0; k < li; ++k)
                    v = v + pa;
                val[p] = v;
            }
            cov = true;
            if (test_edges) {
                float v0 = c00, v1 = d00, v2 = f00;
                for (int k = 0; k < lj; ++k) {
                    v0 = v0 + e0b;
                    v1 = v1 + e1b;
                    v2 = v2 + e2b;
                }
                for (int k = 0; k < li; ++k) {
                    v0 = v0 + e0a;
                    v1 = v1 + e1a;
                    v2 = v2 + e2a;
                }
                cov = edge_pass(v0, t0) && edge_pass(v1, t1) && edge_pass(v2, t2);
            }
        } else {
            // Span: which flat part owns this scanline, and its [xl, xr)
            const int blo = (int)__float_as_uint(rec[REC_BOT + 4]);
            const int bhi = (int)__float_as_uint(rec[REC_BOT + 5]);
            const int tlo = (int)__float_as_uint(rec[REC_TOP + 4]);
            const int thi = (int)__float_as_uint(rec[REC_TOP + 5]);
            const bool inb = (flags & FLAG_BOTTOM) && blo <= py && py < bhi;
            const bool intop = (flags & FLAG_TOP) && tlo < py && py <= thi;
            const int base = inb ? REC_BOT : REC_TOP;
            const float ox = inb ? rec[REC_BOT + 0] : rec[REC_TOP + 0];
            const float oy = inb ? rec[REC_BOT + 1] : rec[REC_TOP + 1];
            const float is1 = inb ? rec[REC_BOT + 2] : rec[REC_TOP + 2];
            const float is2 = inb ? rec[REC_BOT + 3] : rec[REC_TOP + 3];
            (void)base;
            float dy = (pyf - oy) + 0.5f;
            float cx1 = ox + is1 * dy + 0.5f;
            float cx2 = ox + is2 * dy + 0.5f;
            int xl = imax(a.sc_minx, f2i(cx1));
            int xr = imin(a.sc_maxx, f2i(cx2));
            cov = (inb || intop) && xl <= px && px < xr;
            const int steps = cov ? px - xl : 0;
            const float xs = (float)xl + 0.5f, ys = pyf + 0.5f;
#pragma unroll
            for (int p = 0; p < N; ++p) {
                const float pa = rec[REC_PARAM + 3 * p], pb = rec[REC_PARAM + 3 * p + 1],
                            pc = rec[REC_PARAM + 3 * p + 2];
                float v = pa * xs + pb * ys + pc;
                for (int k = 0; k < steps; ++k)
                    v = v + pa;
                val[p] = v;
            }
        }
        if (cov && inside) {
            Pix<NA, NP> pix;
            int q = 0;
            pix.z = NZ ? val[q] : 0.0f;
            q += NZ;
            if (NW) {
                pix.invw = val[q];
                pix.w = 1.0f / pix.invw;
            } else {
                pix.invw = pix.w = 0.0f;
            }
            q += NW;
#pragma unroll
            for (int k = 0; k < NA; ++k)
                pix.avar[k] = val[q + k];
#pragma unroll
            for (int k = 0; k < NP; ++k)
                pix.pvar[k] = val[q + NA + k] * pix.w;
            Prog::shade(st, a, pix, key, px, py);
        }
    }
    if (inside)
        Prog::store(st, a, o);
}

template <class Prog, int NZ, int NW, int NA, int NP>
static hipError_t launch(const RasterArgs &a, hipStream_t s)
{
    uint32_t blocks = (uint32_t)((a.tiles + 3) / 4);
    if (!blocks)
        return hipSuccess;
    hipLaunchKernelGGL((k_raster<Prog, NZ, NW, NA, NP>), dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

// The plane layout each program's kernel expects (see ParamLayout).
bool raster_layout(const swr_pixel_shader_desc &ps, ParamLayout &L)
{
    L.nz = L.nw = L.na = L.np = 0;
    L.set_na = L.set_np = 0;
    switch (ps.program) {
    case SWR_PS_NULL:
    case SWR_PS_COVERAGE:
        break;
    case SWR_PS_RGB:
        if (ps.avar_count < 3)
            return false;
        L.na = L.set_na = 3;
        break;
    case SWR_PS_DEPTH_CHECKER:
        if (!ps.interpolate_z || ps.avar_count < 3 || ps.pvar_count < 2)
            return false;
        L.nz = 1;
        L.nw = 1;
        L.na = L.set_na = 3;
        L.np = L.set_np = 2;
        break;
    case SWR_PS_HASH:
    case SWR_PS_LOG: {
        if (ps.avar_count < 0 || ps.avar_count > 16 || ps.pvar_count < 0 || ps.pvar_count > 16)
            return false;
        int cls = (ps.avar_count <= 4 && ps.pvar_count <= 4) ? 4 : 16;
        L.nz = 1;
        L.nw = 1;
        L.na = L.np = cls;
        L.set_na = ps.avar_count;
        L.set_np = ps.pvar_count;
        break;
    }
    default:
        return false;
    }
    L.count = L.nz + L.nw + L.na + L.np;
    L.stride = (REC_PARAM + 3 * L.count + 3) & ~3;
    return true;
}

hipError_t launch_raster(const swr_pixel_shader_desc &ps, const ParamLayout &L,
                         const RasterArgs &a, hipStream_t s)
{
    switch (ps.program) {
    case SWR_PS_COVERAGE:
        return launch<ProgCoverage, 0, 0, 0, 0>(a, s);
    case SWR_PS_RGB:
        return launch<ProgRGB, 0, 0, 3, 0>(a, s);
    case SWR_PS_DEPTH_CHECKER:
        return launch<ProgDepthChecker, 1, 1, 3, 2>(a, s);
    case SWR_PS_HASH:
        return L.na == 4 ? launch<ProgHash<4, 4>, 1, 1, 4, 4>(a, s)
                         : launch<ProgHash<16, 16>, 1, 1, 16, 16>(a, s);
    case SWR_PS_LOG:
        return L.na == 4 ? launch<ProgLog<4, 4>, 1, 1, 4, 4>(a, s)
                         : launch<ProgLog<16, 16>, 1, 1, 16, 16>(a, s);
    default:
        return hipSuccess;
    }
}

}  // namespace swr
