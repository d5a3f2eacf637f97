// swr_programs.h — the built-in pixel programs (include/swr_hip.h, SWR_PS_*), shared by the
// triangle raster kernel and the line / point fragment kernel.
//
// A program keeps one pixel's target values in registers: load() reads them once, shade()
// runs the reference's drawPixel for one invocation, store() writes back what changed.
#pragma once

#include "swr_common.h"

#pragma clang fp contract(off)

namespace swr {

template <int NZ_, int NW_, int NA_, int NP_>
struct Layout {
    static constexpr int NZ = NZ_, NW = NW_, NA = NA_, NP = NP_;
    static constexpr int N = NZ + NW + NA + NP;
    static constexpr int Z = 0, IW = NZ, A = NZ + NW, P = NZ + NW + NA;
};

__device__ inline uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ inline uint32_t mix32(uint32_t h, uint32_t x) { return (h ^ x) * 16777619u; }
__device__ inline float clamp01(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

// ---- the pixel a program sees ----------------------------------------------------------
// PixelData (PixelData.h:38-146) as accessors.  A program reads z(), invw(), w(), avar(k)
// and pvar(k, w) with w = w() fetched once; each source of pixels supplies them its way.
//
// Triangle pixels: the chained planes v[] in Layout order (z, invw, avar..., pvarTemp...);
// w = 1/invw and pvar = pvarTemp * w at the pixel (PixelData.h:84-103).
template <class L>
struct ChainPixel {
    const float *v;
    __device__ float z() const { return v[L::Z]; }
    __device__ float invw() const { return v[L::IW]; }
    __device__ float w() const { return 1.0f / v[L::IW]; }
    __device__ float avar(int k) const { return v[L::A + k]; }
    __device__ float pvar(int k, float w) const { return v[L::P + k] * w; }
};

// ---- pixel programs (include/swr_hip.h) ---------------------------------------------------
template <class L>
struct ProgCoverage {
    static constexpr int KEY_WORD = REC_KEY;
    struct State { bool hit; };
    __device__ static void load(State &s, const RasterArgs &, size_t) { s.hit = false; }
    template <class PX>
    __device__ static void shade(State &s, const RasterArgs &, const PX &, uint32_t, int, int)
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
template <class L>
struct ProgRGB {
    static constexpr int KEY_WORD = REC_KEY;
    struct State { uint32_t c; bool hit; };
    __device__ static void load(State &s, const RasterArgs &, size_t) { s.hit = false; s.c = 0; }
    template <class PX>
    __device__ static void shade(State &s, const RasterArgs &, const PX &v, uint32_t, int, int)
    {
        uint32_t r = (uint32_t)f2i(v.avar(0) * 255.0f);
        uint32_t g = (uint32_t)f2i(v.avar(1) * 255.0f);
        uint32_t b = (uint32_t)f2i(v.avar(2) * 255.0f);
        s.c = (r << 16) | (g << 8) | b;
        s.hit = true;
    }
    __device__ static void store(const State &s, const RasterArgs &a, size_t o)
    {
        if (s.hit)
            static_cast<uint32_t *>(a.t0)[o] = s.c;
    }
};

template <class L>
struct ProgDepthChecker {
    static constexpr int KEY_WORD = REC_KEY;
    struct State { float depth; uint32_t color; bool hit; };
    __device__ static void load(State &s, const RasterArgs &a, size_t o)
    {
        s.depth = static_cast<const float *>(a.t1)[o];
        s.color = 0;
        s.hit = false;
    }
    template <class PX>
    __device__ static void shade(State &s, const RasterArgs &a, const PX &v, uint32_t, int, int)
    {
        const float z = v.z();
        if (z < s.depth) {
            s.depth = z;
            const float w = v.w();  // only the depth-passing pixels read pvar
            const float u = v.pvar(0, w), t = v.pvar(1, w);
            int cu = f2i(floorf(u * a.checker_cells));
            int cv = f2i(floorf(t * a.checker_cells));
            float k = (((uint32_t)cu + (uint32_t)cv) & 1u) ? 1.0f : 0.5f;
            uint32_t r = (uint32_t)(clamp01(v.avar(0) * k) * 255.0f);
            uint32_t g = (uint32_t)(clamp01(v.avar(1) * k) * 255.0f);
            uint32_t b = (uint32_t)(clamp01(v.avar(2) * k) * 255.0f);
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

template <class L>
struct ProgHash {
    static constexpr int KEY_WORD = REC_KEY;
    struct State { uint32_t h, n, n0; };
    __device__ static void load(State &s, const RasterArgs &a, size_t o)
    {
        s.h = static_cast<const uint32_t *>(a.t0)[o];
        s.n = s.n0 = static_cast<const uint32_t *>(a.t1)[o];
    }
    template <class PX>
    __device__ static void shade(State &s, const RasterArgs &a, const PX &v, uint32_t key, int,
                                 int)
    {
        const float w = v.w();
        uint32_t h = mix32(s.h, key);
        if (a.iz)
            h = mix32(h, fbits(v.z()));
        if (a.iw) {
            h = mix32(h, fbits(w));
            h = mix32(h, fbits(v.invw()));
        }
#pragma unroll
        for (int k = 0; k < L::NA; ++k)
            if (k < a.na_set)
                h = mix32(h, fbits(v.avar(k)));
#pragma unroll
        for (int k = 0; k < L::NP; ++k)
            if (k < a.np_set)
                h = mix32(h, fbits(v.pvar(k, w)));
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

template <class L>
struct ProgLog {
    static constexpr int KEY_WORD = REC_KEY;
    struct State { uint32_t n, n0; };
    __device__ static void load(State &s, const RasterArgs &a, size_t o)
    {
        s.n = s.n0 = static_cast<const uint32_t *>(a.t0)[o];
    }
    template <class PX>
    __device__ static void shade(State &s, const RasterArgs &a, const PX &v, uint32_t key, int x,
                                 int y)
    {
        const float w = v.w();
        uint32_t seq = s.n++;
        unsigned long long slot = atomicAdd(a.log_count, 1ull);
        if (slot < a.log_cap) {
            uint32_t *r = a.log + slot * SWR_LOG_WORDS;
            r[0] = (uint32_t)x;
            r[1] = (uint32_t)y;
            r[2] = key;
            r[3] = seq;
            r[4] = a.iz ? fbits(v.z()) : 0u;
            r[5] = a.iw ? fbits(w) : 0u;
            r[6] = a.iw ? fbits(v.invw()) : 0u;
#pragma unroll
            for (int k = 0; k < SWR_MAX_AVARS; ++k)
                r[7 + k] = (k < L::NA && k < a.na_set) ? fbits(v.avar(k < L::NA ? k : 0)) : 0u;
#pragma unroll
            for (int k = 0; k < SWR_MAX_PVARS; ++k)
                r[7 + SWR_MAX_AVARS + k] =
                    (k < L::NP && k < a.np_set) ? fbits(v.pvar(k < L::NP ? k : 0, w)) : 0u;
        }
    }
    __device__ static void store(const State &s, const RasterArgs &a, size_t o)
    {
        if (s.n != s.n0)
            static_cast<uint32_t *>(a.t0)[o] = s.n;
    }
};

// Visibility buffer: the nearest primitive's source triangle (+1, 0 = none) and its depth,
// under the depth-checker's `z < depth` test.
template <class L>
struct ProgVisibility {
    static constexpr int KEY_WORD = REC_SRC;
    struct State { float depth; uint32_t id; bool hit; };
    __device__ static void load(State &s, const RasterArgs &a, size_t o)
    {
        s.depth = static_cast<const float *>(a.t1)[o];
        s.id = 0;
        s.hit = false;
    }
    template <class PX>
    __device__ static void shade(State &s, const RasterArgs &, const PX &v, uint32_t src, int,
                                 int)
    {
        const float z = v.z();
        if (z < s.depth) {
            s.depth = z;
            s.id = src + 1u;
            s.hit = true;
        }
    }
    __device__ static void store(const State &s, const RasterArgs &a, size_t o)
    {
        if (s.hit) {
            static_cast<float *>(a.t1)[o] = s.depth;
            static_cast<uint32_t *>(a.t0)[o] = s.id;
        }
    }
};

}  // namespace swr
