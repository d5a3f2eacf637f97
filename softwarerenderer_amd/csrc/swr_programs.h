// swr_programs.h — the built-in pixel programs (include/swr_hip.h, SWR_PS_*), shared by the
// triangle raster kernel and the line / point fragment kernel.
//
// A program keeps one pixel's target values in registers: load() reads them once, shade()
// runs the reference's drawPixel for one invocation, store() writes back what changed.
#pragma once

#include "swr_chain.h"
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
template <class L, bool SP = true>
struct ChainPixel {
    // false: every primitive of the draw is a Block primitive (no Span chains to replay)
    static constexpr bool SPANS = SP;
    const float *v;
    uint32_t prim;     // the primitive's record index (its planes), for programs with PLANES set
    __device__ float z() const { return v[L::Z]; }
    __device__ float invw() const { return v[L::IW]; }
    __device__ float w() const { return 1.0f / v[L::IW]; }
    __device__ float avar(int k) const { return v[L::A + k]; }
    __device__ float pvar(int k, float w) const { return v[L::P + k] * w; }
    // the same values from the raw interpolants, for programs that keep a fragment's inputs
    // and shade it later: w() == w_of(w_raw()), pvar(k, w) == pvar_of(p_raw(k), w)
    __device__ float w_raw() const { return v[L::IW]; }
    __device__ float p_raw(int k) const { return v[L::P + k]; }
    __device__ static float w_of(float raw) { return 1.0f / raw; }
    __device__ static float pvar_of(float raw, float w) { return raw * w; }
    // a fragment's planes can be chained again at its pixel (chain_at)
    static constexpr bool REPLAY = true;
};

// Plane p of a primitive record at pixel (x, y) by the chain the raster kernel runs, in the
// same operations: Block (PixelShaderBase.h:55-94): init at the 8x8 block's first pixel
// centre, +b per row, +a per column; Span (PixelShaderBase.h:96-112): init at the span's
// first pixel xl of row y, +a per pixel after it.  (x, y) must be a pixel the primitive
// covers (for Span, xl <= x).
// SPANS = false: the record is known to be a Block primitive.
template <bool SPANS = true>
__device__ inline float chain_at(const float *rec, int p, int x, int y, int sc_minx, int sc_maxx)
{
    const float pa = rec[REC_PARAM + 3 * p], pb = rec[REC_PARAM + 3 * p + 1],
                pc = rec[REC_PARAM + 3 * p + 2];
    const uint32_t flags = __float_as_uint(rec[REC_FLAGS]);
    if (!SPANS || (flags & FLAG_BLOCK)) {
        // y - by additions of b, then x - bx of a (0..7 each), as selects from the running
        // sums: no per-lane trip counts
        const int bx = x & ~(TILE - 1), by = y & ~(TILE - 1);
        const int dy = y - by, dx = x - bx;
        float r = pa * ((float)bx + 0.5f) + pb * ((float)by + 0.5f) + pc, t = r;
#pragma unroll
        for (int k = 1; k < TILE; ++k) {
            t = t + pb;
            r = k <= dy ? t : r;
        }
        t = r;
#pragma unroll
        for (int i = 1; i < TILE; ++i) {
            t = t + pa;
            r = i <= dx ? t : r;
        }
        return r;
    }
    int xl = 0, xr = 0;
    span_row_extent(rec, flags, y, sc_minx, sc_maxx, xl, xr);
    return chain_add(pa * ((float)xl + 0.5f) + pb * ((float)y + 0.5f) + pc, pa, x - xl);
}

// ---- pixel programs (include/swr_hip.h) ---------------------------------------------------
template <class L>
struct ProgCoverage {
    static constexpr int KEY_WORD = REC_KEY;
    static constexpr bool PLANES = false;
    struct State { bool hit; };
    __device__ static void load(State &s, const RasterArgs &, size_t) { s.hit = false; }
    template <class PX>
    __device__ static void shade(State &s, const RasterArgs &, const PX &, uint32_t, int, int)
    {
        s.hit = true;
    }
    template <class PX = void>
    __device__ static void store(const State &s, const RasterArgs &a, size_t o, int x, int y)
    {
        if (s.hit)
            static_cast<uint32_t *>(a.t0)[o] = 1u;
    }
};

// RasterizerTest.cpp:38-48.  Every fragment overwrites the pixel, so only the last one
// counts: for triangles shade keeps its record (the raster kernel chains nothing, CHAINED)
// and store chains its colour planes at the pixel (chain_at) and packs them once.
template <class L>
struct ProgRGB {
    static constexpr int KEY_WORD = REC_KEY;
    static constexpr bool PLANES = true;
    static constexpr int CHAINED = 0;
    struct State { uint32_t c; uint32_t prim; bool hit; };
    __device__ static void load(State &s, const RasterArgs &, size_t)
    {
        s.hit = false;
        s.c = 0;
        s.prim = 0;
    }
    __device__ static uint32_t pack(float r0, float g0, float b0)
    {
        uint32_t r = (uint32_t)f2i(r0 * 255.0f);
        uint32_t g = (uint32_t)f2i(g0 * 255.0f);
        uint32_t b = (uint32_t)f2i(b0 * 255.0f);
        return (r << 16) | (g << 8) | b;
    }
    template <class PX>
    __device__ static void shade(State &s, const RasterArgs &, const PX &v, uint32_t, int, int)
    {
        if constexpr (PX::REPLAY)
            s.prim = v.prim;
        else
            s.c = pack(v.avar(0), v.avar(1), v.avar(2));
        s.hit = true;
    }
    template <class PX>
    __device__ static void store(const State &s, const RasterArgs &a, size_t o, int x, int y)
    {
        if (!s.hit)
            return;
        uint32_t c = s.c;
        if constexpr (PX::REPLAY) {
            const float *rec = a.prims + (size_t)s.prim * (size_t)a.stride;
            c = pack(chain_at<PX::SPANS>(rec, L::A, x, y, a.sc_minx, a.sc_maxx),
                     chain_at<PX::SPANS>(rec, L::A + 1, x, y, a.sc_minx, a.sc_maxx),
                     chain_at<PX::SPANS>(rec, L::A + 2, x, y, a.sc_minx, a.sc_maxx));
        }
        static_cast<uint32_t *>(a.t0)[o] = c;
    }
};

// Only the last depth-passing fragment of a pixel reaches the targets, and its colour is a
// pure function of its interpolants.  So shade keeps only that fragment: for triangles its
// primitive record (the raster kernel chains z alone, CHAINED), and store chains its other
// planes again at the pixel (chain_at: the same operations on the same values); for points
// and lines its raw values.  Store computes the colour once.
template <class L>
struct ProgDepthChecker {
    static constexpr int KEY_WORD = REC_KEY;
    static constexpr bool PLANES = true;
    static constexpr int CHAINED = 1;  // z
    struct State { float depth, wr, p0, p1, a0, a1, a2; uint32_t prim; bool hit; };
    __device__ static void load(State &s, const RasterArgs &a, size_t o)
    {
        s.depth = static_cast<const float *>(a.t1)[o];
        s.wr = s.p0 = s.p1 = s.a0 = s.a1 = s.a2 = 0.0f;
        s.prim = 0;
        s.hit = false;
    }
    template <class PX>
    __device__ static void shade(State &s, const RasterArgs &, const PX &v, uint32_t, int, int)
    {
        const float z = v.z();
        if (z < s.depth) {
            s.depth = z;
            if constexpr (PX::REPLAY) {
                s.prim = v.prim;
            } else {
                s.wr = v.w_raw();
                s.p0 = v.p_raw(0);
                s.p1 = v.p_raw(1);
                s.a0 = v.avar(0);
                s.a1 = v.avar(1);
                s.a2 = v.avar(2);
            }
            s.hit = true;
        }
    }
    template <class PX>
    __device__ static void store(const State &s0, const RasterArgs &a, size_t o, int x, int y)
    {
        if (s0.hit) {
            State s = s0;
            if constexpr (PX::REPLAY) {
                const float *rec = a.prims + (size_t)s.prim * (size_t)a.stride;
                s.wr = chain_at<PX::SPANS>(rec, L::IW, x, y, a.sc_minx, a.sc_maxx);
                s.a0 = chain_at<PX::SPANS>(rec, L::A, x, y, a.sc_minx, a.sc_maxx);
                s.a1 = chain_at<PX::SPANS>(rec, L::A + 1, x, y, a.sc_minx, a.sc_maxx);
                s.a2 = chain_at<PX::SPANS>(rec, L::A + 2, x, y, a.sc_minx, a.sc_maxx);
                s.p0 = chain_at<PX::SPANS>(rec, L::P, x, y, a.sc_minx, a.sc_maxx);
                s.p1 = chain_at<PX::SPANS>(rec, L::P + 1, x, y, a.sc_minx, a.sc_maxx);
            }
            const float w = PX::w_of(s.wr);  // only the depth-passing pixels read pvar
            const float u = PX::pvar_of(s.p0, w), t = PX::pvar_of(s.p1, w);
            int cu = f2i(floorf(u * a.checker_cells));
            int cv = f2i(floorf(t * a.checker_cells));
            float k = (((uint32_t)cu + (uint32_t)cv) & 1u) ? 1.0f : 0.5f;
            uint32_t r = (uint32_t)(clamp01(s.a0 * k) * 255.0f);
            uint32_t g = (uint32_t)(clamp01(s.a1 * k) * 255.0f);
            uint32_t b = (uint32_t)(clamp01(s.a2 * k) * 255.0f);
            static_cast<float *>(a.t1)[o] = s.depth;
            static_cast<uint32_t *>(a.t0)[o] = r | (g << 8) | (b << 16) | 0xFF000000u;
        }
    }
};

template <class L>
struct ProgHash {
    static constexpr int KEY_WORD = REC_KEY;
    static constexpr bool PLANES = false;
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
    template <class PX = void>
    __device__ static void store(const State &s, const RasterArgs &a, size_t o, int x, int y)
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
    static constexpr bool PLANES = false;
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
    template <class PX = void>
    __device__ static void store(const State &s, const RasterArgs &a, size_t o, int x, int y)
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
    static constexpr bool PLANES = false;
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
    template <class PX = void>
    __device__ static void store(const State &s, const RasterArgs &a, size_t o, int x, int y)
    {
        if (s.hit) {
            static_cast<float *>(a.t1)[o] = s.depth;
            static_cast<uint32_t *>(a.t0)[o] = s.id;
        }
    }
};

// ---- the texture program ----------------------------------------------------------------
// Texture::sample and its helpers (Texture.h:35-228) over a TexDev, operation for operation;
// the reference's std::min / std::max / std::clamp argument orders are kept (they decide
// NaN cases) and its Uint8 conversions are (int) truncation then the low 8 bits, as x86-64
// converts.  log2 of the level of detail is rounded once (see tex_log2).

__device__ inline uint32_t tex_ch(uint32_t c, int k) { return (c >> (16 - 8 * k)) & 0xFFu; }
__device__ inline uint32_t tex_pack(uint32_t r, uint32_t g, uint32_t b)
{
    return ((r & 0xFFu) << 16) | ((g & 0xFFu) << 8) | (b & 0xFFu);
}
__device__ inline int tex_imin(int a, int b) { return (b < a) ? b : a; }
__device__ inline float tex_clamp(float v, float lo, float hi)
{
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}

// fmodf(u, 1) then + 1 below zero (Texture.h:43-46): u - trunc(u) is exact for every finite
// float (Sterbenz), and carries the sign of u as fmod does (also for a zero result).
__device__ inline float tex_wrap(float u)
{
    float r = __builtin_copysignf(u - __builtin_truncf(u), u);
    if (r < 0)
        r += 1.0f;
    return r;
}

// log2 rounded once to float: x = m 2^e, m in [sqrt(1/2), sqrt(2)); log(m) = 2 atanh(s),
// s = (m - 1)/(m + 1), as a degree-25 odd series in double.  The C library's log2f (which the
// reference calls) may round the other way by one ulp.
__device__ inline float tex_log2(float xf)
{
    if (!(xf > 0.0f))
        return xf == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (__builtin_isinf(xf))
        return xf;
    const unsigned long long bits = (unsigned long long)__double_as_longlong((double)xf);
    int e = (int)((bits >> 52) & 0x7FFull) - 1022;
    double m = __longlong_as_double((long long)((bits & ~(0x7FFull << 52)) | (1022ull << 52)));
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double q = 1.0 / 25.0;
    q = q * s2 + 1.0 / 23.0;
    q = q * s2 + 1.0 / 21.0;
    q = q * s2 + 1.0 / 19.0;
    q = q * s2 + 1.0 / 17.0;
    q = q * s2 + 1.0 / 15.0;
    q = q * s2 + 1.0 / 13.0;
    q = q * s2 + 1.0 / 11.0;
    q = q * s2 + 1.0 / 9.0;
    q = q * s2 + 1.0 / 7.0;
    q = q * s2 + 1.0 / 5.0;
    q = q * s2 + 1.0 / 3.0;
    q = q * s2 + 1.0;
    const double ln_m = 2.0 * s * q;
    return (float)((double)e + ln_m * 1.4426950408889634074);
}

// sampleBilinear + bilinearLerpColors, Texture.h:148-228
__device__ inline uint32_t tex_bilinear(const TexDev &t, int lv, float u, float v)
{
    const int w = t.w[lv], h = t.h[lv];
    const float px = u * (float)(w - 1), py = v * (float)(h - 1);
    const int x0 = imax(0, f2i(floorf(px))), y0 = imax(0, f2i(floorf(py)));
    const int x1 = tex_imin(w - 1, x0 + 1), y1 = tex_imin(h - 1, y0 + 1);
    const float fx = px - (float)x0, fy = py - (float)y0;
    const uint32_t *p = t.texels + t.off[lv];
    const uint32_t c00 = p[(size_t)y0 * w + x0], c10 = p[(size_t)y0 * w + x1];
    const uint32_t c01 = p[(size_t)y1 * w + x0], c11 = p[(size_t)y1 * w + x1];
    uint32_t out[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float sum = (float)tex_ch(c00, k) * (1.0f - fx) * (1.0f - fy) +
                          (float)tex_ch(c10, k) * fx * (1.0f - fy) +
                          (float)tex_ch(c01, k) * (1.0f - fx) * fy + (float)tex_ch(c11, k) * fx * fy;
        out[k] = (uint32_t)(uint8_t)f2i(sum);
    }
    return tex_pack(out[0], out[1], out[2]);
}

// sampleTrilinear + lerpColors, Texture.h:121-146, 200-205
__device__ inline uint32_t tex_trilinear(const TexDev &t, float u, float v, float rho)
{
    float lod = tex_log2(rho < 1e-6f ? 1e-6f : rho);
    lod = tex_clamp(lod, 0.0f, (float)(t.levels - 1));
    const int base = f2i(floorf(lod));
    const int next = tex_imin(base + 1, t.levels - 1);
    const float frac = tex_clamp(lod - (float)base, 0.0f, 1.0f);
    const uint32_t cb = tex_bilinear(t, base, u, v);
    if (base == next)
        return cb;
    const uint32_t cn = tex_bilinear(t, next, u, v);
    uint32_t out[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int d = (int)tex_ch(cn, k) - (int)tex_ch(cb, k);
        const uint8_t step = (uint8_t)f2i((float)d * frac);
        out[k] = (uint8_t)(tex_ch(cb, k) + step);
    }
    return tex_pack(out[0], out[1], out[2]);
}

// Texture::sample, Texture.h:35-118
__device__ inline uint32_t tex_sample(const TexDev &t, float u, float v, float dudx, float dvdx,
                                      float dudy, float dvdy)
{
    u = tex_wrap(u);
    v = tex_wrap(v);
    const float W = (float)t.w[0], H = (float)t.h[0];
    const float dudx_s = dudx * W, dvdx_s = dvdx * H, dudy_s = dudy * W, dvdy_s = dvdy * H;
    float dx_len = sqrtf(dudx_s * dudx_s + dvdx_s * dvdx_s);
    float dy_len = sqrtf(dudy_s * dudy_s + dvdy_s * dvdy_s);
    dx_len = ref_max(dx_len, 1e-6f);
    dy_len = ref_max(dy_len, 1e-6f);
    const float major = ref_max(dx_len, dy_len), minor = ref_min(dx_len, dy_len);
    const float ratio = ref_min(major / minor, (float)t.max_aniso);
    const int n = imax(1, f2i(ceilf(ratio)));
    if (n <= 1)
        return tex_trilinear(t, u, v, major);
    float mdu, mdv;
    if (dx_len > dy_len) {
        mdu = dudx / dx_len;
        mdv = dvdx / dx_len;
    } else {
        mdu = dudy / dy_len;
        mdv = dvdy / dy_len;
    }
    float r = 0, g = 0, b = 0;
    const float step = 1.0f / (float)n;
    for (int i = 0; i < n; ++i) {
        const float tt = ((float)i + 0.5f) * step - 0.5f;
        const float su = tex_wrap(u + mdu * major * tt);
        const float sv = tex_wrap(v + mdv * major * tt);
        const uint32_t c = tex_trilinear(t, su, sv, minor);
        r += (float)tex_ch(c, 0);
        g += (float)tex_ch(c, 1);
        b += (float)tex_ch(c, 2);
    }
    r = ref_min(255.0f, r / (float)n);
    g = ref_min(255.0f, g / (float)n);
    b = ref_min(255.0f, b / (float)n);
    return tex_pack((uint8_t)f2i(r), (uint8_t)f2i(g), (uint8_t)f2i(b));
}

// The Box demo's shader (Box.cpp:39-53): computePerspectiveDerivatives of pvar 0 and 1
// (PixelData.h:127-145: the planes evaluated directly at the pixel centre), then sample.
template <class L>
struct ProgTexture {
    static constexpr int KEY_WORD = REC_KEY;
    static constexpr bool PLANES = true;
    // every fragment overwrites the pixel: shade keeps the last one's record and store
    // samples once, from the planes chained again at the pixel (chain_at)
    static constexpr int CHAINED = 0;
    struct State { uint32_t prim; bool hit; };
    __device__ static void load(State &s, const RasterArgs &, size_t)
    {
        s.hit = false;
        s.prim = 0;
    }
    __device__ static void derivs(const float *pl, const float *iw, float xf, float yf,
                                  float &ddx, float &ddy)
    {
        const float var = pl[0] * xf + pl[1] * yf + pl[2];
        const float cur = iw[0] * xf + iw[1] * yf + iw[2];
        ddx = (cur * pl[0] - var * iw[0]) / (cur * cur);
        ddy = (cur * pl[1] - var * iw[1]) / (cur * cur);
    }
    template <class PX>
    __device__ static void shade(State &s, const RasterArgs &, const PX &v, uint32_t, int, int)
    {
        s.prim = v.prim;
        s.hit = true;
    }
    template <class PX>
    __device__ static void store(const State &s, const RasterArgs &a, size_t o, int x, int y)
    {
        if (!s.hit)
            return;
        const float *rec = a.prims + (size_t)s.prim * (size_t)a.stride;
        const float w = PX::w_of(chain_at<PX::SPANS>(rec, L::IW, x, y, a.sc_minx, a.sc_maxx));
        const float u =
            PX::pvar_of(chain_at<PX::SPANS>(rec, L::P, x, y, a.sc_minx, a.sc_maxx), w);
        const float t =
            PX::pvar_of(chain_at<PX::SPANS>(rec, L::P + 1, x, y, a.sc_minx, a.sc_maxx), w);
        const float *P = rec + REC_PARAM;
        const float xf = (float)x + 0.5f, yf = (float)y + 0.5f;
        float dudx, dudy, dvdx, dvdy;
        derivs(P + 3 * (L::P + 0), P + 3 * L::IW, xf, yf, dudx, dudy);
        derivs(P + 3 * (L::P + 1), P + 3 * L::IW, xf, yf, dvdx, dvdy);
        static_cast<uint32_t *>(a.t0)[o] = tex_sample(*a.tex, u, t, dudx, dvdx, dudy, dvdy);
    }
};

}  // namespace swr
