// swr_lines_kernel.h — the raster stage of points and lines: one wave64 per 8x8 tile, lane =
// pixel, the tile's fragment records in draw order through the pixel program.  A template
// over the program, like the triangle kernel (swr_raster_kernel.h), so that programs compiled
// outside the library (include/swr/pixel_program.h) instantiate it too.
#pragma once

#include "swr_common.h"
#include "swr_programs.h"

namespace swr {

// fragment record words: pixel x, y (int), key, source primitive, z, w, varyings
constexpr int FR_X = 0, FR_Y = 1, FR_KEY = 2, FR_SRC = 3, FR_Z = 4, FR_W = 5, FR_VAR = 6;

// A pixel of a point or line step: PixelData as pixelDataFromVertex fills it
// (Rasterizer.h:160-173): w = v.w, invw = 1/v.w, pvar = v.pvar (no perspective divide).
template <class L>
struct VertexPixel {
    static constexpr bool SPANS = true;  // (no primitive records to replay)
    float zv, wv;
    float av[L::NA > 0 ? L::NA : 1];
    float pv[L::NP > 0 ? L::NP : 1];
    __device__ float z() const { return zv; }
    __device__ float invw() const { return 1.0f / wv; }
    __device__ float w() const { return wv; }
    __device__ float avar(int k) const { return av[k]; }
    __device__ float pvar(int k, float) const { return pv[k]; }
    __device__ float w_raw() const { return wv; }
    __device__ float p_raw(int k) const { return pv[k]; }
    __device__ static float w_of(float raw) { return raw; }
    __device__ static float pvar_of(float raw, float) { return raw; }
    static constexpr bool REPLAY = false;  // no planes: programs keep the values
    uint32_t prim = 0;
};

// One wave64 per 8x8 tile, lane = pixel: the tile's fragments, 64 at a time, in order.
template <template <class> class ProgT, int NZ, int NW, int NA, int NP>
__global__ void __launch_bounds__(256) k_lines_raster(RasterArgs a)
{
    using L = Layout<NZ, NW, NA, NP>;
    using Prog = ProgT<L>;
    const int lane = threadIdx.x & 63;
    const int tile = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) +
                                                          (threadIdx.x >> 6)));
    if (tile >= a.tiles)
        return;
    if (a.status[ST_OVERFLOW] | a.status[ST_ERROR])
        return;
    const uint32_t beg = a.ranges[2 * tile], end = a.ranges[2 * tile + 1];
    if (beg >= end)
        return;
    const int bx = (tile % a.tiles_x) * TILE, by = (tile / a.tiles_x) * TILE;
    const int px = bx + (lane & 7), py = by + (lane >> 3);
    const bool inside = px < a.width && py < a.height;
    const size_t o = (size_t)py * (size_t)a.pitch + (size_t)px;
    const int kw = Prog::KEY_WORD == REC_SRC ? FR_SRC : FR_KEY;

    typename Prog::State st;
    if (inside)
        Prog::load(st, a, o);
    for (uint32_t c = beg; c < end; c += 64) {
        const uint32_t n = (end - c) < 64u ? (end - c) : 64u;
        uint32_t f = 0;
        int pix = -1;
        if ((uint32_t)lane < n) {
            f = a.vals[c + lane];
            const float *r = a.prims + (size_t)f * (size_t)a.stride;
            const int fx = (int)__float_as_uint(r[FR_X]), fy = (int)__float_as_uint(r[FR_Y]);
            pix = (fy - by) * 8 + (fx - bx);
        }
        for (uint32_t j = 0; j < n; ++j) {
            const int pj = __builtin_amdgcn_readlane(pix, (int)j);
            if (pj != lane || !inside)
                continue;
            const uint32_t fj = (uint32_t)__builtin_amdgcn_readlane((int)f, (int)j);
            const float *r = a.prims + (size_t)fj * (size_t)a.stride;
            VertexPixel<L> p;
            p.zv = r[FR_Z];
            p.wv = r[FR_W];
#pragma unroll
            for (int k = 0; k < L::NA; ++k)
                p.av[k] = r[FR_VAR + k];
#pragma unroll
            for (int k = 0; k < L::NP; ++k)
                p.pv[k] = r[FR_VAR + L::NA + k];
            Prog::shade(st, a, p, __float_as_uint(r[kw]), px, py);
        }
    }
    if (inside)
        Prog::template store<VertexPixel<L>>(st, a, o, px, py);
}

template <template <class> class ProgT, int NZ, int NW, int NA, int NP>
inline hipError_t launch_lr(const RasterArgs &a, hipStream_t s)
{
    const uint32_t blocks = ((uint32_t)a.tiles + 3) / 4;
    hipLaunchKernelGGL((k_lines_raster<ProgT, NZ, NW, NA, NP>), dim3(blocks > 0 ? blocks : 1),
                       dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace swr
