// swr_vertex.h — per-vertex work shared by the triangle, line and point paths: the
// built-in vertex programs, clip masks and plane distances, the clip lerp and the
// perspective divide + viewport transform.
#pragma once

#include "swr_common.h"

#pragma clang fp contract(off)

namespace swr {

struct Pos4 {
    float x, y, z, w;
};

// ---- vertex programs (VertexProcessor.cpp:134-150 + the built-in shaders) ----------------
__device__ inline void mvp_apply(const float *m, float x, float y, float z, Pos4 &o)
{
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float s = 0.0f;  // vector_math.h:55-60 sequential dot product
        s = s + m[4 * i + 0] * x;
        s = s + m[4 * i + 1] * y;
        s = s + m[4 * i + 2] * z;
        s = s + m[4 * i + 3] * 1.0f;
        r[i] = s;
    }
    o.x = r[0];
    o.y = r[1];
    o.z = r[2];
    o.w = r[3];
}

// Carried varyings: var[0 .. lay.na) avar, var[lay.na .. lay.na+lay.np) pvar.  POS = false:
// the varyings only (the position comes from the post-transform vertices).
template <int NV, bool POS = true>
__device__ inline void fetch_vertex(const GeomArgs &a, int32_t index, Pos4 &p, float *var)
{
    p.x = p.y = p.z = p.w = 0.0f;
#pragma unroll
    for (int k = 0; k < NV; ++k)
        var[k] = 0.0f;
    if (a.vs_program == SWR_VS_DUMMY)
        return;
    const float *src = reinterpret_cast<const float *>(a.attr0 + (size_t)a.stride0 * (size_t)index);
    const int na = a.lay.na, np = a.lay.np;
    switch (a.vs_program) {
    case SWR_VS_RAW: {
        p.x = src[0];
        p.y = src[1];
        p.z = src[2];
        p.w = src[3];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if (k < na) {
                if (k < a.vs_na)
                    var[k] = src[4 + k];
            } else if (k < na + np) {
                int q = k - na;
                if (q < a.vs_np)
                    var[k] = src[4 + a.vs_na + q];
            }
        }
        break;
    }
    case SWR_VS_POS_RGB:
        p.x = src[0];
        p.y = src[1];
        p.z = src[2];
        p.w = 1.0f;
#pragma unroll
        for (int k = 0; k < NV && k < 3; ++k)
            if (k < na)
                var[k] = src[3 + k];
        break;
    case SWR_VS_MVP_POS_RGB_UV:
        if (POS)
            mvp_apply(a.mvp, src[0], src[1], src[2], p);
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            if (k < na && k < 3)
                var[k] = src[3 + k];
            else if (k >= na && k < na + np && k - na < 2)
                var[k] = src[6 + (k - na)];
        }
        break;
    case SWR_VS_MVP_POS_UV:
        if (POS)
            mvp_apply(a.mvp, src[0], src[1], src[2], p);
#pragma unroll
        for (int k = 0; k < NV; ++k)
            if (k >= na && k < na + np && k - na < 2)
                var[k] = src[6 + (k - na)];
        break;
    default:
        break;
    }
}

// VertexProcessor::clipMask, VertexProcessor.cpp:122-132
__device__ inline int clip_mask(const Pos4 &v)
{
    int m = 0;
    if (v.w - v.x < 0) m |= 0x01;
    if (v.x + v.w < 0) m |= 0x02;
    if (v.w - v.y < 0) m |= 0x04;
    if (v.y + v.w < 0) m |= 0x08;
    if (v.w - v.z < 0) m |= 0x10;
    if (v.z + v.w < 0) m |= 0x20;
    return m;
}

// a*x + b*y + c*z + d*w for the plane sequence of VertexProcessor.cpp:237-242
__device__ inline float plane_dot(int p, const Pos4 &v)
{
    const float a = p == 0 ? -1.0f : (p == 1 ? 1.0f : 0.0f);
    const float b = p == 2 ? -1.0f : (p == 3 ? 1.0f : 0.0f);
    const float c = p == 4 ? -1.0f : (p == 5 ? 1.0f : 0.0f);
    return a * v.x + b * v.y + c * v.z + 1.0f * v.w;
}

__device__ inline int sgnf(float v) { return (0.0f < v) - (v < 0.0f); }

__device__ inline float lerp1(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// Window-space vertex: VertexProcessor::transformVertices, VertexProcessor.cpp:364-373
struct WinV {
    float x, y, z, w;
};

__device__ inline WinV to_window(const GeomArgs &a, const Pos4 &p)
{
    WinV o;
    float invW = 1.0f / p.w;
    float x = p.x * invW, y = p.y * invW, z = p.z * invW;
    o.x = (a.vp_px * x + a.vp_ox);
    o.y = (a.vp_py * -y + a.vp_oy);
    o.z = a.dz_scale * z + a.dz_bias;
    o.w = p.w;
    return o;
}

}  // namespace swr
