// swr_lines.hip — points and lines on gfx950.
//
// The reference draws a line as a DDA (Rasterizer.h:175-222): steps = max(|dx|, |dy|) of the
// truncated endpoints, one per-step increment per value, each step's vertex handed to the
// pixel shader as it is (pixelDataFromVertex, Rasterizer.h:160-173) when it passes the
// scissor test.  The vertex of step k is a chain of k float additions, so it is produced by
// walking the chain, one thread per line.  A point is a line of one step with no increment
// (drawPointTemplate, Rasterizer.h:149-158).
//
// Draw sequence:
//   count   per primitive: vertex program, LineClipper / clipPoints, viewport transform and
//           a walk of the x/y chain counting the fragments that pass the scissor test and
//           land on the target; totals into the status words
//   scan    exclusive scan of the counts: each primitive's first fragment, in draw order
//   emit    the same walk with every value: fragment records and (tile, fragment) pairs
//   sort    stable radix sort of the pairs by tile (the triangle path's sort)
//   raster  one wave64 per 8x8 tile: lane = pixel, the tile's fragments in draw order
//           through the pixel program (the triangle path's programs)
// Fragment order within a pixel is the reference's invocation order: fragments are numbered
// in (primitive, step) order and the sort is stable.
#include "swr_common.h"
#include "swr_internal.h"
#include "swr_lines_kernel.h"
#include "swr_programs.h"
#include "swr_vertex.h"

namespace swr {

int fragment_stride(const ParamLayout &L) { return FR_VAR + emit_class(L.na + L.np); }

template <int NV>
struct LVert {
    float x, y, z, w;
    float v[NV > 0 ? NV : 1];
};

// LineClipper::clipToPlane, LineClipper.cpp:30-56, plane p of VertexProcessor.cpp:197-202
__device__ inline void clip_line_plane(int p, const Pos4 &v0, const Pos4 &v1, float &t0, float &t1,
                                       bool &clipped)
{
    if (clipped)
        return;
    const float dp0 = plane_dot(p, v0), dp1 = plane_dot(p, v1);
    const bool n0 = dp0 < 0, n1 = dp1 < 0;
    if (n0 && n1) {
        clipped = true;
        return;
    }
    if (n0) {
        const float t = -dp0 / (dp1 - dp0);
        t0 = ref_max(t0, t);
    } else {
        const float t = dp0 / (dp0 - dp1);
        t1 = ref_min(t1, t);
    }
}

// Window-space vertices of primitive i.  False when the primitive draws nothing (skipped,
// clipped, or an index error, which is flagged in the status words).
//   drawElements: VertexProcessor.cpp:78-120 (batches of 1024 primitives) -> clipLines
//     (:167-215) or clipPoints (:152-165) -> transformVertices (:347-377) -> the list entry.
//   IRasterizer seam: drawLineList / drawPointList (Rasterizer.h:116-132).
template <int NV>
__device__ bool prepare_primitive(const LineArgs &A, uint64_t i, LVert<NV> *V)
{
    const GeomArgs &a = A.g;
    const int nv = A.mode == SWR_DRAW_LINE ? 2 : 1;
    const uint64_t ib = (uint64_t)nv * i;
    int32_t ids[2];
    ids[0] = a.indices[ib];
    ids[1] = nv == 2 ? a.indices[ib + 1] : ids[0];
    if (a.window_input) {
        if (ids[0] == -1)
            return false;  // Rasterizer.h:119, 128
#pragma unroll
        for (int q = 0; q < 2; ++q)
            if (q < nv && (ids[q] < 0 || (uint64_t)ids[q] >= a.wvert_count)) {
                atomicOr(&a.status[ST_ERROR], 1ull);
                return false;
            }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const swr_rasterizer_vertex &s = a.wverts[ids[q]];
            V[q].x = s.x;
            V[q].y = s.y;
            V[q].z = s.z;
            V[q].w = s.w;
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const int k2 = k - a.lay.na;
                V[q].v[k] = k < a.lay.na ? s.avar[k & 15] : (k2 < a.lay.np ? s.pvar[k2 & 15] : 0.0f);
            }
        }
        return true;
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
        if (q < nv && (ids[q] < 0 || (a.attr0_bytes && (uint64_t)ids[q] * (uint64_t)a.stride0 +
                                                            a.attr0_need > a.attr0_bytes))) {
            atomicOr(&a.status[ST_ERROR], 1ull);
            return false;
        }
    Pos4 P[2];
    float var[2][NV > 0 ? NV : 1];
#pragma unroll
    for (int q = 0; q < 2; ++q)
        fetch_vertex<NV>(a, ids[q], P[q], var[q]);  // a point fetches its vertex twice
    if (nv == 1) {
        if (clip_mask(P[0]))
            return false;
    } else {
        const int m0 = clip_mask(P[0]), m1 = clip_mask(P[1]);
        const int cm = m0 | m1;
        float t0 = 0.0f, t1 = 1.0f;
        bool clipped = false;
#pragma unroll
        for (int p = 0; p < 6; ++p)
            if (cm & (1 << p))
                clip_line_plane(p, P[0], P[1], t0, t1, clipped);
        if (clipped)
            return false;
        // Helper::interpolateVertex (PolyClipper.h:34-48) from the unclipped endpoints
        const Pos4 A0 = P[0], A1 = P[1];
        float var0[NV > 0 ? NV : 1], var1[NV > 0 ? NV : 1];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            var0[k] = var[0][k];
            var1[k] = var[1][k];
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (!(q == 0 ? m0 : m1))
                continue;
            const float t = q == 0 ? t0 : t1;
            P[q].x = lerp1(A0.x, A1.x, t);
            P[q].y = lerp1(A0.y, A1.y, t);
            P[q].z = lerp1(A0.z, A1.z, t);
            P[q].w = lerp1(A0.w, A1.w, t);
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const bool lerped = k < a.lay.na ? (k < a.vs_na) : ((k - a.lay.na) < a.vs_np);
                var[q][k] = lerped ? lerp1(var0[k], var1[k], t) : 0.0f;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const WinV w = to_window(a, P[q]);
        V[q].x = w.x;
        V[q].y = w.y;
        V[q].z = w.z;
        V[q].w = w.w;
#pragma unroll
        for (int k = 0; k < NV; ++k)
            V[q].v[k] = var[q][k];
    }
    return true;
}

// Rasterizer::scissorTest, Rasterizer.h:143-147
__device__ inline bool scissor_pass(const GeomArgs &a, float x, float y)
{
    return x >= (float)a.sc_minx && x < (float)a.sc_maxx && y >= (float)a.sc_miny &&
           y < (float)a.sc_maxy;
}

// The pixel a step lands on, when the shader would touch the target there.
__device__ inline bool on_target(const GeomArgs &a, float x, float y, int &px, int &py)
{
    if (!scissor_pass(a, x, y))
        return false;
    px = f2i(x);
    py = f2i(y);
    return px >= 0 && py >= 0 && px < a.width && py < a.height;
}

// Walks primitive i.  EMIT = false: returns its fragment count.  EMIT = true: writes its
// fragments from A.first[i] on.
template <int NV, bool EMIT>
__device__ uint32_t walk_primitive(const LineArgs &A, uint64_t i, bool &live)
{
    const GeomArgs &a = A.g;
    LVert<NV> V[2];
    live = prepare_primitive<NV>(A, i, V);
    if (!live)
        return 0;
    const uint32_t key = a.window_input ? (uint32_t)i
                                        : ((uint32_t)(i >> 10) << 15) + (uint32_t)(i & 1023u);
    const uint32_t src = (uint32_t)i;
    uint64_t f = EMIT ? A.first[i] : 0;
    uint32_t n = 0;
    auto emit = [&](int px, int py, const LVert<NV> &v) {
        if (EMIT) {
            float *r = A.frags + f * (uint64_t)A.fstride;
            r[FR_X] = __uint_as_float((uint32_t)px);
            r[FR_Y] = __uint_as_float((uint32_t)py);
            r[FR_KEY] = __uint_as_float(key);
            r[FR_SRC] = __uint_as_float(src);
            r[FR_Z] = v.z;
            r[FR_W] = v.w;
#pragma unroll
            for (int k = 0; k < NV; ++k)
                r[FR_VAR + k] = v.v[k];
            A.keys[f] = (uint32_t)(py >> 3) * (uint32_t)a.tiles_x + (uint32_t)(px >> 3);
            A.vals[f] = (uint32_t)f;
            ++f;
        }
        ++n;
    };
    int px, py;
    if (A.mode == SWR_DRAW_POINT) {
        if (on_target(a, V[0].x, V[0].y, px, py))
            emit(px, py, V[0]);
        return n;
    }
    // Rasterizer::drawLineTemplate / computeVertexStep / stepVertex, Rasterizer.h:175-222
    int64_t adx = (int64_t)f2i(V[1].x) - (int64_t)f2i(V[0].x);
    int64_t ady = (int64_t)f2i(V[1].y) - (int64_t)f2i(V[0].y);
    adx = adx < 0 ? -adx : adx;
    ady = ady < 0 ? -ady : ady;
    const int64_t steps = adx > ady ? adx : ady;
    if (steps > SWR_MAX_LINE_STEPS) {
        atomicOr(&a.status[ST_ERROR], 4ull);
        live = false;
        return 0;
    }
    const float fs = (float)steps;
    LVert<NV> st, v = V[0];
    st.x = (V[1].x - V[0].x) / fs;
    st.y = (V[1].y - V[0].y) / fs;
    st.z = (V[1].z - V[0].z) / fs;
    st.w = (V[1].w - V[0].w) / fs;
#pragma unroll
    for (int k = 0; k < NV; ++k)
        st.v[k] = (V[1].v[k] - V[0].v[k]) / fs;
    for (int64_t s = 0; s < steps; ++s) {
        if (on_target(a, v.x, v.y, px, py))
            emit(px, py, v);
        v.x += st.x;
        v.y += st.y;
        if (EMIT) {
            if (A.iz)
                v.z += st.z;
            if (A.iw)
                v.w += st.w;
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const bool stepped = k < a.lay.na ? (k < a.lay.set_na) : ((k - a.lay.na) < a.lay.set_np);
                if (stepped)
                    v.v[k] += st.v[k];
            }
        }
    }
    return n;
}

__device__ inline unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void __launch_bounds__(256) k_lines_begin(LineArgs A)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ST_WORDS && i != ST_STICKY)
        A.g.status[i] = 0ull;
    for (uint64_t k = i; k < 2ull * A.g.tiles; k += (uint64_t)gridDim.x * blockDim.x)
        A.g.ranges[k] = 0u;
}

__global__ void __launch_bounds__(256) k_lines_count(LineArgs A)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = 0;
    bool live = false;
    if (i < A.prims) {
        n = walk_primitive<0, false>(A, i, live);
        A.count[i] = n;
    }
    const unsigned long long frags = wave_sum(n), prims = wave_sum(live ? 1ull : 0ull);
    if ((threadIdx.x & 63) == 0) {
        if (frags)
            atomicAdd(&A.g.status[ST_PAIRS], frags);
        if (prims)
            atomicAdd(&A.g.status[ST_PRIMS], prims);
    }
}

template <int NV>
__global__ void __launch_bounds__(256) k_lines_emit(LineArgs A)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.prims || A.count[i] == 0)
        return;
    bool live;
    walk_primitive<NV, true>(A, i, live);
}

// ---- launchers ------------------------------------------------------------------------------
static dim3 grid_of(uint64_t n) { return dim3((uint32_t)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1)); }

hipError_t launch_lines_count(const LineArgs &A, hipStream_t s)
{
    const uint64_t clear = A.g.tiles * 2ull > ST_WORDS ? A.g.tiles * 2ull : ST_WORDS;
    const uint64_t cb = (clear + 255) / 256 < 4096 ? (clear + 255) / 256 : 4096;
    hipLaunchKernelGGL(k_lines_begin, dim3((uint32_t)cb), dim3(256), 0, s, A);
    if (A.prims)
        hipLaunchKernelGGL(k_lines_count, grid_of(A.prims), dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_lines_emit(const LineArgs &A, hipStream_t s)
{
    if (!A.prims)
        return hipSuccess;
    const dim3 g = grid_of(A.prims), b(256);
    switch (emit_class(A.g.lay.na + A.g.lay.np)) {
    case 0: hipLaunchKernelGGL(k_lines_emit<0>, g, b, 0, s, A); break;
    case 3: hipLaunchKernelGGL(k_lines_emit<3>, g, b, 0, s, A); break;
    case 5: hipLaunchKernelGGL(k_lines_emit<5>, g, b, 0, s, A); break;
    case 8: hipLaunchKernelGGL(k_lines_emit<8>, g, b, 0, s, A); break;
    default: hipLaunchKernelGGL(k_lines_emit<32>, g, b, 0, s, A); break;
    }
    return hipGetLastError();
}

hipError_t launch_lines_raster(const swr_pixel_shader_desc &ps, const ParamLayout &L,
                               const RasterArgs &a, hipStream_t s)
{
    switch (ps.program) {
    case SWR_PS_COVERAGE: return launch_lr<ProgCoverage, 0, 0, 0, 0>(a, s);
    case SWR_PS_RGB: return launch_lr<ProgRGB, 0, 0, 3, 0>(a, s);
    case SWR_PS_DEPTH_CHECKER: return launch_lr<ProgDepthChecker, 1, 1, 3, 2>(a, s);
    case SWR_PS_VISIBILITY: return launch_lr<ProgVisibility, 1, 0, 0, 0>(a, s);
    case SWR_PS_HASH:
        return L.na == 4 ? launch_lr<ProgHash, 1, 1, 4, 4>(a, s)
                         : launch_lr<ProgHash, 1, 1, 16, 16>(a, s);
    case SWR_PS_LOG:
        return L.na == 4 ? launch_lr<ProgLog, 1, 1, 4, 4>(a, s)
                         : launch_lr<ProgLog, 1, 1, 16, 16>(a, s);
    case SWR_PS_MODULE:
        return (hipError_t)ps.module->launch_fragments(&a, s);
    default: return hipSuccess;
    }
}

}  // namespace swr
