/*
 * ref_driver.cpp — TEST INFRASTRUCTURE ONLY.
 *
 * Drives trenki2/SoftwareRenderer itself (its src/renderer sources, compiled where they lie
 * by oracle/Makefile; nothing is copied) through the entry points swr_oracle.h declares,
 * prefixed swr_ref_.  The built library lands in oracle/_ref/ and is used to pin the C
 * oracle and to generate golden fixtures, and as the "reference" CPU baseline.
 *
 * Shaders are the built-in programs of include/swr_hip.h written as reference shader
 * classes: the pixel shader statics (InterpolateZ/W, AVarCount, PVarCount) and vertex
 * shader statics (AttribCount, AVarCount, PVarCount) are template arguments; the program
 * body is chosen at run time from the pipeline state.
 *
 * The primitive key of each pixel-shader call is observed by splitting each batch the
 * reference's VertexProcessor hands to IRasterizer::drawTriangleList (drawLineList,
 * drawPointList) into single-primitive calls of the reference Rasterizer's own list
 * functions (Rasterizer.h:116-141), setting the key before each.
 */
#include "Renderer.h"

#include <climits>
#include <cmath>
#include <cstring>

extern "C" {
#include "swr_oracle.h"
}

namespace {

const swr_pipeline_state *g_state = nullptr;
uint32_t g_key = 0;

uint32_t fbits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

uint32_t mix32(uint32_t h, uint32_t x) { return (h ^ x) * 16777619u; }
float clamp01(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

template <bool Z, bool W, int A, int P>
struct ProgramShader : public swr::PixelShaderBase<ProgramShader<Z, W, A, P>> {
    static const bool InterpolateZ = Z;
    static const bool InterpolateW = W;
    static const int AVarCount = A;
    static const int PVarCount = P;

    static void drawPixel(const swr::PixelData &p)
    {
        const swr_pixel_shader_desc &ps = g_state->ps;
        if (p.x < 0 || p.y < 0 || p.x >= ps.width || p.y >= ps.height)
            return;
        size_t o = (size_t)p.y * (size_t)ps.pitch + (size_t)p.x;
        switch (ps.program) {
        case SWR_PS_DEPTH_CHECKER: {
            float *depth = static_cast<float *>(ps.target1);
            uint32_t *color = static_cast<uint32_t *>(ps.target0);
            if (p.z < depth[o]) {
                depth[o] = p.z;
                int cu = (int)std::floor(p.pvar[0] * ps.checker_cells);
                int cv = (int)std::floor(p.pvar[1] * ps.checker_cells);
                float k = (((unsigned)cu + (unsigned)cv) & 1u) ? 1.0f : 0.5f;
                unsigned r = (unsigned)(clamp01(p.avar[0] * k) * 255.0f);
                unsigned g = (unsigned)(clamp01(p.avar[1] * k) * 255.0f);
                unsigned b = (unsigned)(clamp01(p.avar[2] * k) * 255.0f);
                color[o] = r | (g << 8) | (b << 16) | 0xFF000000u;
            }
            break;
        }
        case SWR_PS_RGB: {
            int rint = (int)(p.avar[0] * 255);
            int gint = (int)(p.avar[1] * 255);
            int bint = (int)(p.avar[2] * 255);
            static_cast<uint32_t *>(ps.target0)[o] =
                ((uint32_t)rint << 16) | ((uint32_t)gint << 8) | (uint32_t)bint;
            break;
        }
        case SWR_PS_COVERAGE:
            static_cast<uint32_t *>(ps.target0)[o] = 1;
            break;
        case SWR_PS_HASH: {
            uint32_t *hp = static_cast<uint32_t *>(ps.target0);
            uint32_t h = mix32(hp[o], g_key);
            if (Z)
                h = mix32(h, fbits(p.z));
            if (W) {
                h = mix32(h, fbits(p.w));
                h = mix32(h, fbits(p.invw));
            }
            for (int i = 0; i < A; ++i)
                h = mix32(h, fbits(p.avar[i]));
            for (int i = 0; i < P; ++i)
                h = mix32(h, fbits(p.pvar[i]));
            hp[o] = h;
            static_cast<uint32_t *>(ps.target1)[o] += 1;
            break;
        }
        case SWR_PS_LOG: {
            uint32_t seq = static_cast<uint32_t *>(ps.target0)[o]++;
            uint64_t n = (*ps.log_count)++;
            if (n < ps.log_capacity) {
                uint32_t *rec = ps.log + n * SWR_LOG_WORDS;
                std::memset(rec, 0, SWR_LOG_WORDS * 4);
                rec[0] = (uint32_t)p.x;
                rec[1] = (uint32_t)p.y;
                rec[2] = g_key;
                rec[3] = seq;
                if (Z)
                    rec[4] = fbits(p.z);
                if (W) {
                    rec[5] = fbits(p.w);
                    rec[6] = fbits(p.invw);
                }
                for (int i = 0; i < A; ++i)
                    rec[7 + i] = fbits(p.avar[i]);
                for (int i = 0; i < P; ++i)
                    rec[7 + SWR_MAX_AVARS + i] = fbits(p.pvar[i]);
            }
            break;
        }
        default:
            break;
        }
    }
};

template <int Attr, int A, int P>
struct ProgramVertexShader : public swr::VertexShaderBase<ProgramVertexShader<Attr, A, P>> {
    static const int AttribCount = Attr;
    static const int AVarCount = A;
    static const int PVarCount = P;

    static void mvp(const float *m, const float *a, swr::VertexShaderOutput *out)
    {
        float r[4];
        for (int i = 0; i < 4; ++i) {
            float s = 0.0f;
            s += m[4 * i + 0] * a[0];
            s += m[4 * i + 1] * a[1];
            s += m[4 * i + 2] * a[2];
            s += m[4 * i + 3] * 1.0f;
            r[i] = s;
        }
        out->x = r[0];
        out->y = r[1];
        out->z = r[2];
        out->w = r[3];
    }

    static void processVertex(swr::VertexShaderInput in, swr::VertexShaderOutput *out)
    {
        const swr_vertex_shader_desc &vs = g_state->vs;
        const float *a = static_cast<const float *>(in[0]);
        switch (vs.program) {
        case SWR_VS_RAW:
            out->x = a[0];
            out->y = a[1];
            out->z = a[2];
            out->w = a[3];
            for (int i = 0; i < A; ++i)
                out->avar[i] = a[4 + i];
            for (int i = 0; i < P; ++i)
                out->pvar[i] = a[4 + A + i];
            break;
        case SWR_VS_POS_RGB:
            out->x = a[0];
            out->y = a[1];
            out->z = a[2];
            out->w = 1.0f;
            out->avar[0] = a[3];
            out->avar[1] = a[4];
            out->avar[2] = a[5];
            break;
        case SWR_VS_MVP_POS_RGB_UV:
            mvp(vs.mvp, a, out);
            out->avar[0] = a[3];
            out->avar[1] = a[4];
            out->avar[2] = a[5];
            out->pvar[0] = a[6];
            out->pvar[1] = a[7];
            break;
        case SWR_VS_MVP_POS_UV:
            mvp(vs.mvp, a, out);
            out->pvar[0] = a[6];
            out->pvar[1] = a[7];
            break;
        default:
            break;
        }
    }
};

/* Rasterizer whose list entry points tag every primitive with its draw-order key. */
class KeyedRasterizer : public swr::Rasterizer {
public:
    mutable uint32_t batch = 0;

    void drawTriangleList(const swr::RasterizerVertex *v, const int *idx, size_t n) const override
    {
        for (size_t i = 0; i + 3 <= n; i += 3) {
            g_key = (batch << 15) + (uint32_t)(i / 3);
            swr::Rasterizer::drawTriangleList(v, idx + i, 3);
        }
        batch++;
    }
    void drawLineList(const swr::RasterizerVertex *v, const int *idx, size_t n) const override
    {
        for (size_t i = 0; i + 2 <= n; i += 2) {
            g_key = (batch << 15) + (uint32_t)(i / 2);
            swr::Rasterizer::drawLineList(v, idx + i, 2);
        }
        batch++;
    }
    void drawPointList(const swr::RasterizerVertex *v, const int *idx, size_t n) const override
    {
        for (size_t i = 0; i < n; ++i) {
            g_key = (batch << 15) + (uint32_t)i;
            swr::Rasterizer::drawPointList(v, idx + i, 1);
        }
        batch++;
    }
};

/* ---- pixel shader instantiations ---------------------------------------------------- */
typedef void (*set_ps_fn)(swr::Rasterizer &);
template <bool Z, bool W, int A, int P> void set_ps(swr::Rasterizer &r)
{
    r.setPixelShader<ProgramShader<Z, W, A, P>>();
}

struct PsEntry { int a, p; set_ps_fn fn[4]; };
#define PS_ROW(A, P) {A, P, {set_ps<false, false, A, P>, set_ps<false, true, A, P>, \
                             set_ps<true, false, A, P>, set_ps<true, true, A, P>}}
const PsEntry k_ps_table[] = {
    PS_ROW(0, 0), PS_ROW(3, 0), PS_ROW(0, 2), PS_ROW(3, 2), PS_ROW(1, 1), PS_ROW(2, 3),
    PS_ROW(4, 4), PS_ROW(5, 0), PS_ROW(0, 5), PS_ROW(16, 16), PS_ROW(4, 3), PS_ROW(2, 2),
};

bool bind_pixel_shader(swr::Rasterizer &r, const swr_pixel_shader_desc &ps)
{
    int zw = (ps.interpolate_z ? 2 : 0) + (ps.interpolate_w ? 1 : 0);
    for (const PsEntry &e : k_ps_table)
        if (e.a == ps.avar_count && e.p == ps.pvar_count) {
            e.fn[zw](r);
            return true;
        }
    return false;
}

/* ---- vertex shader instantiations (AttribCount is 1 for every built-in program) ------- */
typedef void (*set_vs_fn)(swr::VertexProcessor &);
template <int A, int P> void set_vs(swr::VertexProcessor &v)
{
    v.setVertexShader<ProgramVertexShader<1, A, P>>();
}
template <int A, int... Ps> struct VsRow {
    static constexpr set_vs_fn fn[sizeof...(Ps)] = {set_vs<A, Ps>...};
};
template <int... As> struct VsTable {
    static constexpr const set_vs_fn *rows[sizeof...(As)] = {
        VsRow<As, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16>::fn...};
};
using VsAll = VsTable<0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16>;

bool bind_vertex_shader(swr::VertexProcessor &v, const swr_vertex_shader_desc &vs)
{
    if (vs.avar_count < 0 || vs.avar_count > 16 || vs.pvar_count < 0 || vs.pvar_count > 16)
        return false;
    if (vs.program == SWR_VS_DUMMY) {
        v.setVertexShader<swr::DummyVertexShader>();
        return true;
    }
    VsAll::rows[vs.avar_count][vs.pvar_count](v);
    return true;
}

int setup(KeyedRasterizer &r, const swr_pipeline_state *st)
{
    g_state = st;
    r.setRasterMode(static_cast<swr::RasterMode>(st->raster_mode));
    r.setScissorRect(st->scissor[0], st->scissor[1], st->scissor[2], st->scissor[3]);
    if (!bind_pixel_shader(r, st->ps))
        return SWR_ERR_UNSUPPORTED;
    return SWR_OK;
}

} // namespace

extern "C" int swr_ref_draw_elements(const swr_pipeline_state *st, int mode, size_t count,
                                     const int32_t *indices)
{
    if (!st || mode < 0 || mode > 2)
        return SWR_ERR_INVALID_ARG;
    for (size_t i = 0; i < count; ++i)
        if (indices[i] < 0)
            return SWR_ERR_INDEX_RANGE;
    KeyedRasterizer r;
    int rc = setup(r, st);
    if (rc)
        return rc;
    swr::VertexProcessor v(&r);
    v.setViewport(st->viewport[0], st->viewport[1], st->viewport[2], st->viewport[3]);
    v.setDepthRange(st->depth_near, st->depth_far);
    v.setCullMode(static_cast<swr::CullMode>(st->cull_mode));
    if (!bind_vertex_shader(v, st->vs))
        return SWR_ERR_UNSUPPORTED;
    for (int i = 0; i < SWR_MAX_VERTEX_ATTRIBS; ++i)
        if (st->attrib[i])
            v.setVertexAttribPointer(i, st->stride[i], st->attrib[i]);
    size_t factor = mode == 2 ? 3 : (mode == 1 ? 2 : 1);
    size_t usable = count - count % factor;
    v.drawElements(static_cast<swr::DrawMode>(mode), usable, const_cast<int *>(indices));
    return SWR_OK;
}

extern "C" int swr_ref_draw_triangle_list(const swr_pipeline_state *st,
                                          const swr_rasterizer_vertex *v, size_t nv,
                                          const int32_t *idx, size_t n)
{
    static_assert(sizeof(swr_rasterizer_vertex) == sizeof(swr::RasterizerVertex), "layout");
    (void)nv;
    KeyedRasterizer r;
    int rc = setup(r, st);
    if (rc)
        return rc;
    r.drawTriangleList(reinterpret_cast<const swr::RasterizerVertex *>(v), idx, n);
    return SWR_OK;
}

extern "C" int swr_ref_draw_line_list(const swr_pipeline_state *st, const swr_rasterizer_vertex *v,
                                      size_t nv, const int32_t *idx, size_t n)
{
    (void)nv;
    KeyedRasterizer r;
    int rc = setup(r, st);
    if (rc)
        return rc;
    r.drawLineList(reinterpret_cast<const swr::RasterizerVertex *>(v), idx, n);
    return SWR_OK;
}

extern "C" int swr_ref_draw_point_list(const swr_pipeline_state *st, const swr_rasterizer_vertex *v,
                                       size_t nv, const int32_t *idx, size_t n)
{
    (void)nv;
    KeyedRasterizer r;
    int rc = setup(r, st);
    if (rc)
        return rc;
    r.drawPointList(reinterpret_cast<const swr::RasterizerVertex *>(v), idx, n);
    return SWR_OK;
}

/* TriangleEquations of the reference itself (TriangleEquations.h:47-71), same layout as
 * swr_oracle_triangle_setup. */
extern "C" int swr_ref_triangle_setup(const swr_rasterizer_vertex *v0,
                                      const swr_rasterizer_vertex *v1,
                                      const swr_rasterizer_vertex *v2, int na, int np, float *out)
{
    const auto *a = reinterpret_cast<const swr::RasterizerVertex *>(v0);
    const auto *b = reinterpret_cast<const swr::RasterizerVertex *>(v1);
    const auto *c = reinterpret_cast<const swr::RasterizerVertex *>(v2);
    swr::TriangleEquations q(*a, *b, *c, na, np);
    int k = 0;
    out[k++] = q.area2;
    const swr::EdgeEquation *E[3] = {&q.e0, &q.e1, &q.e2};
    for (auto *e : E) {
        out[k++] = e->a;
        out[k++] = e->b;
        out[k++] = e->c;
    }
    if (q.area2 <= 0)
        return 0;
    const swr::ParameterEquation *P[2] = {&q.z, &q.invw};
    for (auto *p : P) {
        out[k++] = p->a;
        out[k++] = p->b;
        out[k++] = p->c;
    }
    for (int i = 0; i < na; ++i) {
        out[k++] = q.avar[i].a;
        out[k++] = q.avar[i].b;
        out[k++] = q.avar[i].c;
    }
    for (int i = 0; i < np; ++i) {
        out[k++] = q.pvar[i].a;
        out[k++] = q.pvar[i].b;
        out[k++] = q.pvar[i].c;
    }
    return k;
}

/* PolyClipper of the reference itself (PolyClipper.cpp:31-81) on one triangle, with the
 * plane sequence of VertexProcessor.cpp:233-242. */
extern "C" int swr_ref_clip_triangle(const swr_rasterizer_vertex *v0,
                                     const swr_rasterizer_vertex *v1,
                                     const swr_rasterizer_vertex *v2, int na, int np,
                                     swr_rasterizer_vertex *out, int max_out)
{
    std::vector<swr::VertexShaderOutput> verts(3);
    std::memcpy(&verts[0], v0, sizeof(swr::VertexShaderOutput));
    std::memcpy(&verts[1], v1, sizeof(swr::VertexShaderOutput));
    std::memcpy(&verts[2], v2, sizeof(swr::VertexShaderOutput));
    auto mask = [](const swr::VertexShaderOutput &v) {
        int m = 0;
        if (v.w - v.x < 0) m |= 0x01;
        if (v.x + v.w < 0) m |= 0x02;
        if (v.w - v.y < 0) m |= 0x04;
        if (v.y + v.w < 0) m |= 0x08;
        if (v.w - v.z < 0) m |= 0x10;
        if (v.z + v.w < 0) m |= 0x20;
        return m;
    };
    int cm = mask(verts[0]) | mask(verts[1]) | mask(verts[2]);
    swr::PolyClipper clipper;
    clipper.init(&verts, 0, 1, 2, na, np);
    if (cm & 0x01) clipper.clipToPlane(-1, 0, 0, 1);
    if (cm & 0x02) clipper.clipToPlane(1, 0, 0, 1);
    if (cm & 0x04) clipper.clipToPlane(0, -1, 0, 1);
    if (cm & 0x08) clipper.clipToPlane(0, 1, 0, 1);
    if (cm & 0x10) clipper.clipToPlane(0, 0, -1, 1);
    if (cm & 0x20) clipper.clipToPlane(0, 0, 1, 1);
    const std::vector<int> &ind = clipper.indices();
    int cnt = (int)ind.size();
    for (int k = 0; k < cnt && k < max_out; ++k) {
        /* only the fields the clipper wrote are defined; zero the rest as the oracle does */
        swr_rasterizer_vertex o;
        std::memset(&o, 0, sizeof o);
        const swr::VertexShaderOutput &s = verts[ind[k]];
        o.x = s.x;
        o.y = s.y;
        o.z = s.z;
        o.w = s.w;
        for (int i = 0; i < na; ++i)
            o.avar[i] = s.avar[i];
        for (int i = 0; i < np; ++i)
            o.pvar[i] = s.pvar[i];
        out[k] = o;
    }
    return cnt;
}
