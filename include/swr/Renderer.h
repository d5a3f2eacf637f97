/*
 * swr/Renderer.h — the reference's C++ host API (Renderer.h:27-28 umbrella) over the gfx950
 * C ABI (swr_hip.h).  Header-only; link with libswr_hip.so.
 *
 * Names, signatures, defaults and static shader contracts follow the reference:
 *   Rasterizer       Rasterizer.h:52-141      (setRasterMode, setScissorRect, setPixelShader<T>,
 *                                             drawPoint/Line/Triangle, draw*List)
 *   VertexProcessor  VertexProcessor.h:56-91  (setViewport, setDepthRange, setCullMode,
 *                                             setVertexShader<T>, setVertexAttribPointer,
 *                                             drawElements)
 *   PixelShaderBase<T>  PixelShaderBase.h:40-53  statics InterpolateZ/InterpolateW/AVarCount/PVarCount
 *   VertexShaderBase<T> VertexShaderBase.h:33-43 statics AttribCount/AVarCount/PVarCount
 *
 * What changes for a shader class: its body runs on the GPU, so instead of a static
 * drawPixel / processVertex it names a device program (static const int Program) and
 * fills that program's uniforms and render targets in a static bind(desc) function,
 * which is called at every draw (the reference reads a shader's static members when its
 * drawPixel runs).  Errors throw swr::Error (the reference asserts).
 */
#ifndef SWR_RENDERER_HPP
#define SWR_RENDERER_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../swr_hip.h"

namespace swr {

const int BlockSize = SWR_BLOCK_SIZE;        /* IRasterizer.h:33 */
const int MaxAVars = SWR_MAX_AVARS;          /* IRasterizer.h:36 */
const int MaxPVars = SWR_MAX_PVARS;          /* IRasterizer.h:39 */
const int MaxVertexAttribs = SWR_MAX_VERTEX_ATTRIBS; /* VertexConfig.h:34 */

typedef swr_rasterizer_vertex RasterizerVertex;     /* IRasterizer.h:42-53 */

enum class RasterMode { Span = SWR_RASTER_SPAN, Block = SWR_RASTER_BLOCK, Adaptive = SWR_RASTER_ADAPTIVE };
enum class DrawMode { Point = SWR_DRAW_POINT, Line = SWR_DRAW_LINE, Triangle = SWR_DRAW_TRIANGLE };
enum class CullMode { None = SWR_CULL_NONE, CCW = SWR_CULL_CCW, CW = SWR_CULL_CW };

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
    int code;
};

/* PixelShaderBase<T>, PixelShaderBase.h:40-137 */
template <class Derived>
class PixelShaderBase {
public:
    static const int InterpolateZ = false;
    static const int InterpolateW = false;
    static const int AVarCount = 0;
    static const int PVarCount = 0;
    static const int Program = SWR_PS_NULL;
    static void bind(swr_pixel_shader_desc &) {}
};

/* NullPixelShader, PixelShaderBase.h:139 */
class NullPixelShader : public PixelShaderBase<NullPixelShader> {};

/* VertexShaderBase<T>, VertexShaderBase.h:33-51 */
template <class Derived>
class VertexShaderBase {
public:
    static const int AttribCount = 0;
    static const int AVarCount = 0;
    static const int PVarCount = 0;
    static const int Program = SWR_VS_DUMMY;
    static void bind(swr_vertex_shader_desc &) {}
};

/* DummyVertexShader, VertexShaderBase.h:53 */
class DummyVertexShader : public VertexShaderBase<DummyVertexShader> {};

/* IRasterizer, IRasterizer.h:56-71 */
class IRasterizer {
public:
    virtual ~IRasterizer() {}
    virtual void drawPointList(const RasterizerVertex *vertices, const int *indices, size_t indexCount) const = 0;
    virtual void drawLineList(const RasterizerVertex *vertices, const int *indices, size_t indexCount) const = 0;
    virtual void drawTriangleList(const RasterizerVertex *vertices, const int *indices, size_t indexCount) const = 0;
};

namespace detail {
inline void check(swr_context *ctx, int rc, const char *what)
{
    if (rc != SWR_OK)
        throw Error(rc, std::string(what) + ": " + (ctx ? swr_last_error(ctx) : "no context"));
}

template <class PS>
void describe_pixel_shader(swr_pixel_shader_desc &d)
{
    d = swr_pixel_shader_desc();
    d.program = PS::Program;
    d.interpolate_z = PS::InterpolateZ ? 1 : 0;
    d.interpolate_w = PS::InterpolateW ? 1 : 0;
    d.avar_count = PS::AVarCount;
    d.pvar_count = PS::PVarCount;
    d.checker_cells = 8.0f;
    PS::bind(d);
}

template <class VS>
void describe_vertex_shader(swr_vertex_shader_desc &d)
{
    d = swr_vertex_shader_desc();
    d.program = VS::Program;
    d.attrib_count = VS::AttribCount;
    d.avar_count = VS::AVarCount;
    d.pvar_count = VS::PVarCount;
    for (int i = 0; i < 16; ++i)
        d.mvp[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    VS::bind(d);
}
}  // namespace detail

/* Rasterizer, Rasterizer.h:52-141; one gfx950 device context per instance. */
class Rasterizer : public IRasterizer {
public:
    /* Rasterizer.h:67-72: Span, scissor (0,0,0,0), NullPixelShader */
    explicit Rasterizer(int device = 0)
    {
        detail::check(nullptr, swr_create(&m_ctx, device), "swr_create");
        setPixelShader<NullPixelShader>();
    }
    ~Rasterizer() override { swr_destroy(m_ctx); }
    Rasterizer(const Rasterizer &) = delete;
    Rasterizer &operator=(const Rasterizer &) = delete;

    void setRasterMode(RasterMode mode)
    {
        detail::check(m_ctx, swr_set_raster_mode(m_ctx, (int)mode), "setRasterMode");
    }

    void setScissorRect(int x, int y, int width, int height)
    {
        detail::check(m_ctx, swr_set_scissor_rect(m_ctx, x, y, width, height), "setScissorRect");
    }

    template <class PixelShader>
    void setPixelShader()
    {
        m_describe = &detail::describe_pixel_shader<PixelShader>;
        bindShader();
    }

    void drawPoint(const RasterizerVertex &v) const { drawPointList(&v, kIota, 1); }
    void drawLine(const RasterizerVertex &v0, const RasterizerVertex &v1) const
    {
        RasterizerVertex v[2] = {v0, v1};
        drawLineList(v, kIota, 2);
    }
    void drawTriangle(const RasterizerVertex &v0, const RasterizerVertex &v1,
                      const RasterizerVertex &v2) const
    {
        RasterizerVertex v[3] = {v0, v1, v2};
        drawTriangleList(v, kIota, 3);
    }

    /* Lists: primitives whose first index is -1 are skipped.  vertex_count is not part of the
     * reference signature; it is derived from the largest index here. */
    void drawPointList(const RasterizerVertex *vertices, const int *indices, size_t n) const override
    {
        bindShader();
        detail::check(m_ctx, swr_draw_point_list(m_ctx, vertices, extent(indices, n), indices, n),
                      "drawPointList");
    }
    void drawLineList(const RasterizerVertex *vertices, const int *indices, size_t n) const override
    {
        bindShader();
        detail::check(m_ctx, swr_draw_line_list(m_ctx, vertices, extent(indices, n), indices, n),
                      "drawLineList");
    }
    void drawTriangleList(const RasterizerVertex *vertices, const int *indices, size_t n) const override
    {
        bindShader();
        detail::check(m_ctx, swr_draw_triangle_list(m_ctx, vertices, extent(indices, n), indices, n),
                      "drawTriangleList");
    }

    swr_context *context() const { return m_ctx; }
    void bindShader() const
    {
        swr_pixel_shader_desc d;
        m_describe(d);
        detail::check(m_ctx, swr_set_pixel_shader(m_ctx, &d), "setPixelShader");
    }

private:
    static size_t extent(const int *indices, size_t n)
    {
        int mx = -1;
        for (size_t i = 0; i < n; ++i)
            mx = indices[i] > mx ? indices[i] : mx;
        return (size_t)(mx + 1);
    }
    static constexpr int kIota[3] = {0, 1, 2};
    swr_context *m_ctx = nullptr;
    void (*m_describe)(swr_pixel_shader_desc &) = nullptr;
};

/* VertexProcessor, VertexProcessor.h:56-152.  It drives a swr::Rasterizer: geometry and
 * rasterization run as one GPU pipeline on the rasterizer's context. */
class VertexProcessor {
public:
    /* VertexProcessor.cpp:29-35: cull CW, depth range (0,1), DummyVertexShader */
    explicit VertexProcessor(IRasterizer *rasterizer)
    {
        setRasterizer(rasterizer);
        setCullMode(CullMode::CW);
        setDepthRange(0.0f, 1.0f);
        setVertexShader<DummyVertexShader>();
    }

    void setRasterizer(IRasterizer *rasterizer)
    {
        m_r = dynamic_cast<Rasterizer *>(rasterizer);
        if (!m_r)
            throw Error(SWR_ERR_INVALID_ARG, "the gfx950 VertexProcessor drives a swr::Rasterizer");
    }

    void setViewport(int x, int y, int width, int height)
    {
        detail::check(ctx(), swr_set_viewport(ctx(), x, y, width, height), "setViewport");
    }
    void setDepthRange(float n, float f)
    {
        detail::check(ctx(), swr_set_depth_range(ctx(), n, f), "setDepthRange");
    }
    void setCullMode(CullMode mode)
    {
        detail::check(ctx(), swr_set_cull_mode(ctx(), (int)mode), "setCullMode");
    }

    template <class VertexShader>
    void setVertexShader()
    {
        static_assert(VertexShader::AttribCount <= MaxVertexAttribs, "AttribCount");
        m_describe = &detail::describe_vertex_shader<VertexShader>;
        bindShader();
    }

    /* buffer may be host or device memory; its size is not part of the reference signature */
    void setVertexAttribPointer(int index, int stride, const void *buffer)
    {
        detail::check(ctx(), swr_set_vertex_attrib_pointer(ctx(), index, stride, buffer, 0),
                      "setVertexAttribPointer");
    }

    void drawElements(DrawMode mode, size_t count, int *indices) const
    {
        bindShader();
        m_r->bindShader();
        detail::check(ctx(), swr_draw_elements(ctx(), (int)mode, count, indices), "drawElements");
    }

private:
    swr_context *ctx() const { return m_r->context(); }
    void bindShader() const
    {
        swr_vertex_shader_desc d;
        m_describe(d);
        detail::check(ctx(), swr_set_vertex_shader(ctx(), &d), "setVertexShader");
    }
    Rasterizer *m_r = nullptr;
    void (*m_describe)(swr_vertex_shader_desc &) = nullptr;
};

}  // namespace swr

#endif
