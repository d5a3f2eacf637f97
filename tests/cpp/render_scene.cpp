// render_scene.cpp — drives the gfx950 path through the C++ host API the way the reference's
// examples drive theirs (README.md:19-70, Box.cpp:39-202): shader classes with statics, a
// Rasterizer, a VertexProcessor, one drawElements.
//
//   render_scene <scene.bin> <out.bin>
// scene.bin: int32 W, H, raster_mode, cull_mode, vertex_count, index_count, stride;
//            float mvp[16]; vertices (vertex_count * stride bytes); int32 indices
// out.bin:   uint32 color[H*W] (RGBA8), float depth[H*W]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "swr/Renderer.h"

using namespace swr;

class PixelShader : public PixelShaderBase<PixelShader> {
public:
    static const bool InterpolateZ = true;
    static const bool InterpolateW = true;
    static const int AVarCount = 3;
    static const int PVarCount = 2;
    static const int Program = SWR_PS_DEPTH_CHECKER;

    static uint32_t *color;
    static float *depth;
    static int width, height;

    static void bind(swr_pixel_shader_desc &d)
    {
        d.target0 = color;
        d.target1 = depth;
        d.width = width;
        d.height = height;
        d.pitch = width;
        d.checker_cells = 8.0f;
    }
};
uint32_t *PixelShader::color;
float *PixelShader::depth;
int PixelShader::width, PixelShader::height;

class VertexShader : public VertexShaderBase<VertexShader> {
public:
    static const int AttribCount = 1;
    static const int AVarCount = 3;
    static const int PVarCount = 2;
    static const int Program = SWR_VS_MVP_POS_RGB_UV;
    static float modelViewProjectionMatrix[16];

    static void bind(swr_vertex_shader_desc &d)
    {
        for (int i = 0; i < 16; ++i)
            d.mvp[i] = modelViewProjectionMatrix[i];
    }
};
float VertexShader::modelViewProjectionMatrix[16];

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: render_scene scene.bin out.bin\n");
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f)
        return 2;
    int hdr[7];
    if (std::fread(hdr, 4, 7, f) != 7)
        return 2;
    const int W = hdr[0], H = hdr[1], mode = hdr[2], cull = hdr[3];
    const size_t nv = (size_t)hdr[4], ni = (size_t)hdr[5];
    const int stride = hdr[6];
    if (std::fread(VertexShader::modelViewProjectionMatrix, 4, 16, f) != 16)
        return 2;
    std::vector<unsigned char> verts(nv * (size_t)stride);
    std::vector<int> idx(ni);
    if (std::fread(verts.data(), 1, verts.size(), f) != verts.size() ||
        std::fread(idx.data(), 4, ni, f) != ni)
        return 2;
    std::fclose(f);

    std::vector<uint32_t> color((size_t)W * H, 0u);
    std::vector<float> depth((size_t)W * H, 1.0f);
    PixelShader::color = color.data();
    PixelShader::depth = depth.data();
    PixelShader::width = W;
    PixelShader::height = H;

    try {
        Rasterizer r;
        VertexProcessor v(&r);
        r.setRasterMode(static_cast<RasterMode>(mode));
        r.setScissorRect(0, 0, W, H);
        r.setPixelShader<PixelShader>();
        v.setViewport(0, 0, W, H);
        v.setCullMode(static_cast<CullMode>(cull));
        v.setVertexShader<VertexShader>();
        v.setVertexAttribPointer(0, stride, verts.data());
        v.drawElements(DrawMode::Triangle, idx.size(), idx.data());
    } catch (const Error &e) {
        std::fprintf(stderr, "swr error %d: %s\n", e.code, e.what());
        return 1;
    }

    f = std::fopen(argv[2], "wb");
    if (!f)
        return 2;
    std::fwrite(color.data(), 4, color.size(), f);
    std::fwrite(depth.data(), 4, depth.size(), f);
    std::fclose(f);
    std::printf("rendered %zu triangles at %dx%d\n", idx.size() / 3, W, H);
    return 0;
}
