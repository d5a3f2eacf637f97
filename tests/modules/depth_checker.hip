// The depth-tested checkerboard shader of the library's SWR_PS_DEPTH_CHECKER, written the way
// a reference user writes a shader: the test and the colour per fragment, at the pixel.
#include <swr/pixel_program.h>

struct CheckerUniforms {
    uint32_t *color;
    float *depth;
    int32_t pitch;
    float cells;
};

struct DepthChecker : swr::PixelProgramBase<DepthChecker> {
    static const int InterpolateZ = true;
    static const int InterpolateW = true;
    static const int AVarCount = 3;
    static const int PVarCount = 2;
    __device__ static void drawPixel(const swr::PixelData &p, const void *uniforms)
    {
        const CheckerUniforms &u = *static_cast<const CheckerUniforms *>(uniforms);
        const size_t o = (size_t)p.y * (size_t)u.pitch + (size_t)p.x;
        if (!(p.z < u.depth[o]))
            return;
        u.depth[o] = p.z;
        const int cu = swr::f2i(floorf(p.pvar[0] * u.cells));
        const int cv = swr::f2i(floorf(p.pvar[1] * u.cells));
        const float k = (((uint32_t)cu + (uint32_t)cv) & 1u) ? 1.0f : 0.5f;
        const uint32_t r = (uint32_t)(swr::clamp01(p.avar[0] * k) * 255.0f);
        const uint32_t g = (uint32_t)(swr::clamp01(p.avar[1] * k) * 255.0f);
        const uint32_t b = (uint32_t)(swr::clamp01(p.avar[2] * k) * 255.0f);
        u.color[o] = r | (g << 8) | (b << 16) | 0xFF000000u;
    }
};

SWR_PIXEL_MODULE(swr_module_depth_checker, DepthChecker)
