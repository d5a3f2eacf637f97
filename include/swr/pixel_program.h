/*
 * swr/pixel_program.h — pixel programs compiled outside libswr_hip.so (hipcc, gfx950 only).
 *
 * The reference's pixel shaders are classes with four static constants, InterpolateZ,
 * InterpolateW, AVarCount and PVarCount, and a static drawPixel(const PixelData &) that the
 * rasterizer calls for every covered pixel (PixelShaderBase.h:36-137, Rasterizer.h:90-96).
 * The same class, with drawPixel a __device__ function that also receives the program's
 * uniforms, compiles into a module of its own:
 *
 *     #include <swr/pixel_program.h>
 *     struct Tint : swr::PixelProgramBase<Tint> {
 *         static const int InterpolateZ = true;
 *         static const int AVarCount = 3;
 *         __device__ static void drawPixel(const swr::PixelData &p, const void *uniforms)
 *         {
 *             const TintUniforms &u = *static_cast<const TintUniforms *>(uniforms);
 *             ...  // read / write this pixel's memory, e.g. u.color[p.y * u.pitch + p.x]
 *         }
 *     };
 *     SWR_PIXEL_MODULE(tint_module, Tint)
 *
 * built with the library's flags (softwarerenderer_amd.build.build_pixel_module runs them):
 *     hipcc -O3 -std=c++17 -fPIC -shared --offload-arch=gfx950 -ffp-contract=off
 *           -fno-fast-math -fhip-fp32-correctly-rounded-divide-sqrt
 *           -Xclang -target-feature -Xclang -packed-fp32-ops -I<repo>/include tint.hip
 * The entry point `int tint_module(swr_pixel_module *out)` fills the module descriptor that
 * swr_set_pixel_shader takes with program SWR_PS_MODULE (include/swr_hip.h).
 *
 * Order and values are the reference's: the module instantiates the library's own raster
 * kernels (one wave per 8x8 block, lane = pixel), so each pixel sees its fragments in draw
 * order, with z, 1/w, the affine varyings and pvar * w produced by the reference's
 * incremental chains bit for bit (PixelShaderBase.h:55-112, PixelData.h:51-125).  A pixel
 * belongs to one lane for the whole draw: a program that reads and writes only its own
 * pixel's memory sees its earlier fragments' writes, as in the reference.  Writes to other
 * pixels' memory race, as they do in the reference's OpenMP build.
 */
#ifndef SWR_PIXEL_PROGRAM_H
#define SWR_PIXEL_PROGRAM_H

#include "../swr_hip.h"
#include "../../softwarerenderer_amd/csrc/swr_lines_kernel.h"
#include "../../softwarerenderer_amd/csrc/swr_raster_kernel.h"

namespace swr {

/* PixelData, PixelData.h:34-49: what drawPixel receives.  For triangles `planes` points at
 * the primitive's plane equations (a, b, c per value) for computePerspectiveDerivatives;
 * points and lines have none (null). */
struct PixelData {
    int x, y;
    float z, w, invw;
    float avar[SWR_MAX_AVARS];
    float pvar[SWR_MAX_PVARS];
    const float *planes_invw;   /* the 1/w plane, or null  */
    const float *planes_pvar;   /* pvar plane 0 (plane i at planes_pvar + 3 i), or null */

    /* PixelData::computePerspectiveDerivatives, PixelData.h:127-145: d(pvar)/dx and
     * d(pvar)/dy at the pixel centre by the quotient rule, in the reference's operations. */
    __device__ void computePerspectiveDerivatives(int varIndex, float &ddx, float &ddy) const
    {
        const float *pl = planes_pvar + 3 * varIndex, *iw = planes_invw;
        const float xf = (float)x + 0.5f, yf = (float)y + 0.5f;
        const float var = pl[0] * xf + pl[1] * yf + pl[2];
        const float cur = iw[0] * xf + iw[1] * yf + iw[2];
        ddx = (cur * pl[0] - var * iw[0]) / (cur * cur);
        ddy = (cur * pl[1] - var * iw[1]) / (cur * cur);
    }
};

/* The statics of PixelShaderBase.h:44-53 with their defaults. */
template <class Derived>
struct PixelProgramBase {
    static const int InterpolateZ = false;
    static const int InterpolateW = false;
    static const int AVarCount = 0;
    static const int PVarCount = 0;
};

/* A shader class as a program of the raster kernels: every fragment goes straight to
 * drawPixel; nothing is kept per pixel between fragments (the program's memory is its own). */
template <class Shader>
struct ModuleProgram {
    static_assert(Shader::AVarCount >= 0 && Shader::AVarCount <= SWR_MAX_AVARS, "AVarCount");
    static_assert(Shader::PVarCount >= 0 && Shader::PVarCount <= SWR_MAX_PVARS, "PVarCount");
    static constexpr int NZ = Shader::InterpolateZ ? 1 : 0;
    // triangles interpolate 1/w for the perspective varyings too (PixelData.h:61-65)
    static constexpr int NW = (Shader::InterpolateW || Shader::PVarCount > 0) ? 1 : 0;
    static constexpr int NA = Shader::AVarCount, NP = Shader::PVarCount;

    template <class L>
    struct Prog {
        static constexpr int KEY_WORD = REC_KEY;
        static constexpr bool PLANES = true;
        struct State {};
        __device__ static void load(State &, const RasterArgs &, size_t) {}
        template <class PX>
        __device__ static void shade(State &, const RasterArgs &a, const PX &v, uint32_t, int x,
                                     int y)
        {
            PixelData p;
            p.x = x;
            p.y = y;
            p.z = NZ ? v.z() : 0.0f;
            // triangles: PixelData::init / stepX (w = 1 / invw); points and lines:
            // pixelDataFromVertex (Rasterizer.h:160-173: w as stepped, invw = 1 / w, and
            // only with InterpolateW)
            const bool has_w = PX::REPLAY ? NW != 0 : (Shader::InterpolateW != 0);
            p.w = has_w ? v.w() : 0.0f;
            p.invw = has_w ? v.invw() : 0.0f;
#pragma unroll
            for (int k = 0; k < SWR_MAX_AVARS; ++k)
                p.avar[k] = 0.0f;
#pragma unroll
            for (int k = 0; k < SWR_MAX_PVARS; ++k)
                p.pvar[k] = 0.0f;
            if constexpr (NA > 0) {
#pragma unroll
                for (int k = 0; k < NA; ++k)
                    p.avar[k] = v.avar(k);
            }
            if constexpr (NP > 0) {
#pragma unroll
                for (int k = 0; k < NP; ++k)
                    p.pvar[k] = v.pvar(k, p.w);
            }
            const float *rec =
                PX::REPLAY ? a.prims + (size_t)v.prim * (size_t)a.stride : nullptr;
            p.planes_invw = rec && NW ? rec + REC_PARAM + 3 * L::IW : nullptr;
            p.planes_pvar = rec && NP ? rec + REC_PARAM + 3 * L::P : nullptr;
            Shader::drawPixel(p, a.user);
        }
        template <class PX>
        __device__ static void store(const State &, const RasterArgs &, size_t, int, int) {}
    };

    static int launch_triangles(const void *args, void *stream, int waves)
    {
        return (int)launch<Prog, NZ, NW, NA, NP>(*static_cast<const RasterArgs *>(args),
                                                 static_cast<hipStream_t>(stream), waves);
    }
    static int launch_fragments(const void *args, void *stream)
    {
        return (int)launch_lr<Prog, NZ, NW, NA, NP>(*static_cast<const RasterArgs *>(args),
                                                    static_cast<hipStream_t>(stream));
    }
    static int fill(swr_pixel_module *m, const char *name)
    {
        if (!m)
            return SWR_ERR_INVALID_ARG;
        m->abi = SWR_PIXEL_MODULE_ABI;
        m->args_bytes = (uint32_t)sizeof(RasterArgs);
        m->interpolate_z = Shader::InterpolateZ ? 1 : 0;
        m->interpolate_w = Shader::InterpolateW ? 1 : 0;
        m->avar_count = Shader::AVarCount;
        m->pvar_count = Shader::PVarCount;
        m->launch_triangles = &launch_triangles;
        m->launch_fragments = &launch_fragments;
        m->name = name;
        return SWR_OK;
    }
};

}  // namespace swr

/* The module's entry point: extern "C" int entry(swr_pixel_module *out). */
#define SWR_PIXEL_MODULE(entry, Shader)                                                    \
    extern "C" __attribute__((visibility("default"))) int entry(swr_pixel_module *out)   \
    {                                                                                      \
        return ::swr::ModuleProgram<Shader>::fill(out, #Shader);                           \
    }

#endif /* SWR_PIXEL_PROGRAM_H */
