/*
 * swr/vertex_program.h — vertex programs compiled outside libswr_hip.so (hipcc, gfx950 only).
 *
 * The reference's vertex shaders are classes with three static constants, AttribCount,
 * AVarCount and PVarCount, and a static processVertex(VertexShaderInput, VertexShaderOutput *)
 * that VertexProcessor calls once per referenced vertex (VertexShaderBase.h:33-51,
 * VertexProcessor.cpp:86-117, 134-150).  The same class, with processVertex a __device__
 * function that also receives the program's uniforms, compiles into a module of its own:
 *
 *     #include <swr/vertex_program.h>
 *     struct Wave : swr::VertexProgramBase<Wave> {
 *         static const int AttribCount = 2;
 *         static const int AVarCount = 3;
 *         __device__ static void processVertex(swr::VertexShaderInput in,
 *                                              swr::VertexShaderOutput *out,
 *                                              const void *uniforms)
 *         {
 *             const float *p = static_cast<const float *>(in[0]);   // attribute 0
 *             const float *c = static_cast<const float *>(in[1]);   // attribute 1
 *             ...  // out->x, y, z, w (clip space), out->avar[0..2]
 *         }
 *     };
 *     SWR_VERTEX_MODULE(wave_module, Wave)
 *
 * built with the library's flags (softwarerenderer_amd.build.build_vertex_module runs them).
 * The entry point `int wave_module(swr_vertex_module *out)` fills the module descriptor that
 * swr_set_vertex_shader takes with program SWR_VS_MODULE (include/swr_hip.h).  The library
 * runs processVertex once per vertex of the attribute buffers and rasterizes the outputs with
 * the reference's clipping, culling and draw order, so results are the reference's for any
 * program that is a pure function of its inputs and uniforms (the reference's VertexCache
 * already relies on that).
 */
#ifndef SWR_VERTEX_PROGRAM_H
#define SWR_VERTEX_PROGRAM_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../swr_hip.h"

namespace swr {

/* VertexConfig.h:37-40: one pointer per vertex attribute to this vertex's element, and the
 * output vertex (RasterizerVertex, IRasterizer.h:42-53). */
typedef const void *VertexShaderInput[SWR_MAX_VERTEX_ATTRIBS];
typedef swr_rasterizer_vertex VertexShaderOutput;

/* The statics of VertexShaderBase.h:37-43 with their defaults. */
template <class Derived>
struct VertexProgramBase {
    static const int AttribCount = 0;
    static const int AVarCount = 0;
    static const int PVarCount = 0;
};

struct VertexModuleArgs {
    const char *attr[SWR_MAX_VERTEX_ATTRIBS];
    int64_t stride[SWR_MAX_VERTEX_ATTRIBS];
    uint32_t count;
    float *out;
    const void *uniforms;
};

/* One thread per vertex: the attribute pointers of VertexProcessor::initVertexInput
 * (VertexProcessor.cpp:140-150), processVertex, then the output row {x, y, z, w, avar[A],
 * pvar[P]}.  Outputs the program leaves unwritten are 0 (the reference leaves them
 * uninitialised). */
template <class Shader>
__global__ void __launch_bounds__(256) k_vertex_module(VertexModuleArgs a)
{
    constexpr int NA = Shader::AVarCount, NP = Shader::PVarCount;
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.count)
        return;
    VertexShaderInput in;
#pragma unroll
    for (int i = 0; i < SWR_MAX_VERTEX_ATTRIBS; ++i)
        in[i] = i < Shader::AttribCount ? a.attr[i] + a.stride[i] * (int64_t)v : nullptr;
    VertexShaderOutput o;
    o.x = o.y = o.z = o.w = 0.0f;
#pragma unroll
    for (int k = 0; k < SWR_MAX_AVARS; ++k)
        o.avar[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < SWR_MAX_PVARS; ++k)
        o.pvar[k] = 0.0f;
    Shader::processVertex(in, &o, a.uniforms);
    float *row = a.out + (size_t)v * (size_t)(4 + NA + NP);
    row[0] = o.x;
    row[1] = o.y;
    row[2] = o.z;
    row[3] = o.w;
#pragma unroll
    for (int k = 0; k < NA; ++k)
        row[4 + k] = o.avar[k];
#pragma unroll
    for (int k = 0; k < NP; ++k)
        row[4 + NA + k] = o.pvar[k];
}

template <class Shader>
struct VertexModuleProgram {
    static_assert(Shader::AttribCount >= 1 && Shader::AttribCount <= SWR_MAX_VERTEX_ATTRIBS,
                  "AttribCount");
    static_assert(Shader::AVarCount >= 0 && Shader::AVarCount <= SWR_MAX_AVARS, "AVarCount");
    static_assert(Shader::PVarCount >= 0 && Shader::PVarCount <= SWR_MAX_PVARS, "PVarCount");

    static int shade(const void *const *attribs, const int64_t *strides, uint32_t count,
                     float *out, const void *uniforms, void *stream)
    {
        if (!count)
            return (int)hipSuccess;
        VertexModuleArgs a;
        for (int i = 0; i < SWR_MAX_VERTEX_ATTRIBS; ++i) {
            a.attr[i] =
                i < Shader::AttribCount ? static_cast<const char *>(attribs[i]) : nullptr;
            a.stride[i] = i < Shader::AttribCount ? strides[i] : 0;
        }
        a.count = count;
        a.out = out;
        a.uniforms = uniforms;
        hipLaunchKernelGGL(k_vertex_module<Shader>, dim3((count + 255) / 256), dim3(256), 0,
                           static_cast<hipStream_t>(stream), a);
        return (int)hipGetLastError();
    }
    static int fill(swr_vertex_module *m, const char *name)
    {
        if (!m)
            return SWR_ERR_INVALID_ARG;
        m->abi = SWR_VERTEX_MODULE_ABI;
        m->attrib_count = Shader::AttribCount;
        m->avar_count = Shader::AVarCount;
        m->pvar_count = Shader::PVarCount;
        m->shade = &shade;
        m->name = name;
        return SWR_OK;
    }
};

}  // namespace swr

/* The module's entry point: extern "C" int entry(swr_vertex_module *out). */
#define SWR_VERTEX_MODULE(entry, Shader)                                                   \
    extern "C" __attribute__((visibility("default"))) int entry(swr_vertex_module *out)  \
    {                                                                                      \
        return ::swr::VertexModuleProgram<Shader>::fill(out, #Shader);                     \
    }

#endif /* SWR_VERTEX_PROGRAM_H */
