// Vertex shaders written the way a reference user writes one: the MVP transform of the
// examples (clip = M * (p, 1), colour as affine varyings, uv as perspective varyings), the
// matrix a uniform in device memory.  MvpInterleaved reads one attribute buffer of
// {px, py, pz, r, g, b, u, v}; MvpSplit reads the position from attribute 0 and
// {r, g, b, u, v} from attribute 1.  The dot products are sequential, as the reference's
// vector_math.h:55-60 (no contraction: the library's flags).
#include <swr/vertex_program.h>

struct MvpUniforms {
    float m[16];  // row-major
};

__device__ inline void mvp(const float *m, const float *p, swr::VertexShaderOutput *out)
{
    float r[4];
    for (int i = 0; i < 4; ++i) {
        float s = 0.0f;
        s = s + m[4 * i + 0] * p[0];
        s = s + m[4 * i + 1] * p[1];
        s = s + m[4 * i + 2] * p[2];
        s = s + m[4 * i + 3] * 1.0f;
        r[i] = s;
    }
    out->x = r[0];
    out->y = r[1];
    out->z = r[2];
    out->w = r[3];
}

struct MvpInterleaved : swr::VertexProgramBase<MvpInterleaved> {
    static const int AttribCount = 1;
    static const int AVarCount = 3;
    static const int PVarCount = 2;
    __device__ static void processVertex(swr::VertexShaderInput in, swr::VertexShaderOutput *out,
                                         const void *uniforms)
    {
        const MvpUniforms &u = *static_cast<const MvpUniforms *>(uniforms);
        const float *v = static_cast<const float *>(in[0]);
        mvp(u.m, v, out);
        out->avar[0] = v[3];
        out->avar[1] = v[4];
        out->avar[2] = v[5];
        out->pvar[0] = v[6];
        out->pvar[1] = v[7];
    }
};

struct MvpSplit : swr::VertexProgramBase<MvpSplit> {
    static const int AttribCount = 2;
    static const int AVarCount = 3;
    static const int PVarCount = 2;
    __device__ static void processVertex(swr::VertexShaderInput in, swr::VertexShaderOutput *out,
                                         const void *uniforms)
    {
        const MvpUniforms &u = *static_cast<const MvpUniforms *>(uniforms);
        const float *p = static_cast<const float *>(in[0]);
        const float *c = static_cast<const float *>(in[1]);
        mvp(u.m, p, out);
        for (int k = 0; k < 3; ++k)
            out->avar[k] = c[k];
        out->pvar[0] = c[3];
        out->pvar[1] = c[4];
    }
};

SWR_VERTEX_MODULE(swr_vmodule_mvp_interleaved, MvpInterleaved)
SWR_VERTEX_MODULE(swr_vmodule_mvp_split, MvpSplit)
