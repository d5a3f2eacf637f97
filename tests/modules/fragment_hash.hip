// Pixel programs compiled outside the library, for the module tests: per pixel, an
// order-sensitive hash of every fragment's PixelData (x, y and the interpolated values) and
// the invocation count, read and written at the pixel as a reference shader would.
#include <swr/pixel_program.h>

struct HashUniforms {
    uint32_t *hash;
    uint32_t *count;
    int32_t pitch;
};

__device__ inline uint32_t hmix(uint32_t h, uint32_t x) { return (h ^ x) * 16777619u; }

template <class Derived>
__device__ inline void hash_pixel(const swr::PixelData &p, const void *uniforms)
{
    const HashUniforms &u = *static_cast<const HashUniforms *>(uniforms);
    const size_t o = (size_t)p.y * (size_t)u.pitch + (size_t)p.x;
    uint32_t h = u.hash[o];
    h = hmix(h, (uint32_t)p.x);
    h = hmix(h, (uint32_t)p.y);
    if (Derived::InterpolateZ)
        h = hmix(h, __float_as_uint(p.z));
    if (Derived::InterpolateW) {
        h = hmix(h, __float_as_uint(p.w));
        h = hmix(h, __float_as_uint(p.invw));
    }
    for (int k = 0; k < Derived::AVarCount; ++k)
        h = hmix(h, __float_as_uint(p.avar[k]));
    for (int k = 0; k < Derived::PVarCount; ++k)
        h = hmix(h, __float_as_uint(p.pvar[k]));
    u.hash[o] = h;
    u.count[o] += 1u;
}

// z, w, 3 affine and 2 perspective varyings (the sphere's layout)
struct HashZW32 : swr::PixelProgramBase<HashZW32> {
    static const int InterpolateZ = true;
    static const int InterpolateW = true;
    static const int AVarCount = 3;
    static const int PVarCount = 2;
    __device__ static void drawPixel(const swr::PixelData &p, const void *u)
    {
        hash_pixel<HashZW32>(p, u);
    }
};

// perspective varyings without InterpolateW (1/w is still interpolated for them)
struct HashP4 : swr::PixelProgramBase<HashP4> {
    static const int PVarCount = 4;
    __device__ static void drawPixel(const swr::PixelData &p, const void *u)
    {
        hash_pixel<HashP4>(p, u);
    }
};

SWR_PIXEL_MODULE(swr_module_hash_zw32, HashZW32)
SWR_PIXEL_MODULE(swr_module_hash_p4, HashP4)
