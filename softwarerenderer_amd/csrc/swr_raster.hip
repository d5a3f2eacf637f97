// swr_raster.hip — the raster stage's launchers for the built-in pixel programs (the kernel
// is in swr_raster_kernel.h) and the record layout each program expects.
#include "swr_internal.h"
#include "swr_raster_kernel.h"

namespace swr {

// The programs the tuner may give another group size (the benchmark and visibility ones).
template <template <class> class ProgT, int NZ, int NW, int NA, int NP>
static hipError_t launch_tuned(const RasterArgs &a, hipStream_t s, int waves, int group)
{
    switch (group) {
    case 1: return launch<ProgT, NZ, NW, NA, NP, 1>(a, s, waves);
    case 2: return launch<ProgT, NZ, NW, NA, NP, 2>(a, s, waves);
    case 4: return launch<ProgT, NZ, NW, NA, NP, 4>(a, s, waves);
    default: return launch<ProgT, NZ, NW, NA, NP>(a, s, waves);
    }
}

// The plane layout each program's kernel expects (see ParamLayout).
bool raster_layout(const swr_pixel_shader_desc &ps, ParamLayout &L)
{
    L.nz = L.nw = L.na = L.np = 0;
    L.set_na = L.set_np = 0;
    switch (ps.program) {
    case SWR_PS_NULL:
    case SWR_PS_COVERAGE:
        break;
    case SWR_PS_RGB:
        if (ps.avar_count < 3)
            return false;
        L.na = L.set_na = 3;
        break;
    case SWR_PS_DEPTH_CHECKER:
        if (!ps.interpolate_z || ps.avar_count < 3 || ps.pvar_count < 2)
            return false;
        L.nz = 1;
        L.nw = 1;
        L.na = L.set_na = 3;
        L.np = L.set_np = 2;
        break;
    case SWR_PS_VISIBILITY:
        if (!ps.interpolate_z)
            return false;
        L.nz = 1;
        break;
    case SWR_PS_TEXTURE:
        if (ps.pvar_count < 2)
            return false;
        L.nw = 1;
        L.np = L.set_np = 2;
        break;
    case SWR_PS_MODULE:
        // the program's own counts (PixelData::init interpolates 1/w for the perspective
        // varyings too, PixelData.h:61-65)
        if (ps.avar_count < 0 || ps.avar_count > SWR_MAX_AVARS || ps.pvar_count < 0 ||
            ps.pvar_count > SWR_MAX_PVARS)
            return false;
        L.nz = ps.interpolate_z ? 1 : 0;
        L.nw = (ps.interpolate_w || ps.pvar_count > 0) ? 1 : 0;
        L.na = L.set_na = ps.avar_count;
        L.np = L.set_np = ps.pvar_count;
        break;
    case SWR_PS_HASH:
    case SWR_PS_LOG: {
        if (ps.avar_count < 0 || ps.avar_count > 16 || ps.pvar_count < 0 || ps.pvar_count > 16)
            return false;
        int cls = (ps.avar_count <= 4 && ps.pvar_count <= 4) ? 4 : 16;
        L.nz = 1;
        L.nw = 1;
        L.na = L.np = cls;
        L.set_na = ps.avar_count;
        L.set_np = ps.pvar_count;
        break;
    }
    default:
        return false;
    }
    L.count = L.nz + L.nw + L.na + L.np;
    L.stride = (REC_PARAM + 3 * L.count + 3) & ~3;
    return true;
}

hipError_t launch_raster(const swr_pixel_shader_desc &ps, const ParamLayout &L,
                         const RasterArgs &a, hipStream_t s, int waves, int group)
{
    if (waves != 1 && waves != 2 && waves != 4 && waves != 8)
        waves = 4;
    switch (ps.program) {
    case SWR_PS_COVERAGE:
        return launch<ProgCoverage, 0, 0, 0, 0>(a, s, waves);
    case SWR_PS_RGB:
        return launch<ProgRGB, 0, 0, 3, 0>(a, s, waves);
    case SWR_PS_DEPTH_CHECKER:
        return launch_tuned<ProgDepthChecker, 1, 1, 3, 2>(a, s, waves, group);
    case SWR_PS_VISIBILITY:
        return launch_tuned<ProgVisibility, 1, 0, 0, 0>(a, s, waves, group);
    case SWR_PS_TEXTURE:
        return launch<ProgTexture, 0, 1, 0, 2>(a, s, waves);
    case SWR_PS_HASH:
        return L.na == 4 ? launch<ProgHash, 1, 1, 4, 4>(a, s, waves)
                         : launch<ProgHash, 1, 1, 16, 16>(a, s, waves);
    case SWR_PS_LOG:
        return L.na == 4 ? launch<ProgLog, 1, 1, 4, 4>(a, s, waves)
                         : launch<ProgLog, 1, 1, 16, 16>(a, s, waves);
    case SWR_PS_MODULE:
        return (hipError_t)ps.module->launch_triangles(&a, s, waves);
    default:
        return hipSuccess;
    }
}

}  // namespace swr
