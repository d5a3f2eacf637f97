// swr_internal.h — launchers shared between the kernel files and the C-ABI layer.
#pragma once

#include "swr_common.h"

namespace swr {

struct SortBuffers {
    uint32_t *keys_a, *vals_a, *keys_b, *vals_b;
    uint32_t *hist, *hist_ex, *scan_tmp;
    uint32_t *ranges;
    uint32_t blocks;
    const uint32_t *chunk_kept = nullptr;  // pairs kept per expansion chunk (nullptr: one run)
    uint32_t chunk = 0;                    // slots per expansion chunk
    uint32_t *sorted_keys, *sorted_vals;  // set by sort_pairs
};

// geometry (swr_geometry.hip)
hipError_t launch_begin(const GeomArgs &a, hipStream_t s);
hipError_t launch_geom_count(const GeomArgs &a, hipStream_t s);
hipError_t launch_geom_emit(const GeomArgs &a, hipStream_t s);
hipError_t launch_expand(const GeomArgs &a, hipStream_t s, int chunk);

int emit_class(int nvary);  // varyings carried per vertex: 0, 3, 5, 8 or 32

// points and lines (swr_lines.hip)
struct LineArgs {
    GeomArgs g;             // vertex input, VertexProcessor / Rasterizer state, layout, status,
                            // ranges, tiles
    int mode;               // SWR_DRAW_POINT or SWR_DRAW_LINE
    uint64_t prims;         // primitives: index pairs (lines) or indices (points)
    int iz, iw;             // InterpolateZ / InterpolateW: which values the DDA steps
    uint32_t *count;        // [prims] fragments of each primitive (count pass)
    const uint32_t *first;  // [prims] exclusive scan of count
    float *frags;           // fragment records (emit pass)
    int fstride;            // words per fragment record
    uint32_t *keys, *vals;  // (tile, fragment) pairs in fragment order
};
int fragment_stride(const ParamLayout &L);
hipError_t launch_lines_count(const LineArgs &a, hipStream_t s);
hipError_t launch_lines_emit(const LineArgs &a, hipStream_t s);
hipError_t launch_lines_raster(const swr_pixel_shader_desc &ps, const ParamLayout &L,
                               const RasterArgs &a, hipStream_t s);

// scans and sort (swr_sort.hip)
uint64_t scan_tmp_elems(uint64_t n);
hipError_t scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *tmp, hipStream_t s);
hipError_t scan_pairs(const uint2 *in, uint2 *out, uint64_t n, uint2 *tmp,
                      unsigned long long *status, uint64_t cap_prims, uint64_t cap_pairs,
                      hipStream_t s);
// the (appended fans, live) scan of the input triangles, compacting the live ones into
// `live`, and scan_pairs of the slot counts, in shared launches; tmp holds
// scan_tmp_elems(na) + scan_tmp_elems(ng) 8-byte elements
hipError_t scan_fans_and_pairs(const uint2 *app, uint2 *app_ex, uint64_t na, uint32_t *live,
                               const uint2 *grp, uint2 *grp_ex, uint64_t ng, void *tmp,
                               unsigned long long *status, uint64_t cap_prims,
                               uint64_t cap_pairs, hipStream_t s);
int radix_passes(uint32_t tiles);
hipError_t sort_pairs(SortBuffers &sb, const unsigned long long *status, uint32_t tiles,
                      hipStream_t s);
hipError_t launch_ranges(const SortBuffers &sb, const unsigned long long *status,
                         uint32_t tiles, hipStream_t s);
uint32_t pair_sentinel(uint32_t tiles);

// context access for the kernel files' own entry points (swr_api.cpp)
hipStream_t context_stream(swr_context *c);
// The context's second stream with a fork and a join event (false: none, one stream only).
bool context_side(swr_context *c, hipStream_t *side, hipEvent_t *fork, hipEvent_t *join);
int context_fail(swr_context *c, int code, const char *msg);
// The status code of the context's last failure (SWR_ERR_HIP if none was recorded).
int context_last_code(swr_context *c);
bool context_is_device_ptr(const void *p);
// Device buffer `slot` (0 or 1) of at least `bytes` owned by the context, zero when first
// handed out at this size.  nullptr (and the context's error set) on failure.
void *context_zeroed_scratch(swr_context *c, int slot, size_t bytes);
// Nine words the differentiation entry points keep in the context: what the last forward
// stored its antialiasing decisions for.
uint64_t *context_diff_tag(swr_context *c);
// Device copy of a differentiation scene table (`bytes` of host data), uploaded only when
// it differs from the last one (a synchronous, stream-ordered copy; refused while the
// context stream is being captured).  *gen changes with every upload.  nullptr (and the
// context's error set) on failure.
const void *context_diff_table(swr_context *c, const void *host, size_t bytes, uint64_t *gen);

// raster (swr_raster.hip)
bool raster_layout(const swr_pixel_shader_desc &ps, ParamLayout &L);
hipError_t launch_raster(const swr_pixel_shader_desc &ps, const ParamLayout &L,
                         const RasterArgs &a, hipStream_t s, int waves, int group);

}  // namespace swr
