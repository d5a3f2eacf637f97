// swr_api.cpp — the C ABI (include/swr_hip.h): context state, device arenas, and the
// launch sequence of one draw.
//
// Draw sequence for triangles (all on the context stream):
//   begin      clear tile ranges and the status words
//   count      geometry pass 1: per input triangle, fan size and (prims, pairs) per phase
//   scans      exclusive scans -> draw-order positions; totals + arena overflow published
//   emit       geometry pass 2: primitive records and (tile, primitive) pairs in draw order
//   sort       stable radix sort of the pairs by tile id
//   ranges     per-tile [begin, end)
//   raster     one wave64 per 8x8 tile
// A synchronous draw reads the status words back; when an arena overflowed, nothing was
// rasterized (every later kernel checks the overflow word first), so the arenas are grown
// to the published totals and the draw is issued again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "swr_common.h"
#include "swr_internal.h"

using namespace swr;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

enum Stage { S_GEOM, S_SCAN, S_EMIT, S_SORT, S_RANGES, S_RASTER, S_COUNT };

}  // namespace

struct swr_context {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;               // a second stream for independent kernels
    hipEvent_t fork = nullptr, join = nullptr;
    std::string err;
    int err_code = SWR_OK;                    // code of the last failure (with err)

    // Rasterizer (Rasterizer.h:67-72 defaults)
    int raster_mode = SWR_RASTER_SPAN;
    int sc[4] = {0, 0, 0, 0};
    swr_pixel_shader_desc ps{};
    swr_pixel_module module{};                // SWR_PS_MODULE: the bound module (ps.module)

    // VertexProcessor (VertexProcessor.cpp:29-35 defaults)
    int vp[4] = {0, 0, 0, 0};
    float dn = 0.0f, df = 1.0f;
    int cull = SWR_CULL_CW;
    swr_vertex_shader_desc vs{};
    swr_vertex_module vmodule{};              // SWR_VS_MODULE: the bound module (vs.module)
    struct Attr {
        const void *ptr = nullptr;
        int stride = 0;
        size_t bytes = 0;
    } attr[SWR_MAX_VERTEX_ATTRIBS];

    // arenas
    DevBuf idx, attr0, wverts, app, app_ex, live, grp, grp_ex, scan_tmp, prims, pk_a, pv_a, pk_b, pv_b,
        ranges, hist, hist_ex, hist_tmp, status, tgt[3], log, logcount,
        bigq,       // per-primitive first pair index
        clipq,      // triangles queued for the clipper
        clipq_cnt,  // its per-region counters
        chunk_kept, // pairs kept per expansion chunk
        vwin, vmask,  // post-transform vertices of the vertex pass
        vmod_out,     // rows a vertex module wrote
        zscratch[2];  // scratch of the differentiation kernels
    uint64_t prim_cap = 0, pair_cap = 0;
    unsigned long long *status_host = nullptr;

    bool profiling = false;
    hipEvent_t ev[S_COUNT + 1] = {};
    swr_draw_stats stats{};
    bool async_pending = false;
    swr_tuning tune{};
    uint64_t diff_tag[9] = {};
    DevBuf diff_table;                        // scene table of the last multi-mesh swr_diff
    std::vector<unsigned char> diff_table_host;
    uint64_t diff_table_gen = 0;
};

// A texture and its mip chain on the device (include/swr_hip.h, swr_texture_create).
struct swr_texture {
    swr_context *ctx = nullptr;
    DevBuf texels;  // every level, level k at off[k]
    DevBuf desc;    // TexDev
};

// A batched draw's device tables (include/swr_hip.h, swr_batch_create).
struct swr_batch {
    swr_context *ctx = nullptr;
    DevBuf scenes;               // SceneDev[nscenes]
    DevBuf bmap;                 // u32 scene of each virtual 1024-triangle batch
    uint32_t nscenes = 0, batches = 0;
    uint64_t vtris = 0;          // virtual triangles: every scene padded to whole batches
    uint64_t tris = 0;           // input triangles
    uint64_t bins[3] = {0, 0, 0};  // sum of the scenes' bins of 8, 16 and 32 pixels
    unsigned tmask = 7;          // bit i: every scene provides target i
    bool spans = false;          // some scene draws in Span or Adaptive mode
};

namespace {

int fail(swr_context *c, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int fail(swr_context *c, int code, const char *fmt, ...)
{
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
        c->err_code = code;
    }
    return code;
}

#define HIP_TRY(c, expr)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ == hipErrorStreamCaptureUnsupported)                                           \
            return fail((c), SWR_ERR_CAPACITY,                                                \
                        "%s: a device arena must grow while the stream is being captured "    \
                        "(reserve it before capturing: swr_reserve or one uncaptured call)",  \
                        #expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((c), SWR_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

bool stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return cs != hipStreamCaptureStatusNone;
}

hipError_t ensure(swr_context *c, DevBuf &b, size_t bytes)
{
    if (bytes == 0)
        bytes = 256;
    if (b.bytes >= bytes)
        return hipSuccess;
    // growing synchronizes, frees and allocates: none of that may enter a stream capture
    if (stream_capturing(c->stream))
        return hipErrorStreamCaptureUnsupported;
    if (b.p) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess)
            return e;
        (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return e;
    }
    b.bytes = want;
    return hipSuccess;
}

bool is_device_ptr(const void *p)
{
    if (!p)
        return false;
    hipPointerAttribute_t at;
    hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

uint64_t vs_need_bytes(const swr_vertex_shader_desc &vs)
{
    switch (vs.program) {
    case SWR_VS_RAW: return 4ull * (4 + vs.avar_count + vs.pvar_count);
    case SWR_VS_POS_RGB: return 24;
    case SWR_VS_MVP_POS_RGB_UV:
    case SWR_VS_MVP_POS_UV: return 32;
    default: return 0;
    }
}

// Stages a host or device buffer: device memory is used in place, host memory is copied.
int resolve_in(swr_context *c, const void *src, size_t bytes, DevBuf &stage, const void **out)
{
    if (!src || is_device_ptr(src)) {
        *out = src;
        return SWR_OK;
    }
    HIP_TRY(c, ensure(c, stage, bytes));
    HIP_TRY(c, hipMemcpyAsync(stage.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    *out = stage.p;
    return SWR_OK;
}

// Vertex attribute 0 on the device for a drawElements over indices[0 .. n): device data is
// used in place; host data is copied, sized from buffer_bytes or from the largest index.
int resolve_attr0(swr_context *c, const int32_t *indices, size_t n, const void **dattr,
                  uint64_t *attr_bytes)
{
    *dattr = nullptr;
    *attr_bytes = 0;
    if (c->vs.program == SWR_VS_DUMMY)
        return SWR_OK;
    const auto &A = c->attr[0];
    if (!A.ptr)
        return fail(c, SWR_ERR_INVALID_ARG, "vertex attribute 0 is not set");
    if (is_device_ptr(A.ptr)) {
        *dattr = A.ptr;
        *attr_bytes = A.bytes;
        return SWR_OK;
    }
    size_t bytes = A.bytes;
    if (!bytes) {
        if (is_device_ptr(indices))
            return fail(c, SWR_ERR_INVALID_ARG,
                        "host vertex data with device indices needs buffer_bytes");
        int32_t mx = 0;
        for (size_t i = 0; i < n; ++i) {
            if (indices[i] < 0)
                return fail(c, SWR_ERR_INDEX_RANGE, "negative index %d at %zu", indices[i], i);
            mx = std::max(mx, indices[i]);
        }
        bytes = (size_t)mx * (size_t)A.stride + vs_need_bytes(c->vs);
    }
    int rc = resolve_in(c, A.ptr, bytes, c->attr0, dattr);
    if (rc)
        return rc;
    *attr_bytes = bytes;
    return SWR_OK;
}

struct TargetMap {
    void *dev[3] = {nullptr, nullptr, nullptr};
    void *host[3] = {nullptr, nullptr, nullptr};
    size_t bytes[3] = {0, 0, 0};
    uint32_t *log = nullptr;
    uint32_t *log_host = nullptr;
    size_t log_bytes = 0;
    unsigned long long *logcount = nullptr;
    unsigned long long *logcount_host = nullptr;
};

int targets_needed(int program, bool need[3])
{
    need[0] = need[1] = need[2] = false;
    switch (program) {
    case SWR_PS_DEPTH_CHECKER:
    case SWR_PS_HASH:
    case SWR_PS_VISIBILITY:
        need[0] = need[1] = true;
        break;
    case SWR_PS_RGB:
    case SWR_PS_COVERAGE:
    case SWR_PS_LOG:
    case SWR_PS_TEXTURE:
        need[0] = true;
        break;
    default:
        break;
    }
    return 0;
}

int map_targets(swr_context *c, TargetMap &m)
{
    const swr_pixel_shader_desc &ps = c->ps;
    if (ps.width < 0 || ps.height < 0 || ps.pitch < ps.width)
        return fail(c, SWR_ERR_INVALID_ARG, "bad target size %dx%d pitch %d", ps.width, ps.height,
                    ps.pitch);
    // a module's uniforms are read by its kernels as they are: nothing stages host memory
    if (ps.program == SWR_PS_MODULE && ps.uniforms && !is_device_ptr(ps.uniforms))
        return fail(c, SWR_ERR_INVALID_ARG, "a pixel module's uniforms must be device memory");
    bool need[3];
    targets_needed(ps.program, need);
    void *const t[3] = {ps.target0, ps.target1, ps.target2};
    size_t bytes = (size_t)ps.pitch * (size_t)ps.height * 4;
    for (int i = 0; i < 3; ++i) {
        if (!need[i])
            continue;
        if (!t[i])
            return fail(c, SWR_ERR_INVALID_ARG, "pixel program %d needs target%d", ps.program, i);
        if (is_device_ptr(t[i])) {
            m.dev[i] = t[i];
        } else {
            HIP_TRY(c, ensure(c, c->tgt[i], bytes));
            HIP_TRY(c, hipMemcpyAsync(c->tgt[i].p, t[i], bytes, hipMemcpyHostToDevice, c->stream));
            m.dev[i] = c->tgt[i].p;
            m.host[i] = t[i];
            m.bytes[i] = bytes;
        }
    }
    if (ps.program == SWR_PS_LOG) {
        if (!ps.log_count || (ps.log_capacity && !ps.log))
            return fail(c, SWR_ERR_INVALID_ARG, "log program needs log and log_count");
        size_t lb = (size_t)ps.log_capacity * SWR_LOG_WORDS * 4;
        if (ps.log_capacity == 0 || is_device_ptr(ps.log)) {
            m.log = ps.log;
        } else {
            HIP_TRY(c, ensure(c, c->log, lb));
            m.log = (uint32_t *)c->log.p;
            m.log_host = ps.log;
            m.log_bytes = lb;
        }
        if (is_device_ptr(ps.log_count)) {
            m.logcount = (unsigned long long *)ps.log_count;
        } else {
            HIP_TRY(c, ensure(c, c->logcount, 8));
            HIP_TRY(c, hipMemcpyAsync(c->logcount.p, ps.log_count, 8, hipMemcpyHostToDevice,
                                      c->stream));
            m.logcount = (unsigned long long *)c->logcount.p;
            m.logcount_host = (unsigned long long *)ps.log_count;
        }
    }
    return SWR_OK;
}

int unmap_targets(swr_context *c, const TargetMap &m)
{
    for (int i = 0; i < 3; ++i)
        if (m.host[i])
            HIP_TRY(c, hipMemcpyAsync(m.host[i], m.dev[i], m.bytes[i], hipMemcpyDeviceToHost,
                                      c->stream));
    if (m.log_host)
        HIP_TRY(c, hipMemcpyAsync(m.log_host, m.log, m.log_bytes, hipMemcpyDeviceToHost, c->stream));
    if (m.logcount_host)
        HIP_TRY(c, hipMemcpyAsync(m.logcount_host, m.logcount, 8, hipMemcpyDeviceToHost, c->stream));
    return SWR_OK;
}

void record(swr_context *c, int i)
{
    if (c->profiling)
        (void)hipEventRecord(c->ev[i], c->stream);
}

int check_status(swr_context *c)
{
    unsigned long long err = c->status_host[ST_ERROR];
    if (err & 1ull)
        return fail(c, SWR_ERR_INDEX_RANGE, "an index addresses outside the vertex data");
    if (err & 2ull)
        return fail(c, SWR_ERR_CLIP_OVERFLOW, "a clipped polygon exceeded %d vertices",
                    SWR_MAX_CLIP_VERTS);
    if (c->status_host[ST_OVERFLOW])
        return fail(c, SWR_ERR_CAPACITY, "arena overflow (prims %llu / %llu, pairs %llu / %llu)",
                    c->status_host[ST_PRIMS], (unsigned long long)c->prim_cap,
                    c->status_host[ST_PAIRS], (unsigned long long)c->pair_cap);
    return SWR_OK;
}

// Failures of earlier asynchronous draws (the device's ST_STICKY word, in the last status
// copy), reported once: the device word is cleared when they are reported.
int take_sticky(swr_context *c)
{
    const unsigned long long f = c->status_host[ST_STICKY];
    if (!f)
        return SWR_OK;
    c->status_host[ST_STICKY] = 0;
    HIP_TRY(c, hipMemsetAsync(static_cast<unsigned long long *>(c->status.p) + ST_STICKY, 0, 8,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const unsigned long long err = f >> 1;
    if (err & 1ull)
        return fail(c, SWR_ERR_INDEX_RANGE,
                    "an asynchronous draw: an index addresses outside the vertex data");
    if (err & 2ull)
        return fail(c, SWR_ERR_CLIP_OVERFLOW,
                    "an asynchronous draw: a clipped polygon exceeded %d vertices",
                    SWR_MAX_CLIP_VERTS);
    return fail(c, SWR_ERR_CAPACITY,
                "an asynchronous draw overflowed the context's arenas and drew nothing "
                "(reserve them first: swr_reserve or a synchronous draw of the same scene)");
}

void collect_stats(swr_context *c)
{
    c->stats.raster_primitives = c->status_host[ST_PRIMS];
    c->stats.tile_pairs = c->status_host[ST_KEPT];
    c->stats.pairs_emitted = c->status_host[ST_PAIRS];
    if (c->profiling) {
        for (int i = 0; i < S_COUNT; ++i) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]) == hipSuccess)
                c->stats.stage_ms[i] = ms;
            else
                c->stats.stage_ms[i] = 0.0f;
        }
    }
}

// Device memory one triangle draw works in, beyond its inputs and render targets.
struct Arenas {
    uint2 *app = nullptr, *app_ex = nullptr;
    uint32_t *live = nullptr, *clipq = nullptr, *clipq_cnt = nullptr;
    uint2 *grp = nullptr, *grp_ex = nullptr;
    void *scan_tmp = nullptr;
    uint32_t *ranges = nullptr;
    float *prims = nullptr;
    uint32_t *bigq = nullptr, *pk_a = nullptr, *pv_a = nullptr, *pk_b = nullptr, *pv_b = nullptr;
    uint32_t *hist = nullptr, *hist_ex = nullptr, *hist_tmp = nullptr;
    uint32_t *chunk_kept = nullptr;
    uint64_t prim_cap = 0, pair_cap = 0;
    uint32_t G = 1;  // radix sort workgroups
};

// Sizes of the arenas that depend only on the input (triangles, batches, tiles).
struct FixedSizes {
    size_t app, live, clipq, clipq_cnt, grp, scan_tmp, ranges;
};

FixedSizes fixed_sizes(size_t T, uint64_t ngrp, uint32_t tiles)
{
    FixedSizes f;
    f.app = T * 8;
    f.live = T * 4;
    const size_t blocks = (T + 255) / 256, per_region = ((blocks + 63) / 64) * 256;
    f.clipq = 64 * per_region * 4;
    f.clipq_cnt = DRAW_COUNTER_WORDS * 4;
    f.grp = ngrp * 8;
    f.scan_tmp = (scan_tmp_elems(T) + scan_tmp_elems(ngrp)) * 8;
    f.ranges = (size_t)tiles * 8;
    return f;
}

uint32_t sort_blocks(uint64_t pair_cap)
{
    return (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(1, pair_cap / 4096));
}

size_t hist_bytes(uint32_t G) { return (size_t)256 * G * 4; }
// kept-pair counts of the expansion chunks (256 slots at the least)
size_t chunk_kept_bytes(uint64_t pair_cap) { return (pair_cap / 256 + 1) * 4; }
// the expansion chunk launch_expand uses for a tuning value
int expand_chunk_of(int tuned)
{
    return (tuned == 256 || tuned == 512 || tuned == 2048 || tuned == 4096) ? tuned : 1024;
}
size_t hist_tmp_bytes(uint32_t G) { return scan_tmp_elems((uint64_t)256 * G) * 4; }

// ---- the no-allocation draw's scratch ------------------------------------------------------
// One array per GPU, in the code object's device memory: swr_draw_elements_noalloc carves all
// of its arenas from it and allocates nothing.  In host code `scratch` names a host-side
// stand-in, not device memory: its device address is looked up once per device
// (hipGetSymbolAddress with that device current, at swr_create) and only that address is
// handed to kernels.
static __device__ __align__(256) unsigned char scratch[SWR_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};
bool g_scratch_done[MAX_DEVICES] = {};

// The calling thread's current device is `device`.  Returns hipSuccess once the address is
// known (the lookup runs once per device).
hipError_t resolve_scratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device])
        return g_scratch_ptr[device] ? hipSuccess : hipErrorInvalidSymbol;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_done[device] = true;
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    if (e != hipSuccess)
        (void)hipGetLastError();
    return e;
}

void *scratch_of(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

// Lays every arena of a draw out in [base, base + bytes): the input-sized ones first, then
// primitive records and tile pairs from what remains.  Returns false when the input-sized
// arenas and a minimal record / pair capacity do not fit.
bool carve_arenas(unsigned char *base, size_t bytes, size_t T, uint64_t ngrp, uint32_t tiles,
                  int stride_words, Arenas &A)
{
    const FixedSizes f = fixed_sizes(T, ngrp, tiles);
    size_t off = 0;
    auto take = [&](size_t n) -> unsigned char * {
        unsigned char *p = base + off;
        off += (std::max<size_t>(n, 1) + 255) & ~(size_t)255;
        return p;
    };
    const size_t fixed_total = 2 * ((f.app + 255) & ~(size_t)255) + ((f.live + 255) & ~(size_t)255) +
                               ((f.clipq + 255) & ~(size_t)255) +
                               ((f.clipq_cnt + 255) & ~(size_t)255) +
                               2 * ((f.grp + 255) & ~(size_t)255) +
                               ((f.scan_tmp + 255) & ~(size_t)255) + ((f.ranges + 255) & ~(size_t)255);
    if (fixed_total >= bytes)
        return false;
    A.app = (uint2 *)take(f.app);
    A.app_ex = (uint2 *)take(f.app);
    A.live = (uint32_t *)take(f.live);
    A.clipq = (uint32_t *)take(f.clipq);
    A.clipq_cnt = (uint32_t *)take(f.clipq_cnt);
    A.grp = (uint2 *)take(f.grp);
    A.grp_ex = (uint2 *)take(f.grp);
    A.scan_tmp = take(f.scan_tmp);
    A.ranges = (uint32_t *)take(f.ranges);
    // records: every input triangle and an eighth more for clip fans; pairs: the rest
    const size_t rec_bytes = (size_t)stride_words * 4 + 4;  // record + its first-pair index
    A.prim_cap = T + T / 8 + 256;
    const size_t slack = 12 * 256;  // alignment of the arenas below
    if (off + A.prim_cap * rec_bytes + slack >= bytes)
        return false;
    size_t rest = bytes - off - A.prim_cap * rec_bytes - slack;
    uint64_t pairs = rest / 16;
    A.G = sort_blocks(pairs);
    const size_t hb = 2 * hist_bytes(A.G) + hist_tmp_bytes(A.G) + chunk_kept_bytes(pairs) + 256;
    if (hb >= rest)
        return false;
    pairs = (rest - hb) / 16;
    if (pairs < 4096)
        return false;
    A.pair_cap = pairs;
    A.prims = (float *)take(A.prim_cap * (size_t)stride_words * 4);
    A.bigq = (uint32_t *)take(A.prim_cap * 4);
    A.pk_a = (uint32_t *)take(A.pair_cap * 4);
    A.pv_a = (uint32_t *)take(A.pair_cap * 4);
    A.pk_b = (uint32_t *)take(A.pair_cap * 4);
    A.pv_b = (uint32_t *)take(A.pair_cap * 4);
    A.hist = (uint32_t *)take(hist_bytes(A.G));
    A.hist_ex = (uint32_t *)take(hist_bytes(A.G));
    A.hist_tmp = (uint32_t *)take(hist_tmp_bytes(A.G));
    A.chunk_kept = (uint32_t *)take(chunk_kept_bytes(A.pair_cap));
    return off <= bytes;
}

// The triangle draw: drawElements (window_input = false), the IRasterizer seam, or a batch
// of scenes (bt != nullptr: inputs, state and targets come from the scene table).
// noalloc: the asynchronous draw whose arenas are carved from the device's scratch array
// (nothing is allocated, freed or set by a memset; sizes beyond the array fail at once).
int draw_triangles(swr_context *c, const int32_t *indices, size_t count, bool window_input,
                   const swr_rasterizer_vertex *wverts, size_t nverts, bool sync,
                   const swr_batch *bt = nullptr, bool noalloc = false)
{
    const size_t T = bt ? bt->vtris : count / 3;
    std::memset(&c->stats, 0, sizeof c->stats);
    c->stats.input_primitives = bt ? bt->tris : T;
    // screen bins of 8 << lvl pixels (the tuning's bin size)
    const int lvl = c->tune.bin_size == 32 ? 2 : (c->tune.bin_size == 16 ? 1 : 0);
    const int bin_shift = 3 + lvl, bin = 1 << bin_shift;
    c->stats.tiles = bt ? bt->bins[lvl]
                        : (uint64_t)((c->ps.width + bin - 1) / bin) *
                              (uint64_t)((c->ps.height + bin - 1) / bin);
    if (T == 0 || c->ps.program == SWR_PS_NULL)
        return SWR_OK;  // nothing observable happens
    if (T > (size_t)(1u << 31) / 8)
        return fail(c, SWR_ERR_CAPACITY, "too many triangles (%zu)", T);
    ParamLayout L;
    if (!raster_layout(c->ps, L))
        return fail(c, SWR_ERR_UNSUPPORTED,
                    "pixel program %d does not support z=%d w=%d avars=%d pvars=%d",
                    c->ps.program, c->ps.interpolate_z, c->ps.interpolate_w, c->ps.avar_count,
                    c->ps.pvar_count);
    const swr_pixel_shader_desc &ps = c->ps;
    const int W = bt ? 0 : ps.width, H = bt ? 0 : ps.height;
    const int tiles_x = (W + bin - 1) / bin, tiles_y = (H + bin - 1) / bin;
    const uint32_t tiles = bt ? (uint32_t)bt->bins[lvl] : (uint32_t)tiles_x * (uint32_t)tiles_y;
    if (tiles == 0)
        return SWR_OK;
    if (!sync && !bt) {
        // nothing is staged or copied back for an asynchronous draw
        if (!is_device_ptr(indices))
            return fail(c, SWR_ERR_INVALID_ARG, "async draws take device buffers only: indices");
        bool need_t[3];
        targets_needed(ps.program, need_t);
        void *const t[3] = {ps.target0, ps.target1, ps.target2};
        for (int i = 0; i < 3; ++i)
            if (need_t[i] && !is_device_ptr(t[i]))
                return fail(c, SWR_ERR_INVALID_ARG,
                            "async draws take device buffers only: target%d", i);
        if (ps.program == SWR_PS_LOG &&
            ((ps.log_capacity && !is_device_ptr(ps.log)) || !is_device_ptr(ps.log_count)))
            return fail(c, SWR_ERR_INVALID_ARG,
                        "async draws take device buffers only: log and log_count");
        if (c->vs.program != SWR_VS_DUMMY && !is_device_ptr(c->attr[0].ptr))
            return fail(c, SWR_ERR_INVALID_ARG,
                        "async draws take device buffers only: vertex attribute 0");
    }

    TargetMap tm;
    int rc = SWR_OK;
    if (bt) {
        bool need_t[3];
        targets_needed(ps.program, need_t);
        for (int i = 0; i < 3; ++i)
            if (need_t[i] && !((bt->tmask >> i) & 1u))
                return fail(c, SWR_ERR_INVALID_ARG,
                            "pixel program %d needs target%d in every scene", ps.program, i);
        if (ps.program == SWR_PS_LOG)
            return fail(c, SWR_ERR_UNSUPPORTED, "the log program is not batched");
    } else {
        rc = map_targets(c, tm);
        if (rc)
            return rc;
    }

    // inputs
    const void *didx = nullptr;
    const void *dattr = nullptr;
    const void *dwv = nullptr;
    uint64_t attr_bytes = 0;
    const uint64_t need = vs_need_bytes(c->vs);
    if (!bt) {
        rc = resolve_in(c, indices, T * 3 * sizeof(int32_t), c->idx, &didx);
        if (rc)
            return rc;
    }
    if (bt) {
        // per-scene inputs, validated by swr_batch_create
    } else if (window_input) {
        rc = resolve_in(c, wverts, nverts * sizeof(swr_rasterizer_vertex), c->wverts, &dwv);
        if (rc)
            return rc;
    } else {
        rc = resolve_attr0(c, indices, T * 3, &dattr, &attr_bytes);
        if (rc)
            return rc;
    }

    const uint32_t batches = (uint32_t)((T + BATCH - 1) / BATCH);
    const uint64_t ngrp = (uint64_t)batches * 2048;
    Arenas A;
    if (noalloc) {
        unsigned char *base = static_cast<unsigned char *>(scratch_of(c->device));
        if (!base)
            return fail(c, SWR_ERR_HIP, "the device address of the scratch array is unknown");
        if (!carve_arenas(base, SWR_NOALLOC_SCRATCH_BYTES, T, ngrp, tiles, L.stride, A))
            return fail(c, SWR_ERR_CAPACITY,
                        "%zu triangles on %u tiles do not fit the %d-byte scratch array of "
                        "swr_draw_elements_noalloc",
                        T, tiles, SWR_NOALLOC_SCRATCH_BYTES);
    } else {
        const FixedSizes f = fixed_sizes(T, ngrp, tiles);
        HIP_TRY(c, ensure(c, c->app, f.app));
        HIP_TRY(c, ensure(c, c->clipq, f.clipq));
        HIP_TRY(c, ensure(c, c->clipq_cnt, f.clipq_cnt));
        HIP_TRY(c, ensure(c, c->app_ex, f.app));
        HIP_TRY(c, ensure(c, c->live, f.live));
        HIP_TRY(c, ensure(c, c->grp, f.grp));
        HIP_TRY(c, ensure(c, c->grp_ex, f.grp));
        HIP_TRY(c, ensure(c, c->scan_tmp, f.scan_tmp));
        HIP_TRY(c, ensure(c, c->ranges, f.ranges));
        A.app = (uint2 *)c->app.p;
        A.app_ex = (uint2 *)c->app_ex.p;
        A.live = (uint32_t *)c->live.p;
        A.clipq = (uint32_t *)c->clipq.p;
        A.clipq_cnt = (uint32_t *)c->clipq_cnt.p;
        A.grp = (uint2 *)c->grp.p;
        A.grp_ex = (uint2 *)c->grp_ex.p;
        A.scan_tmp = c->scan_tmp.p;
        A.ranges = (uint32_t *)c->ranges.p;
        c->prim_cap = std::max<uint64_t>(c->prim_cap, T + 256);
        c->pair_cap = std::max<uint64_t>(c->pair_cap, 2 * T + 4096);
    }

    // the vertex pass: every vertex of attribute 0 shaded once, when the buffer's vertex
    // count is known (tuning: 0 automatic = when each vertex is referenced at least twice on
    // average, so that the pass saves more vertex-program runs than it adds; 1 on, -1 off)
    uint32_t vcount = 0;
    if (!bt && !window_input && !noalloc && c->vs.program != SWR_VS_DUMMY && attr_bytes &&
        attr_bytes >= need && c->attr[0].stride > 0 && c->tune.vertex_pass >= 0) {
        const uint64_t nv = (attr_bytes - need) / (uint64_t)c->attr[0].stride + 1;
        const bool want = c->tune.vertex_pass > 0 || 2 * nv <= 3 * (uint64_t)T;
        if (want && nv < (1ull << 31)) {
            vcount = (uint32_t)nv;
            HIP_TRY(c, ensure(c, c->vwin, (size_t)vcount * 16));
            HIP_TRY(c, ensure(c, c->vmask, (size_t)vcount));
        }
    }

    for (int attempt = 0; attempt < 4; ++attempt) {
        if (!noalloc) {
            HIP_TRY(c, ensure(c, c->prims, c->prim_cap * (size_t)L.stride * 4));
            HIP_TRY(c, ensure(c, c->bigq, c->prim_cap * 4));
            HIP_TRY(c, ensure(c, c->pk_a, c->pair_cap * 4));
            HIP_TRY(c, ensure(c, c->pv_a, c->pair_cap * 4));
            HIP_TRY(c, ensure(c, c->pk_b, c->pair_cap * 4));
            HIP_TRY(c, ensure(c, c->pv_b, c->pair_cap * 4));
            A.G = sort_blocks(c->pair_cap);
            HIP_TRY(c, ensure(c, c->hist, hist_bytes(A.G)));
            HIP_TRY(c, ensure(c, c->hist_ex, hist_bytes(A.G)));
            HIP_TRY(c, ensure(c, c->hist_tmp, hist_tmp_bytes(A.G)));
            HIP_TRY(c, ensure(c, c->chunk_kept, chunk_kept_bytes(c->pair_cap)));
            A.prims = (float *)c->prims.p;
            A.bigq = (uint32_t *)c->bigq.p;
            A.pk_a = (uint32_t *)c->pk_a.p;
            A.pv_a = (uint32_t *)c->pv_a.p;
            A.pk_b = (uint32_t *)c->pk_b.p;
            A.pv_b = (uint32_t *)c->pv_b.p;
            A.hist = (uint32_t *)c->hist.p;
            A.hist_ex = (uint32_t *)c->hist_ex.p;
            A.hist_tmp = (uint32_t *)c->hist_tmp.p;
            A.chunk_kept = (uint32_t *)c->chunk_kept.p;
            A.prim_cap = c->prim_cap;
            A.pair_cap = c->pair_cap;
        }
        const uint32_t G = A.G;

        GeomArgs g;
        std::memset(&g, 0, sizeof g);
        g.indices = (const int32_t *)didx;
        g.attr0 = (const uint8_t *)dattr;
        g.stride0 = c->attr[0].stride;
        g.attr0_bytes = attr_bytes;
        g.attr0_need = need;
        std::memcpy(g.mvp, c->vs.mvp, sizeof g.mvp);
        g.vs_program = c->vs.program;
        g.vs_na = c->vs.avar_count;
        g.vs_np = c->vs.pvar_count;
        g.tris = (uint32_t)T;
        g.batches = batches;
        g.raster_mode = c->raster_mode;
        g.cull_mode = c->cull;
        g.sc_minx = c->sc[0];
        g.sc_miny = c->sc[1];
        g.sc_maxx = c->sc[0] + c->sc[2];
        g.sc_maxy = c->sc[1] + c->sc[3];
        // VertexProcessor.cpp:50-53, 373 (host float arithmetic, one rounding per op)
        volatile float px = (float)c->vp[2] / 2.0f, py = (float)c->vp[3] / 2.0f;
        g.vp_px = px;
        g.vp_py = py;
        g.vp_ox = (float)c->vp[0] + px;
        g.vp_oy = (float)c->vp[1] + py;
        volatile float dzs = 0.5f * (c->df - c->dn), dzb = 0.5f * (c->dn + c->df);
        g.dz_scale = dzs;
        g.dz_bias = dzb;
        g.width = W;
        g.height = H;
        g.tiles_x = tiles_x;
        g.tiles_y = tiles_y;
        g.bin_shift = bin_shift;
        // 16x16 bins carry their blocks' classes when primitive indices fit 24 bits
        g.cls4 = (bin_shift == 4 && A.prim_cap <= (uint64_t)PAIR_CLS4_PRIM_MASK + 1) ? 1 : 0;
        g.lay = L;
        g.window_input = window_input ? 1 : 0;
        g.wverts = (const swr_rasterizer_vertex *)dwv;
        g.wvert_count = nverts;
        g.app = A.app;
        g.app_ex = A.app_ex;
        g.live = window_input ? nullptr : A.live;
        g.grp = A.grp;
        g.grp_ex = A.grp_ex;
        g.prims = A.prims;
        g.prim_cap = A.prim_cap;
        g.pair_keys = A.pk_a;
        g.pair_vals = A.pv_a;
        g.pair_cap = A.pair_cap;
        g.status = (unsigned long long *)c->status.p;
        g.ranges = A.ranges;
        g.tiles = tiles;
        g.sentinel = pair_sentinel(tiles);
        g.prim_pb = A.bigq;
        g.chunk_kept = A.chunk_kept;
        g.clipq = A.clipq;
        g.clipq_cnt = A.clipq_cnt;
        if (bt) {
            g.scenes = (const SceneDev *)bt->scenes.p;
            g.batch_scene = (const uint32_t *)bt->bmap.p;
            g.nscenes = bt->nscenes;
        }
        if (vcount) {
            g.vwin = (float4 *)c->vwin.p;
            g.vmask = (uint8_t *)c->vmask.p;
            g.nverts = vcount;
        }

        record(c, 0);
        HIP_TRY(c, launch_begin(g, c->stream));
        HIP_TRY(c, launch_geom_count(g, c->stream));
        record(c, 1);
        if (!window_input)
            HIP_TRY(c, scan_fans_and_pairs(g.app, A.app_ex, T, A.live, g.grp, A.grp_ex, ngrp,
                                           A.scan_tmp, g.status, A.prim_cap, A.pair_cap,
                                           c->stream));
        else
            HIP_TRY(c, scan_pairs(g.grp, A.grp_ex, ngrp, (uint2 *)A.scan_tmp, g.status,
                                  A.prim_cap, A.pair_cap, c->stream));
        record(c, 2);
        HIP_TRY(c, launch_geom_emit(g, c->stream));
        HIP_TRY(c, launch_expand(g, c->stream, c->tune.expand_chunk));
        record(c, 3);
        SortBuffers sb;
        sb.keys_a = A.pk_a;
        sb.vals_a = A.pv_a;
        sb.keys_b = A.pk_b;
        sb.vals_b = A.pv_b;
        sb.hist = A.hist;
        sb.hist_ex = A.hist_ex;
        sb.scan_tmp = A.hist_tmp;
        sb.ranges = A.ranges;
        sb.blocks = G;
        sb.chunk_kept = A.chunk_kept;
        sb.chunk = (uint32_t)expand_chunk_of(c->tune.expand_chunk);
        HIP_TRY(c, sort_pairs(sb, g.status, tiles, c->stream));
        record(c, 4);
        HIP_TRY(c, launch_ranges(sb, g.status, tiles, c->stream));
        record(c, 5);

        RasterArgs r;
        std::memset(&r, 0, sizeof r);
        r.prims = A.prims;
        r.stride = L.stride;
        r.ranges = A.ranges;
        r.vals = sb.sorted_vals;
        r.status = g.status;
        r.tiles = (int)tiles;
        r.tiles_x = tiles_x;
        r.bin_shift = bin_shift;
        r.cls4 = g.cls4;
        r.waves_per_bin = c->tune.waves_per_bin ? c->tune.waves_per_bin : 4;
        if (r.waves_per_bin > (1 << (2 * lvl)))
            r.waves_per_bin = 1 << (2 * lvl);
        r.width = W;
        r.height = H;
        r.pitch = ps.pitch;
        r.sc_minx = g.sc_minx;
        r.sc_maxx = g.sc_maxx;
        r.na_set = L.set_na;
        r.np_set = L.set_np;
        r.iz = ps.interpolate_z ? 1 : 0;
        r.iw = ps.interpolate_w ? 1 : 0;
        r.t0 = tm.dev[0];
        r.t1 = tm.dev[1];
        r.t2 = tm.dev[2];
        r.log = tm.log;
        r.log_cap = ps.log_capacity;
        r.log_count = tm.logcount;
        r.checker_cells = ps.checker_cells;
        r.user = ps.uniforms;
        if (bt) {
            r.scenes = g.scenes;
            r.nscenes = g.nscenes;
        }
        r.xcd = c->tune.xcd_order >= 0 ? 1 : 0;
        // Block mode draws every triangle with Block traversal (Rasterizer.h:433-445)
        r.spans = bt ? (bt->spans ? 1 : 0) : (c->raster_mode != SWR_RASTER_BLOCK ? 1 : 0);
        r.queue = c->tune.raster_queue > 0 ? A.clipq_cnt + CLIPQ_COUNTERS : nullptr;
        r.fold = sync ? 0 : 1;
        if (ps.program == SWR_PS_TEXTURE)
            r.tex = (const TexDev *)static_cast<const swr_texture *>(ps.texture)->desc.p;
        HIP_TRY(c, launch_raster(ps, L, r, c->stream, c->tune.raster_waves ? c->tune.raster_waves : 4,
                                 c->tune.raster_group));
        record(c, 6);
        HIP_TRY(c, hipMemcpyAsync(c->status_host, c->status.p, ST_WORDS * 8, hipMemcpyDeviceToHost,
                                  c->stream));
        if (!sync) {
            c->async_pending = true;
            return SWR_OK;
        }
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (c->status_host[ST_OVERFLOW] && !(c->status_host[ST_ERROR])) {
            if (c->status_host[ST_OVERFLOW] & 1ull)
                c->prim_cap = c->status_host[ST_PRIMS] + c->status_host[ST_PRIMS] / 4 + 256;
            if (c->status_host[ST_OVERFLOW] & 2ull)
                c->pair_cap = c->status_host[ST_PAIRS] + c->status_host[ST_PAIRS] / 4 + 4096;
            continue;  // nothing was rasterized: run the draw again with larger arenas
        }
        break;
    }
    rc = take_sticky(c);
    if (rc)
        return rc;
    rc = check_status(c);
    if (rc)
        return rc;
    if (!bt)
        rc = unmap_targets(c, tm);
    if (rc)
        return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    collect_stats(c);
    return SWR_OK;
}

// Points and lines: drawElements (window_input = false) or the IRasterizer list entries.
// The count pass publishes the fragment total, which sizes the fragment arena before the
// emit pass (one host round trip per draw).
int draw_lines_points(swr_context *c, int mode, const int32_t *indices, size_t count,
                      bool window_input, const swr_rasterizer_vertex *wverts, size_t nverts)
{
    const size_t per = mode == SWR_DRAW_LINE ? 2 : 1;
    const size_t P = count / per;  // a partial last line is dropped
    std::memset(&c->stats, 0, sizeof c->stats);
    c->stats.input_primitives = P;
    c->stats.tiles = (uint64_t)((c->ps.width + TILE - 1) / TILE) *
                     (uint64_t)((c->ps.height + TILE - 1) / TILE);
    if (P == 0 || c->ps.program == SWR_PS_NULL)
        return SWR_OK;
    if (c->ps.program == SWR_PS_TEXTURE)
        return fail(c, SWR_ERR_UNSUPPORTED,
                    "the texture program needs a triangle's planes: points and lines have none");
    if (P > (size_t)(1u << 30))
        return fail(c, SWR_ERR_CAPACITY, "too many primitives (%zu)", P);
    ParamLayout L;
    if (!raster_layout(c->ps, L))
        return fail(c, SWR_ERR_UNSUPPORTED,
                    "pixel program %d does not support z=%d w=%d avars=%d pvars=%d",
                    c->ps.program, c->ps.interpolate_z, c->ps.interpolate_w, c->ps.avar_count,
                    c->ps.pvar_count);
    const swr_pixel_shader_desc &ps = c->ps;
    const int W = ps.width, H = ps.height;
    const int tiles_x = (W + TILE - 1) / TILE, tiles_y = (H + TILE - 1) / TILE;
    const uint32_t tiles = (uint32_t)tiles_x * (uint32_t)tiles_y;
    if (tiles == 0)
        return SWR_OK;

    TargetMap tm;
    int rc = map_targets(c, tm);
    if (rc)
        return rc;
    const void *didx = nullptr, *dattr = nullptr, *dwv = nullptr;
    uint64_t attr_bytes = 0;
    rc = resolve_in(c, indices, P * per * sizeof(int32_t), c->idx, &didx);
    if (rc)
        return rc;
    if (window_input)
        rc = resolve_in(c, wverts, nverts * sizeof(swr_rasterizer_vertex), c->wverts, &dwv);
    else
        rc = resolve_attr0(c, indices, P * per, &dattr, &attr_bytes);
    if (rc)
        return rc;

    HIP_TRY(c, ensure(c, c->app, P * 4));
    HIP_TRY(c, ensure(c, c->app_ex, P * 4));
    HIP_TRY(c, ensure(c, c->scan_tmp, scan_tmp_elems(P) * 8));
    HIP_TRY(c, ensure(c, c->ranges, (size_t)tiles * 8));
    HIP_TRY(c, ensure(c, c->status, ST_WORDS * 8));

    LineArgs A;
    std::memset(&A, 0, sizeof A);
    GeomArgs &g = A.g;
    g.indices = (const int32_t *)didx;
    g.attr0 = (const uint8_t *)dattr;
    g.stride0 = c->attr[0].stride;
    g.attr0_bytes = attr_bytes;
    g.attr0_need = vs_need_bytes(c->vs);
    std::memcpy(g.mvp, c->vs.mvp, sizeof g.mvp);
    g.vs_program = c->vs.program;
    g.vs_na = c->vs.avar_count;
    g.vs_np = c->vs.pvar_count;
    g.sc_minx = c->sc[0];
    g.sc_miny = c->sc[1];
    g.sc_maxx = c->sc[0] + c->sc[2];
    g.sc_maxy = c->sc[1] + c->sc[3];
    // VertexProcessor.cpp:50-53, 373 (host float arithmetic, one rounding per op)
    volatile float px = (float)c->vp[2] / 2.0f, py = (float)c->vp[3] / 2.0f;
    g.vp_px = px;
    g.vp_py = py;
    g.vp_ox = (float)c->vp[0] + px;
    g.vp_oy = (float)c->vp[1] + py;
    volatile float dzs = 0.5f * (c->df - c->dn), dzb = 0.5f * (c->dn + c->df);
    g.dz_scale = dzs;
    g.dz_bias = dzb;
    g.width = W;
    g.height = H;
    g.tiles_x = tiles_x;
    g.tiles_y = tiles_y;
    g.bin_shift = 3;
    g.lay = L;
    g.window_input = window_input ? 1 : 0;
    g.wverts = (const swr_rasterizer_vertex *)dwv;
    g.wvert_count = nverts;
    g.status = (unsigned long long *)c->status.p;
    g.ranges = (uint32_t *)c->ranges.p;
    g.tiles = tiles;
    A.mode = mode;
    A.prims = P;
    A.iz = ps.interpolate_z ? 1 : 0;
    A.iw = ps.interpolate_w ? 1 : 0;
    A.count = (uint32_t *)c->app.p;
    A.first = (const uint32_t *)c->app_ex.p;
    A.fstride = fragment_stride(L);

    record(c, 0);
    HIP_TRY(c, launch_lines_count(A, c->stream));
    record(c, 1);
    HIP_TRY(c, scan_u32(A.count, (uint32_t *)c->app_ex.p, P, (uint32_t *)c->scan_tmp.p, c->stream));
    record(c, 2);
    HIP_TRY(c, hipMemcpyAsync(c->status_host, c->status.p, ST_WORDS * 8, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    rc = take_sticky(c);
    if (rc)
        return rc;
    if (c->status_host[ST_ERROR] & 4ull)
        return fail(c, SWR_ERR_CAPACITY, "a line spans more than %d steps", SWR_MAX_LINE_STEPS);
    rc = check_status(c);
    if (rc)
        return rc;
    const uint64_t F = c->status_host[ST_PAIRS];
    if (F > (1ull << 31))
        return fail(c, SWR_ERR_CAPACITY, "too many fragments (%llu)", (unsigned long long)F);

    if (F) {
        c->pair_cap = std::max<uint64_t>(c->pair_cap, F);
        HIP_TRY(c, ensure(c, c->prims, F * (size_t)A.fstride * 4));
        HIP_TRY(c, ensure(c, c->pk_a, c->pair_cap * 4));
        HIP_TRY(c, ensure(c, c->pv_a, c->pair_cap * 4));
        HIP_TRY(c, ensure(c, c->pk_b, c->pair_cap * 4));
        HIP_TRY(c, ensure(c, c->pv_b, c->pair_cap * 4));
        uint32_t G = (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(1, c->pair_cap / 4096));
        HIP_TRY(c, ensure(c, c->hist, (size_t)256 * G * 4));
        HIP_TRY(c, ensure(c, c->hist_ex, (size_t)256 * G * 4));
        HIP_TRY(c, ensure(c, c->hist_tmp, scan_tmp_elems((uint64_t)256 * G) * 4));
        A.frags = (float *)c->prims.p;
        A.keys = (uint32_t *)c->pk_a.p;
        A.vals = (uint32_t *)c->pv_a.p;
        HIP_TRY(c, launch_lines_emit(A, c->stream));
        record(c, 3);
        SortBuffers sb;
        sb.keys_a = (uint32_t *)c->pk_a.p;
        sb.vals_a = (uint32_t *)c->pv_a.p;
        sb.keys_b = (uint32_t *)c->pk_b.p;
        sb.vals_b = (uint32_t *)c->pv_b.p;
        sb.hist = (uint32_t *)c->hist.p;
        sb.hist_ex = (uint32_t *)c->hist_ex.p;
        sb.scan_tmp = (uint32_t *)c->hist_tmp.p;
        sb.ranges = (uint32_t *)c->ranges.p;
        sb.blocks = G;
        HIP_TRY(c, sort_pairs(sb, g.status, tiles, c->stream));
        record(c, 4);
        HIP_TRY(c, launch_ranges(sb, g.status, tiles, c->stream));
        record(c, 5);

        RasterArgs r;
        std::memset(&r, 0, sizeof r);
        r.prims = (const float *)c->prims.p;
        r.stride = A.fstride;
        r.ranges = (const uint32_t *)c->ranges.p;
        r.vals = sb.sorted_vals;
        r.status = g.status;
        r.tiles = (int)tiles;
        r.tiles_x = tiles_x;
        r.bin_shift = 3;
        r.width = W;
        r.height = H;
        r.pitch = ps.pitch;
        r.sc_minx = g.sc_minx;
        r.sc_maxx = g.sc_maxx;
        r.na_set = L.set_na;
        r.np_set = L.set_np;
        r.iz = A.iz;
        r.iw = A.iw;
        r.t0 = tm.dev[0];
        r.t1 = tm.dev[1];
        r.t2 = tm.dev[2];
        r.log = tm.log;
        r.log_cap = ps.log_capacity;
        r.log_count = tm.logcount;
        r.checker_cells = ps.checker_cells;
        r.user = ps.uniforms;
        HIP_TRY(c, launch_lines_raster(ps, L, r, c->stream));
        record(c, 6);
        HIP_TRY(c, hipMemcpyAsync(c->status_host, c->status.p, ST_WORDS * 8,
                                  hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        rc = check_status(c);
        if (rc)
            return rc;
    }
    rc = unmap_targets(c, tm);
    if (rc)
        return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->stats.raster_primitives = c->status_host[ST_PRIMS];
    c->stats.tile_pairs = F;
    c->stats.pairs_emitted = F;
    if (c->profiling && F)
        collect_stats(c);
    return SWR_OK;
}

}  // namespace

namespace swr {

hipStream_t context_stream(swr_context *c)
{
    (void)hipSetDevice(c->device);
    return c->stream;
}

bool context_side(swr_context *c, hipStream_t *side, hipEvent_t *fork, hipEvent_t *join)
{
    if (!c->side || !c->fork || !c->join)
        return false;
    *side = c->side;
    *fork = c->fork;
    *join = c->join;
    return true;
}

int context_fail(swr_context *c, int code, const char *msg) { return fail(c, code, "%s", msg); }

int context_last_code(swr_context *c) { return c->err_code != SWR_OK ? c->err_code : SWR_ERR_HIP; }

bool context_is_device_ptr(const void *p) { return is_device_ptr(p); }

uint64_t *context_diff_tag(swr_context *c) { return c->diff_tag; }

const void *context_diff_table(swr_context *c, const void *host, size_t bytes, uint64_t *gen)
{
    const unsigned char *h = static_cast<const unsigned char *>(host);
    if (c->diff_table.p && c->diff_table_host.size() == bytes &&
        std::memcmp(c->diff_table_host.data(), h, bytes) == 0) {
        *gen = c->diff_table_gen;
        return c->diff_table.p;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        fail(c, SWR_ERR_INVALID_ARG, "the swr_diff scene table changed while the stream is "
                                     "being captured (run the call once before capturing)");
        return nullptr;
    }
    // queued kernels may still read the old table: wait for them, then copy in stream order
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess)
        e = ensure(c, c->diff_table, bytes);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->diff_table.p, h, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        fail(c, SWR_ERR_HIP, "scene table upload failed: %s", hipGetErrorString(e));
        return nullptr;
    }
    c->diff_table_host.assign(h, h + bytes);
    *gen = ++c->diff_table_gen;
    return c->diff_table.p;
}

void *context_zeroed_scratch(swr_context *c, int slot, size_t bytes)
{
    DevBuf &b = c->zscratch[slot & 1];
    if (b.bytes < bytes) {
        if (stream_capturing(c->stream)) {
            fail(c, SWR_ERR_CAPACITY, "scratch of %zu bytes must grow while the stream is being "
                                      "captured (run the call once before capturing)", bytes);
            return nullptr;
        }
        hipError_t e = ensure(c, b, bytes);
        if (e == hipSuccess)  // ordered before the kernels that use it on this stream
            e = hipMemsetAsync(b.p, 0, b.bytes, c->stream);
        if (e != hipSuccess) {
            fail(c, SWR_ERR_HIP, "scratch allocation failed: %s", hipGetErrorString(e));
            return nullptr;
        }
    }
    return b.p;
}

}  // namespace swr

// ---- C ABI ------------------------------------------------------------------------------
extern "C" {

int swr_create(swr_context **out, int device)
{
    if (!out)
        return SWR_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
        return SWR_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess)
        return SWR_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return SWR_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return SWR_ERR_NO_DEVICE;
    swr_context *c = new swr_context();
    c->device = device;
    c->ps.program = SWR_PS_NULL;
    c->ps.checker_cells = 8.0f;
    c->vs.program = SWR_VS_DUMMY;
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void **)&c->status_host, ST_WORDS * 8, hipHostMallocDefault) != hipSuccess) {
        delete c;
        return SWR_ERR_HIP;
    }
    std::memset(c->status_host, 0, ST_WORDS * 8);
    c->stream = c->own;
    // the no-allocation draw's scratch array: its address on this device, once per device
    // (a failure only disables swr_draw_elements_noalloc, which then reports it)
    (void)resolve_scratch(device);
    // the status words start at zero (the sticky word is only ever cleared by the host)
    if (ensure(c, c->status, ST_WORDS * 8) != hipSuccess ||
        hipMemsetAsync(c->status.p, 0, c->status.bytes, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        if (c->status.p)
            (void)hipFree(c->status.p);
        (void)hipHostFree(c->status_host);
        (void)hipStreamDestroy(c->own);
        delete c;
        return SWR_ERR_HIP;
    }
    for (int i = 0; i <= S_COUNT; ++i)
        (void)hipEventCreate(&c->ev[i]);
    // optional: without it the kernels it would run run on the context stream
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        c->side = nullptr;
    }
    *out = c;
    return SWR_OK;
}

int swr_destroy(swr_context *c)
{
    if (!c)
        return SWR_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    DevBuf *bufs[] = {&c->idx, &c->attr0, &c->wverts, &c->app, &c->app_ex, &c->live, &c->grp,
                      &c->grp_ex,
                      &c->scan_tmp, &c->prims, &c->pk_a, &c->pv_a, &c->pk_b, &c->pv_b,
                      &c->ranges, &c->hist, &c->hist_ex, &c->hist_tmp, &c->status, &c->tgt[0],
                      &c->tgt[1], &c->tgt[2], &c->log, &c->logcount, &c->bigq, &c->clipq,
                      &c->clipq_cnt, &c->chunk_kept, &c->vwin, &c->vmask, &c->vmod_out,
                      &c->zscratch[0], &c->zscratch[1], &c->diff_table};
    for (DevBuf *b : bufs)
        if (b->p)
            (void)hipFree(b->p);
    for (int i = 0; i <= S_COUNT; ++i)
        if (c->ev[i])
            (void)hipEventDestroy(c->ev[i]);
    if (c->status_host)
        (void)hipHostFree(c->status_host);
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipStreamDestroy(c->side);
    }
    if (c->fork)
        (void)hipEventDestroy(c->fork);
    if (c->join)
        (void)hipEventDestroy(c->join);
    if (c->own)
        (void)hipStreamDestroy(c->own);
    delete c;
    return SWR_OK;
}

const char *swr_last_error(const swr_context *c) { return c ? c->err.c_str() : "null context"; }

int swr_set_stream(swr_context *c, void *stream)
{
    if (!c)
        return SWR_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    HIP_TRY(c, hipStreamSynchronize(c->stream));  // work already queued stays ordered
    c->stream = (hipStream_t)stream;
    return SWR_OK;
}

int swr_synchronize(swr_context *c)
{
    if (!c)
        return SWR_ERR_INVALID_ARG;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->async_pending) {
        c->async_pending = false;
        // each asynchronous draw folded its own failures into the sticky word
        int rc = take_sticky(c);
        if (rc)
            return rc;
        collect_stats(c);
    }
    return SWR_OK;
}

int swr_set_raster_mode(swr_context *c, int mode)
{
    if (!c || mode < SWR_RASTER_SPAN || mode > SWR_RASTER_ADAPTIVE)
        return fail(c, SWR_ERR_INVALID_ARG, "bad raster mode %d", mode);
    c->raster_mode = mode;
    return SWR_OK;
}

int swr_set_scissor_rect(swr_context *c, int x, int y, int w, int h)
{
    if (!c)
        return SWR_ERR_INVALID_ARG;
    c->sc[0] = x;
    c->sc[1] = y;
    c->sc[2] = w;
    c->sc[3] = h;
    return SWR_OK;
}

int swr_set_pixel_shader(swr_context *c, const swr_pixel_shader_desc *d)
{
    if (!c || !d)
        return fail(c, SWR_ERR_INVALID_ARG, "null pixel shader");
    if (d->program < SWR_PS_NULL || d->program > SWR_PS_MODULE)
        return fail(c, SWR_ERR_UNSUPPORTED, "pixel program %d is not built", d->program);
    if (d->program == SWR_PS_TEXTURE) {
        const swr_texture *t = static_cast<const swr_texture *>(d->texture);
        if (!t || t->ctx != c)
            return fail(c, SWR_ERR_INVALID_ARG, "the texture program needs a texture of this context");
    }
    swr_pixel_shader_desc ps = *d;
    if (d->program == SWR_PS_MODULE) {
        // a program built outside the library: it must have been compiled against this
        // library's raster argument block; its statics replace the descriptor's counts
        const swr_pixel_module *m = d->module;
        if (!m)
            return fail(c, SWR_ERR_INVALID_ARG, "SWR_PS_MODULE needs a module");
        if (m->abi != SWR_PIXEL_MODULE_ABI || m->args_bytes != (uint32_t)sizeof(RasterArgs))
            return fail(c, SWR_ERR_UNSUPPORTED,
                        "pixel module %s was built for another library (abi %u, %u-byte "
                        "arguments; this library: abi %u, %zu bytes)",
                        m->name ? m->name : "?", m->abi, m->args_bytes, SWR_PIXEL_MODULE_ABI,
                        sizeof(RasterArgs));
        if (!m->launch_triangles || !m->launch_fragments)
            return fail(c, SWR_ERR_INVALID_ARG, "pixel module without launchers");
        if (m->avar_count < 0 || m->avar_count > SWR_MAX_AVARS || m->pvar_count < 0 ||
            m->pvar_count > SWR_MAX_PVARS)
            return fail(c, SWR_ERR_UNSUPPORTED, "pixel module with %d avars / %d pvars",
                        m->avar_count, m->pvar_count);
        ps.interpolate_z = m->interpolate_z ? 1 : 0;
        ps.interpolate_w = m->interpolate_w ? 1 : 0;
        ps.avar_count = m->avar_count;
        ps.pvar_count = m->pvar_count;
    }
    ParamLayout L;
    if (ps.program != SWR_PS_NULL && !raster_layout(ps, L))
        return fail(c, SWR_ERR_UNSUPPORTED, "pixel program %d cannot run with these counts",
                    ps.program);
    if (ps.program == SWR_PS_MODULE) {
        c->module = *d->module;
        ps.module = &c->module;
    }
    c->ps = ps;
    return SWR_OK;
}

int swr_set_viewport(swr_context *c, int x, int y, int w, int h)
{
    if (!c)
        return SWR_ERR_INVALID_ARG;
    c->vp[0] = x;
    c->vp[1] = y;
    c->vp[2] = w;
    c->vp[3] = h;
    return SWR_OK;
}

int swr_set_depth_range(swr_context *c, float n, float f)
{
    if (!c)
        return SWR_ERR_INVALID_ARG;
    c->dn = n;
    c->df = f;
    return SWR_OK;
}

int swr_set_cull_mode(swr_context *c, int mode)
{
    if (!c || mode < SWR_CULL_NONE || mode > SWR_CULL_CW)
        return fail(c, SWR_ERR_INVALID_ARG, "bad cull mode %d", mode);
    c->cull = mode;
    return SWR_OK;
}

int swr_set_vertex_shader(swr_context *c, const swr_vertex_shader_desc *d)
{
    if (!c || !d)
        return fail(c, SWR_ERR_INVALID_ARG, "null vertex shader");
    if (d->program < SWR_VS_DUMMY || d->program > SWR_VS_MODULE)
        return fail(c, SWR_ERR_UNSUPPORTED, "vertex program %d is not built", d->program);
    if (d->program == SWR_VS_MODULE) {
        const swr_vertex_module *m = d->module;
        if (!m || m->abi != SWR_VERTEX_MODULE_ABI || !m->shade)
            return fail(c, SWR_ERR_INVALID_ARG,
                        "vertex module missing or built for another library version");
        if (m->attrib_count < 1 || m->attrib_count > SWR_MAX_VERTEX_ATTRIBS ||
            m->avar_count < 0 || m->avar_count > SWR_MAX_AVARS || m->pvar_count < 0 ||
            m->pvar_count > SWR_MAX_PVARS)
            return fail(c, SWR_ERR_INVALID_ARG, "vertex module counts out of range");
        if (d->uniforms && !is_device_ptr(d->uniforms))
            return fail(c, SWR_ERR_INVALID_ARG, "vertex module uniforms must be device memory");
        c->vmodule = *m;
        c->vs = *d;
        c->vs.module = &c->vmodule;
        c->vs.attrib_count = m->attrib_count;
        c->vs.avar_count = m->avar_count;
        c->vs.pvar_count = m->pvar_count;
        return SWR_OK;
    }
    if (d->attrib_count < 0 || d->attrib_count > SWR_MAX_VERTEX_ATTRIBS || d->avar_count < 0 ||
        d->avar_count > SWR_MAX_AVARS || d->pvar_count < 0 || d->pvar_count > SWR_MAX_PVARS)
        return fail(c, SWR_ERR_INVALID_ARG, "vertex shader counts out of range");
    c->vs = *d;
    return SWR_OK;
}

int swr_set_vertex_attrib_pointer(swr_context *c, int index, int stride, const void *buffer,
                                  size_t buffer_bytes)
{
    if (!c || index < 0 || index >= SWR_MAX_VERTEX_ATTRIBS || stride < 0)
        return fail(c, SWR_ERR_INVALID_ARG, "bad vertex attribute %d", index);
    c->attr[index].ptr = buffer;
    c->attr[index].stride = stride;
    c->attr[index].bytes = buffer_bytes;
    return SWR_OK;
}

// A draw whose vertex program is a module: every vertex of the attribute buffers is shaded
// once into rows {x, y, z, w, avar, pvar} (VertexProcessor.cpp:134-150 per vertex; the
// reference's VertexCache shades each referenced index once, and vertex programs are pure),
// then the draw reads those rows as SWR_VS_RAW.  The context's vertex state is restored
// when the guard goes out of scope.
struct VertexModuleDraw {
    swr_context *c = nullptr;
    swr_vertex_shader_desc vs{};
    swr_context::Attr a0{};
    ~VertexModuleDraw()
    {
        if (c) {
            c->vs = vs;
            c->attr[0] = a0;
        }
    }
};

int shade_vertex_module(swr_context *c, VertexModuleDraw &g)
{
    const swr_vertex_module &m = c->vmodule;
    uint64_t nv = ~0ull;
    const void *ptrs[SWR_MAX_VERTEX_ATTRIBS] = {};
    int64_t strides[SWR_MAX_VERTEX_ATTRIBS] = {};
    for (int i = 0; i < m.attrib_count; ++i) {
        const swr_context::Attr &A = c->attr[i];
        if (!A.ptr || !is_device_ptr(A.ptr) || !A.bytes || A.stride <= 0)
            return fail(c, SWR_ERR_INVALID_ARG,
                        "vertex module %s: attribute %d must be device memory with its size "
                        "and a positive stride",
                        m.name ? m.name : "", i);
        nv = std::min<uint64_t>(nv, A.bytes / (uint64_t)A.stride);
        ptrs[i] = A.ptr;
        strides[i] = A.stride;
    }
    if (nv >= (1ull << 31))
        return fail(c, SWR_ERR_CAPACITY, "vertex module over %llu vertices",
                    (unsigned long long)nv);
    const uint64_t row = 4ull * (4 + m.avar_count + m.pvar_count);
    HIP_TRY(c, ensure(c, c->vmod_out, std::max<uint64_t>(nv, 1) * row));
    const int rc = m.shade(ptrs, strides, (uint32_t)nv, static_cast<float *>(c->vmod_out.p),
                           c->vs.uniforms, c->stream);
    if (rc != (int)hipSuccess)
        return fail(c, SWR_ERR_HIP, "vertex module %s: %s", m.name ? m.name : "",
                    hipGetErrorString((hipError_t)rc));
    g.c = c;
    g.vs = c->vs;
    g.a0 = c->attr[0];
    c->vs.program = SWR_VS_RAW;
    c->vs.attrib_count = 1;
    c->attr[0].ptr = c->vmod_out.p;
    c->attr[0].stride = (int)row;
    c->attr[0].bytes = (size_t)(nv * row);
    return SWR_OK;
}

int swr_draw_elements(swr_context *c, int mode, size_t count, const int32_t *indices)
{
    if (!c || (count && !indices))
        return fail(c, SWR_ERR_INVALID_ARG, "null indices");
    if (mode != SWR_DRAW_TRIANGLE && mode != SWR_DRAW_POINT && mode != SWR_DRAW_LINE)
        return fail(c, SWR_ERR_INVALID_ARG, "bad draw mode %d", mode);
    (void)hipSetDevice(c->device);
    VertexModuleDraw vmd;
    if (c->vs.program == SWR_VS_MODULE) {
        const int rc = shade_vertex_module(c, vmd);
        if (rc)
            return rc;
    }
    if (mode != SWR_DRAW_TRIANGLE)
        return draw_lines_points(c, mode, indices, count, false, nullptr, 0);
    return draw_triangles(c, indices, count, false, nullptr, 0, true);
}

int swr_draw_elements_async(swr_context *c, int mode, size_t count, const int32_t *indices)
{
    if (!c || (count && !indices))
        return fail(c, SWR_ERR_INVALID_ARG, "null indices");
    if (mode != SWR_DRAW_TRIANGLE)
        return fail(c, SWR_ERR_UNSUPPORTED, "async draws take triangles");
    (void)hipSetDevice(c->device);
    VertexModuleDraw vmd;
    if (c->vs.program == SWR_VS_MODULE) {
        const int rc = shade_vertex_module(c, vmd);
        if (rc)
            return rc;
    }
    return draw_triangles(c, indices, count, false, nullptr, 0, false);
}

int swr_draw_elements_noalloc(swr_context *c, int mode, size_t count, const int32_t *indices)
{
    if (!c || (count && !indices))
        return fail(c, SWR_ERR_INVALID_ARG, "null indices");
    if (mode != SWR_DRAW_TRIANGLE)
        return fail(c, SWR_ERR_UNSUPPORTED, "the no-allocation draw takes triangles");
    if (c->vs.program == SWR_VS_MODULE)
        return fail(c, SWR_ERR_UNSUPPORTED,
                    "the no-allocation draw does not take vertex modules (their output rows "
                    "need device memory)");
    (void)hipSetDevice(c->device);
    return draw_triangles(c, indices, count, false, nullptr, 0, false, nullptr, true);
}

int swr_reserve(swr_context *c, size_t primitives, size_t tile_pairs)
{
    if (!c)
        return SWR_ERR_INVALID_ARG;
    c->prim_cap = std::max<uint64_t>(c->prim_cap, primitives);
    c->pair_cap = std::max<uint64_t>(c->pair_cap, tile_pairs);
    return SWR_OK;
}

int swr_draw_triangle_list(swr_context *c, const swr_rasterizer_vertex *v, size_t nv,
                           const int32_t *idx, size_t n)
{
    if (!c || (n && (!v || !idx)))
        return fail(c, SWR_ERR_INVALID_ARG, "null vertices or indices");
    (void)hipSetDevice(c->device);
    return draw_triangles(c, idx, n, true, v, nv, true);
}

int swr_draw_line_list(swr_context *c, const swr_rasterizer_vertex *v, size_t nv,
                       const int32_t *idx, size_t n)
{
    if (!c || (n && (!v || !idx)))
        return fail(c, SWR_ERR_INVALID_ARG, "null vertices or indices");
    (void)hipSetDevice(c->device);
    return draw_lines_points(c, SWR_DRAW_LINE, idx, n, true, v, nv);
}

int swr_draw_point_list(swr_context *c, const swr_rasterizer_vertex *v, size_t nv,
                        const int32_t *idx, size_t n)
{
    if (!c || (n && (!v || !idx)))
        return fail(c, SWR_ERR_INVALID_ARG, "null vertices or indices");
    (void)hipSetDevice(c->device);
    return draw_lines_points(c, SWR_DRAW_POINT, idx, n, true, v, nv);
}

int swr_batch_create(swr_context *c, const swr_scene *sc, size_t n, swr_batch **out)
{
    if (!c || !out || (n && !sc))
        return fail(c, SWR_ERR_INVALID_ARG, "null batch arguments");
    *out = nullptr;
    if (n > (1u << 20))
        return fail(c, SWR_ERR_CAPACITY, "too many scenes (%zu)", n);
    (void)hipSetDevice(c->device);
    std::vector<SceneDev> tab(n);
    std::vector<uint32_t> bmap;
    uint64_t bins[3] = {0, 0, 0}, tris = 0;
    uint32_t batches = 0;
    unsigned tmask = 7;
    for (size_t i = 0; i < n; ++i) {
        const swr_scene &s = sc[i];
        const uint64_t T = s.index_count / 3;
        if (s.cull_mode < SWR_CULL_NONE || s.cull_mode > SWR_CULL_CW ||
            s.raster_mode < SWR_RASTER_SPAN || s.raster_mode > SWR_RASTER_ADAPTIVE)
            return fail(c, SWR_ERR_INVALID_ARG, "scene %zu: bad cull or raster mode", i);
        if (s.width < 0 || s.height < 0 || s.pitch < s.width)
            return fail(c, SWR_ERR_INVALID_ARG, "scene %zu: bad target size", i);
        if (T && (!is_device_ptr(s.indices) || !is_device_ptr(s.attrib0) ||
                  s.attrib0_stride < 0 || s.attrib0_bytes == 0))
            return fail(c, SWR_ERR_INVALID_ARG,
                        "scene %zu: indices and attribute 0 must be device buffers with a size",
                        i);
        void *const t[3] = {s.target0, s.target1, s.target2};
        for (int k = 0; k < 3; ++k) {
            if (!t[k])
                tmask &= ~(1u << k);
            else if (!is_device_ptr(t[k]))
                return fail(c, SWR_ERR_INVALID_ARG, "scene %zu: target%d is not device memory",
                            i, k);
        }
        SceneDev &d = tab[i];
        std::memset(&d, 0, sizeof d);
        d.indices = s.indices;
        d.attr0 = (const uint8_t *)s.attrib0;
        d.attr0_bytes = s.attrib0_bytes;
        d.t0 = s.target0;
        d.t1 = s.target1;
        d.t2 = s.target2;
        std::memcpy(d.mvp, s.mvp, sizeof d.mvp);
        // VertexProcessor.cpp:50-53 and 373, as for a single draw
        volatile float px = (float)s.viewport[2] / 2.0f, py = (float)s.viewport[3] / 2.0f;
        d.vp_px = px;
        d.vp_py = py;
        d.vp_ox = (float)s.viewport[0] + px;
        d.vp_oy = (float)s.viewport[1] + py;
        volatile float dzs = 0.5f * (s.depth_far - s.depth_near),
                       dzb = 0.5f * (s.depth_near + s.depth_far);
        d.dz_scale = dzs;
        d.dz_bias = dzb;
        d.stride0 = s.attrib0_stride;
        d.cull_mode = s.cull_mode;
        d.raster_mode = s.raster_mode;
        d.sc_minx = s.scissor[0];
        d.sc_miny = s.scissor[1];
        d.sc_maxx = s.scissor[0] + s.scissor[2];
        d.sc_maxy = s.scissor[1] + s.scissor[3];
        d.width = s.width;
        d.height = s.height;
        d.pitch = s.pitch;
        for (int l = 0; l < 3; ++l) {
            const int bin = TILE << l;
            d.bins_x[l] = (s.width + bin - 1) / bin;
            d.bins_y[l] = (s.height + bin - 1) / bin;
            d.bin0[l] = (uint32_t)bins[l];
        }
        d.tris = (uint32_t)T;
        d.batch0 = batches;
        const uint32_t nb = (uint32_t)((T + BATCH - 1) / BATCH);
        for (uint32_t b = 0; b < nb; ++b)
            bmap.push_back((uint32_t)i);
        batches += nb;
        for (int l = 0; l < 3; ++l)
            bins[l] += (uint64_t)d.bins_x[l] * (uint64_t)d.bins_y[l];
        tris += T;
        if (T > (1u << 31) / 8 || (uint64_t)batches * BATCH > (1u << 31) / 8 ||
            bins[0] >= 0xFFFFFFF0ull)
            return fail(c, SWR_ERR_CAPACITY, "batch too large at scene %zu", i);
    }
    swr_batch *b = new swr_batch();
    b->ctx = c;
    b->nscenes = (uint32_t)n;
    b->batches = batches;
    b->vtris = (uint64_t)batches * BATCH;
    b->tris = tris;
    for (int l = 0; l < 3; ++l)
        b->bins[l] = bins[l];
    b->tmask = n ? tmask : 0;
    for (size_t i = 0; i < n; ++i)
        b->spans = b->spans || sc[i].raster_mode != SWR_RASTER_BLOCK;
    hipError_t e = ensure(c, b->scenes, std::max<size_t>(1, n) * sizeof(SceneDev));
    if (e == hipSuccess)
        e = ensure(c, b->bmap, std::max<size_t>(1, bmap.size()) * 4);
    if (e == hipSuccess && n)
        e = hipMemcpy(b->scenes.p, tab.data(), n * sizeof(SceneDev), hipMemcpyHostToDevice);
    if (e == hipSuccess && !bmap.empty())
        e = hipMemcpy(b->bmap.p, bmap.data(), bmap.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        swr_batch_destroy(b);
        return fail(c, SWR_ERR_HIP, "batch upload failed: %s", hipGetErrorString(e));
    }
    *out = b;
    return SWR_OK;
}

int swr_batch_draw(swr_context *c, const swr_batch *b, int async)
{
    if (!c || !b || b->ctx != c)
        return fail(c, SWR_ERR_INVALID_ARG, "batch does not belong to this context");
    if (c->vs.program == SWR_VS_MODULE)
        return fail(c, SWR_ERR_UNSUPPORTED, "batched draws do not take vertex modules");
    (void)hipSetDevice(c->device);
    return draw_triangles(c, nullptr, 0, false, nullptr, 0, async == 0, b);
}

int swr_batch_destroy(swr_batch *b)
{
    if (!b)
        return SWR_ERR_INVALID_ARG;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->scenes.p)
        (void)hipFree(b->scenes.p);
    if (b->bmap.p)
        (void)hipFree(b->bmap.p);
    delete b;
    return SWR_OK;
}

int swr_texture_create(swr_context *c, const uint32_t *texels, int width, int height, int pitch,
                       int max_anisotropy, swr_texture **out)
{
    if (!c || !texels || !out || width < 1 || height < 1 || pitch < width)
        return fail(c, SWR_ERR_INVALID_ARG, "bad texture arguments");
    *out = nullptr;
    (void)hipSetDevice(c->device);
    // level 0 on the host (device texels are copied down: the chain is built once)
    std::vector<uint32_t> base((size_t)width * height);
    if (is_device_ptr(texels)) {
        HIP_TRY(c, hipMemcpy2D(base.data(), (size_t)width * 4, texels, (size_t)pitch * 4,
                               (size_t)width * 4, height, hipMemcpyDeviceToHost));
    } else {
        for (int y = 0; y < height; ++y)
            std::memcpy(&base[(size_t)y * width], texels + (size_t)y * pitch, (size_t)width * 4);
    }
    TexDev td{};
    // std::clamp(maxAnisotropy, 1, MAX_ANISOTROPY), Texture.h:15-16
    td.max_aniso = max_anisotropy < 1 ? 1 : (max_anisotropy > 16 ? 16 : max_anisotropy);
    std::vector<uint32_t> all;
    for (uint32_t &p : base)
        p &= 0xFFFFFFu;
    std::vector<uint32_t> cur = base;
    int w = width, h = height, lv = 0;
    for (;;) {
        td.w[lv] = w;
        td.h[lv] = h;
        td.off[lv] = all.size();
        all.insert(all.end(), cur.begin(), cur.end());
        if ((w <= 1 && h <= 1) || lv + 1 >= TEX_MAX_LEVELS)
            break;
        // Texture::generateMipmaps, Texture.h:268-287: 2x2 box filter, a missing neighbour
        // repeats the top-left texel, each channel averaged with >> 2
        const int nw = std::max(1, w / 2), nh = std::max(1, h / 2);
        std::vector<uint32_t> next((size_t)nw * nh);
        auto ch = [](uint32_t p, int k) { return (p >> (16 - 8 * k)) & 0xFFu; };
        for (int y = 0; y < nh; ++y)
            for (int x = 0; x < nw; ++x) {
                const uint32_t p00 = cur[(size_t)(2 * y) * w + 2 * x];
                const uint32_t p10 = 2 * x + 1 < w ? cur[(size_t)(2 * y) * w + 2 * x + 1] : p00;
                const uint32_t p01 = 2 * y + 1 < h ? cur[(size_t)(2 * y + 1) * w + 2 * x] : p00;
                const uint32_t p11 = (2 * x + 1 < w && 2 * y + 1 < h)
                                         ? cur[(size_t)(2 * y + 1) * w + 2 * x + 1]
                                         : p00;
                uint32_t o = 0;
                for (int k = 0; k < 3; ++k)
                    o |= (((ch(p00, k) + ch(p10, k) + ch(p01, k) + ch(p11, k)) >> 2) & 0xFFu)
                         << (16 - 8 * k);
                next[(size_t)y * nw + x] = o;
            }
        cur.swap(next);
        w = nw;
        h = nh;
        ++lv;
    }
    td.levels = lv + 1;
    swr_texture *t = new swr_texture;
    t->ctx = c;
    hipError_t e = ensure(c, t->texels, all.size() * 4);
    if (e == hipSuccess)
        e = ensure(c, t->desc, sizeof(TexDev));
    if (e == hipSuccess)
        e = hipMemcpy(t->texels.p, all.data(), all.size() * 4, hipMemcpyHostToDevice);
    td.texels = (const uint32_t *)t->texels.p;
    if (e == hipSuccess)
        e = hipMemcpy(t->desc.p, &td, sizeof td, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(t->texels.p);
        (void)hipFree(t->desc.p);
        delete t;
        return fail(c, SWR_ERR_HIP, "texture upload failed: %s", hipGetErrorString(e));
    }
    *out = t;
    return SWR_OK;
}

int swr_texture_destroy(swr_texture *t)
{
    if (!t)
        return SWR_OK;
    (void)hipSetDevice(t->ctx->device);
    (void)hipStreamSynchronize(t->ctx->stream);
    (void)hipFree(t->texels.p);
    (void)hipFree(t->desc.p);
    delete t;
    return SWR_OK;
}

int swr_set_tuning(swr_context *c, const swr_tuning *t)
{
    if (!c || !t)
        return fail(c, SWR_ERR_INVALID_ARG, "null tuning");
    auto pow2_in = [](int v, int lo, int hi) { return v == 0 || (v >= lo && v <= hi && !(v & (v - 1))); };
    if (!pow2_in(t->raster_waves, 1, 8) || !pow2_in(t->raster_group, 1, 8) ||
        !pow2_in(t->expand_chunk, 256, 4096) || t->xcd_order < -1 || t->xcd_order > 1 ||
        t->vertex_pass < -1 || t->vertex_pass > 1 || !pow2_in(t->bin_size, 8, 32) ||
        !pow2_in(t->waves_per_bin, 1, 4) || t->raster_queue < -1 ||
        t->raster_queue > 1)
        return fail(c, SWR_ERR_INVALID_ARG, "tuning value out of range");
    c->tune = *t;
    return SWR_OK;
}

int swr_get_tuning(swr_context *c, swr_tuning *t)
{
    if (!c || !t)
        return SWR_ERR_INVALID_ARG;
    *t = c->tune;
    return SWR_OK;
}

int swr_get_draw_stats(swr_context *c, swr_draw_stats *out)
{
    if (!c || !out)
        return SWR_ERR_INVALID_ARG;
    *out = c->stats;
    return SWR_OK;
}

int swr_set_profiling(swr_context *c, int enable)
{
    if (!c)
        return SWR_ERR_INVALID_ARG;
    c->profiling = enable != 0;
    return SWR_OK;
}

}  // extern "C"
