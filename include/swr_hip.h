/*
 * swr_hip.h — C ABI of the MI355X (gfx950) draw path.
 *
 * This is the boundary a caller of the reference's Rasterizer / VertexProcessor binds to.
 * Every entry point names the reference interface it replaces (path:line in
 * trenki2/SoftwareRenderer, src/renderer/...).  Plain C types only: pointers, sizes, ints.
 *
 * Conventions
 *   - Every function returns an int status: SWR_OK (0) or a negative SWR_ERR_* code.
 *     swr_last_error() gives a human-readable message for the last failure on a context.
 *   - Buffers may be host or device memory (detected with hipPointerGetAttributes).
 *     Host buffers are staged through device arenas owned by the context; device buffers
 *     are used in place.  Render targets in host memory are copied in before and out after
 *     the draw (the pixel programs read-modify-write them, as the reference's shaders do).
 *   - swr_draw_* are synchronous, like the reference: they return once the render targets
 *     hold the final values.  swr_draw_*_async enqueue on the context stream and return.
 *   - One context per host thread; a context is not thread-safe (the reference is not either).
 */
#ifndef SWR_HIP_H
#define SWR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------ */
#define SWR_OK                 0
#define SWR_ERR_INVALID_ARG   -1  /* bad enum, negative size, null where a pointer is required */
#define SWR_ERR_UNSUPPORTED   -2  /* program / varying count combination not built          */
#define SWR_ERR_HIP           -3  /* a HIP runtime call failed                                 */
#define SWR_ERR_NO_DEVICE     -4  /* no gfx950 device                                          */
#define SWR_ERR_CAPACITY      -5  /* an arena could not be grown / a fixed limit was exceeded  */
#define SWR_ERR_INDEX_RANGE   -6  /* an index addresses past a sized attribute buffer          */
#define SWR_ERR_CLIP_OVERFLOW -7  /* a clipped polygon exceeded SWR_MAX_CLIP_VERTS vertices     */

/* ---- limits (IRasterizer.h:33-39, VertexConfig.h:34) -------------------------------- */
#define SWR_BLOCK_SIZE         8
#define SWR_MAX_AVARS         16
#define SWR_MAX_PVARS         16
#define SWR_MAX_VERTEX_ATTRIBS 8
#define SWR_BATCH_PRIMITIVES 1024  /* VertexProcessor.cpp:110 */
#define SWR_MAX_CLIP_VERTS    32   /* polygon vertices after 6 planes; 3 + 6*4 is the convex bound */
/* DDA steps of one line (Rasterizer.h:177-179: max(|dx|, |dy|) of the truncated endpoints).
 * A longer line fails the draw with SWR_ERR_CAPACITY: the reference would walk it step by
 * step (NaN or huge window coordinates give ~2^31 steps). */
#define SWR_MAX_LINE_STEPS   (1 << 20)

/* ---- enums (same values as the reference's enum classes) ---------------------------- */
/* RasterMode, Rasterizer.h:45-49 */
#define SWR_RASTER_SPAN      0
#define SWR_RASTER_BLOCK     1
#define SWR_RASTER_ADAPTIVE  2
/* DrawMode, VertexProcessor.h:42-46 */
#define SWR_DRAW_POINT       0
#define SWR_DRAW_LINE        1
#define SWR_DRAW_TRIANGLE    2
/* CullMode, VertexProcessor.h:49-53 */
#define SWR_CULL_NONE        0
#define SWR_CULL_CCW         1
#define SWR_CULL_CW          2

/* ---- RasterizerVertex (IRasterizer.h:42-53): 36 floats, 144 bytes -------------------- */
typedef struct swr_rasterizer_vertex {
    float x, y, z, w;
    float avar[SWR_MAX_AVARS];
    float pvar[SWR_MAX_PVARS];
} swr_rasterizer_vertex;

/* ---- vertex programs ------------------------------------------------------------------
 * The reference binds a C++ static function (VertexProcessor.h:77-85).  Across a C ABI a
 * shader is a device program id plus its uniforms, or a vertex module (SWR_VS_MODULE,
 * below); the statics AttribCount / AVarCount / PVarCount keep their meaning
 * (VertexShaderBase.h:37-43): the clipper interpolates only the first avar_count /
 * pvar_count varyings (PolyClipper.h:42-45).
 *
 *   SWR_VS_DUMMY          writes nothing (DummyVertexShader, VertexShaderBase.h:53)
 *   SWR_VS_RAW            attrib 0 holds {x,y,z,w, avar[avar_count], pvar[pvar_count]} floats
 *   SWR_VS_POS_RGB        attrib 0 = {x,y,z,r,g,b}: out = (x,y,z,1), avar[0..2] = rgb
 *                         (Benchmark.cpp:32-49, VertexProcessorTest.cpp:59-76)
 *   SWR_VS_MVP_POS_RGB_UV attrib 0 = {px,py,pz,r,g,b,u,v}: out = M*(p,1), avar = rgb, pvar = uv
 *   SWR_VS_MVP_POS_UV     attrib 0 = {px,py,pz,nx,ny,nz,u,v} (ObjData::VertexArrayData):
 *                         out = M*(p,1), pvar = uv (Box.cpp:67-88)
 * M is row-major (mvp[4*r+c]); out[r] = (((0 + M[r][0]*x) + M[r][1]*y) + M[r][2]*z) + M[r][3]*1,
 * the sequential dot product of vector_math.h:55-60, 548-556, with no contraction.
 */
#define SWR_VS_DUMMY           0
#define SWR_VS_RAW             1
#define SWR_VS_POS_RGB         2
#define SWR_VS_MVP_POS_RGB_UV  3
#define SWR_VS_MVP_POS_UV      4
#define SWR_VS_MODULE          5

/* A vertex shader compiled outside the library (include/swr/vertex_program.h): the
 * reference's VertexShaderBase<Derived> with processVertex (VertexShaderBase.h:33-51) as a
 * __device__ function that also receives the program's uniforms.  The module's entry point
 * (SWR_VERTEX_MODULE) fills this struct, which swr_set_vertex_shader copies.  `shade` runs
 * processVertex for vertices 0 .. vertex_count-1, vertex v reading attribute i at
 * attribs[i] + strides[i] * v (VertexProcessor.cpp:134-150), and writes row v of `out`:
 * {x, y, z, w, avar[avar_count], pvar[pvar_count]} floats.  A draw with a vertex module
 * shades every vertex of its attribute buffers once (the buffers must be device memory with
 * their sizes given to swr_set_vertex_attrib_pointer; the vertex count is the number of
 * whole strides in the smallest of them) and then draws those rows as SWR_VS_RAW would:
 * the same results as the reference, whose VertexCache shades each referenced index once
 * (VertexProcessor.cpp:86-117; vertex programs are pure functions of their inputs).  The
 * asynchronous draw enqueues the shading on the context stream; swr_draw_elements_noalloc
 * and batched draws do not take vertex modules.  The module's shared library must stay
 * loaded while the shader is bound; uniforms may be null for a program that reads none. */
#define SWR_VERTEX_MODULE_ABI  1u
typedef struct swr_vertex_module {
    uint32_t abi;              /* SWR_VERTEX_MODULE_ABI                                 */
    int32_t  attrib_count;     /* VertexShaderBase::AttribCount (1..SWR_MAX_VERTEX_ATTRIBS) */
    int32_t  avar_count;       /* VertexShaderBase::AVarCount (0..SWR_MAX_AVARS)        */
    int32_t  pvar_count;       /* VertexShaderBase::PVarCount (0..SWR_MAX_PVARS)        */
    int    (*shade)(const void *const *attribs, const int64_t *strides, uint32_t vertex_count,
                    float *out, const void *uniforms, void *stream);
    const char *name;
} swr_vertex_module;

typedef struct swr_vertex_shader_desc {
    int32_t program;
    int32_t attrib_count;      /* VertexShaderBase::AttribCount */
    int32_t avar_count;        /* VertexShaderBase::AVarCount   */
    int32_t pvar_count;        /* VertexShaderBase::PVarCount   */
    float   mvp[16];           /* uniform block of the MVP programs */
    const swr_vertex_module *module;  /* SWR_VS_MODULE: the module (its counts replace the
                                         three above)                                    */
    const void *uniforms;      /* SWR_VS_MODULE: the program's uniforms, device memory  */
} swr_vertex_shader_desc;

/* ---- pixel programs -------------------------------------------------------------------
 * The reference binds PixelShader::drawPixel (Rasterizer.h:90-96, PixelShaderBase.h:114-119).
 * The statics InterpolateZ / InterpolateW / AVarCount / PVarCount keep their meaning
 * (PixelShaderBase.h:44-53).  Targets are width x height pixels with a row pitch in pixels;
 * pixels outside the target are never written.
 *
 *   SWR_PS_NULL           no-op (NullPixelShader, PixelShaderBase.h:139)
 *   SWR_PS_DEPTH_CHECKER  target1 f32 depth, target0 u32 RGBA8:
 *                         if (z < depth) { depth = z; color = pack(avar[0..2] * checker(pvar[0..1])) }
 *                         checker = ((floor(u*cells) + floor(v*cells)) odd) ? 1 : 0.5
 *   SWR_PS_RGB            target0 u32: (int)(r*255) << 16 | (int)(g*255) << 8 | (int)(b*255)
 *                         (RasterizerTest.cpp:38-48, VertexProcessorTest.cpp:39-49)
 *   SWR_PS_COVERAGE       target0 u32 = 1 (Benchmark.cpp:22-25)
 *   SWR_PS_HASH           target0 u32 order-sensitive hash, target1 u32 invocation count:
 *                         h = mix(h, key), then mix over the bits of z (if InterpolateZ),
 *                         w and invw (if InterpolateW), avar[0..A), pvar[0..P);
 *                         mix(h, x) = (h ^ x) * 16777619
 *   SWR_PS_LOG            target0 u32 invocation count; appends one record per invocation to
 *                         log (SWR_LOG_WORDS u32 words each) at an atomic slot:
 *                         {x, y, key, seq, z, w, invw, avar[16], pvar[16]} (floats as bits);
 *                         *log_count receives the number of invocations even past capacity.
 *   SWR_PS_VISIBILITY     target0 u32 source triangle + 1 (0 = none), target1 f32 depth:
 *                         if (z < depth) { depth = z; id = triangle + 1 }, where triangle is
 *                         the input triangle (index triple) the primitive was clipped from;
 *                         the visibility buffer the differentiable path reads
 *   SWR_PS_TEXTURE        target0 u32 0x00RRGGBB: the textured shader of the Box demo
 *                         (Box.cpp:39-53, TextureExample.h:17-29): the derivatives of
 *                         pvar[0..1] from the triangle's planes (PixelData.h:127-145), then
 *                         Texture::sample (Texture.h:35-118: wrap, anisotropic footprint, up
 *                         to max_anisotropy trilinear taps, Texture.h:121-200) of `texture`.
 *                         Needs PVarCount >= 2 (u, v) and triangles (a point or line has no
 *                         triangle planes: SWR_ERR_UNSUPPORTED).  log2 of the level of
 *                         detail is correctly rounded (the reference calls the C library's
 *                         log2f, which may round the other way by one ulp)
 *   SWR_PS_MODULE         a program compiled outside the library: `module` (below) with
 *                         its own drawPixel, built against include/swr/pixel_program.h;
 *                         the module's statics replace the four counts of the descriptor;
 *                         `uniforms` (device memory) is handed to every drawPixel call.
 *                         No target is required: the program writes what it writes.
 * "key" identifies the primitive in draw order: (batch << 15) + slot, where slot is the
 * triangle's position in the batch's index list after clipping (VertexProcessor.cpp:254-261);
 * for swr_draw_triangle_list it is the triangle's position in the list.
 */
#define SWR_PS_NULL            0
#define SWR_PS_DEPTH_CHECKER   1
#define SWR_PS_RGB             2
#define SWR_PS_COVERAGE        3
#define SWR_PS_HASH            4
#define SWR_PS_LOG             5
#define SWR_PS_VISIBILITY      6
#define SWR_PS_TEXTURE         7
#define SWR_PS_MODULE          8

/* A pixel program compiled outside the library: the device counterpart of a class derived
 * from PixelShaderBase (PixelShaderBase.h:36-119), whose static drawPixel(const PixelData &)
 * the rasterizer calls for every covered pixel in the reference's order (for each pixel, its
 * primitives in draw order).  include/swr/pixel_program.h builds it from such a class; the
 * module's entry point (SWR_PIXEL_MODULE) fills this struct, which swr_set_pixel_shader
 * copies.  The two launchers enqueue the raster stage of a triangle draw and of a point or
 * line draw; `args` is the library's raster argument block, of args_bytes bytes.  The
 * module's shared library must stay loaded while the shader is bound. */
#define SWR_PIXEL_MODULE_ABI   1u
typedef struct swr_pixel_module {
    uint32_t abi;              /* SWR_PIXEL_MODULE_ABI                                  */
    uint32_t args_bytes;       /* size of the raster argument block it was built with   */
    int32_t  interpolate_z;    /* PixelShaderBase::InterpolateZ                         */
    int32_t  interpolate_w;    /* PixelShaderBase::InterpolateW                         */
    int32_t  avar_count;       /* PixelShaderBase::AVarCount (0..SWR_MAX_AVARS)         */
    int32_t  pvar_count;       /* PixelShaderBase::PVarCount (0..SWR_MAX_PVARS)         */
    int    (*launch_triangles)(const void *args, void *stream, int waves);
    int    (*launch_fragments)(const void *args, void *stream);
    const char *name;
} swr_pixel_module;

#define SWR_LOG_WORDS          (7 + SWR_MAX_AVARS + SWR_MAX_PVARS)
#define SWR_HASH_SEED          2166136261u

typedef struct swr_pixel_shader_desc {
    int32_t program;
    int32_t interpolate_z;     /* PixelShaderBase::InterpolateZ */
    int32_t interpolate_w;     /* PixelShaderBase::InterpolateW */
    int32_t avar_count;        /* PixelShaderBase::AVarCount    */
    int32_t pvar_count;        /* PixelShaderBase::PVarCount    */
    int32_t width, height;     /* target size in pixels        */
    int32_t pitch;             /* row pitch in pixels           */
    void   *target0;
    void   *target1;
    void   *target2;
    uint32_t *log;             /* SWR_PS_LOG records            */
    uint64_t  log_capacity;    /* in records                    */
    uint64_t *log_count;       /* device or host u64            */
    float   checker_cells;     /* SWR_PS_DEPTH_CHECKER uniform  */
    int32_t reserved;
    const void *texture;       /* SWR_PS_TEXTURE: a swr_texture */
    const swr_pixel_module *module;  /* SWR_PS_MODULE                 */
    const void *uniforms;      /* SWR_PS_MODULE: device memory handed to drawPixel */
} swr_pixel_shader_desc;

/* ---- context ------------------------------------------------------------------------ */
typedef struct swr_context swr_context;

/* Creates a context on HIP device `device` with the reference's default state:
 * Span, scissor (0,0,0,0), NullPixelShader (Rasterizer.h:67-72); cull CW, depth (0,1),
 * DummyVertexShader (VertexProcessor.cpp:29-35); viewport (0,0,0,0). */
int  swr_create(swr_context **out, int device);
int  swr_destroy(swr_context *ctx);
const char *swr_last_error(const swr_context *ctx);
/* Use `stream` (a hipStream_t; NULL is the null stream) for all later work.  A context
 * starts on a non-blocking stream of its own. */
int  swr_set_stream(swr_context *ctx, void *stream);
/* Waits for the context stream.  Reports, once, the first failure of any asynchronous draw
 * since the last report (SWR_ERR_CAPACITY for an arena overflow: that draw drew nothing;
 * SWR_ERR_INDEX_RANGE, SWR_ERR_CLIP_OVERFLOW); a synchronous draw reports them too. */
int  swr_synchronize(swr_context *ctx);

/* ---- Rasterizer (Rasterizer.h:52-141) ------------------------------------------------ */
int  swr_set_raster_mode(swr_context *ctx, int mode);                        /* Rasterizer.h:75-78 */
int  swr_set_scissor_rect(swr_context *ctx, int x, int y, int w, int h);     /* Rasterizer.h:81-87 */
int  swr_set_pixel_shader(swr_context *ctx, const swr_pixel_shader_desc *d);/* Rasterizer.h:90-96 */
/* IRasterizer seam: window-space vertices; triangles whose first index is -1 are skipped.
 * IRasterizer.h:70 / Rasterizer.h:134-141 (drawTriangle, :111-114, is a list of one). */
int  swr_draw_triangle_list(swr_context *ctx, const swr_rasterizer_vertex *vertices,
                            size_t vertex_count, const int32_t *indices, size_t index_count);
/* IRasterizer.h:66 / Rasterizer.h:125-132: index pairs; a pair whose first index is -1 is
 * skipped.  Lines are DDA-walked on the device (Rasterizer.h:175-222). */
int  swr_draw_line_list(swr_context *ctx, const swr_rasterizer_vertex *vertices,
                        size_t vertex_count, const int32_t *indices, size_t index_count);
/* IRasterizer.h:62 / Rasterizer.h:116-123: one pixel per point that passes the scissor
 * test (Rasterizer.h:149-158); index -1 is skipped. */
int  swr_draw_point_list(swr_context *ctx, const swr_rasterizer_vertex *vertices,
                         size_t vertex_count, const int32_t *indices, size_t index_count);

/* ---- VertexProcessor (VertexProcessor.h:56-91) ---------------------------------------- */
int  swr_set_viewport(swr_context *ctx, int x, int y, int w, int h);         /* VertexProcessor.cpp:43-54 */
int  swr_set_depth_range(swr_context *ctx, float n, float f);                /* VertexProcessor.cpp:56-60 */
int  swr_set_cull_mode(swr_context *ctx, int mode);                          /* VertexProcessor.cpp:62-65 */
int  swr_set_vertex_shader(swr_context *ctx, const swr_vertex_shader_desc *d);/* VertexProcessor.h:77-85 */
/* VertexProcessor.cpp:67-76.  buffer_bytes = 0 means "unknown" (no index range check, as in
 * the reference); otherwise indices addressing past the buffer fail with SWR_ERR_INDEX_RANGE. */
int  swr_set_vertex_attrib_pointer(swr_context *ctx, int index, int stride,
                                   const void *buffer, size_t buffer_bytes);
/* VertexProcessor.cpp:78-120.  count indices; triangles use count/3 complete triples, lines
 * count/2 pairs, points every index.  Lines and points: clipLines / clipPoints
 * (VertexProcessor.cpp:152-215), then the line and point lists above. */
int  swr_draw_elements(swr_context *ctx, int mode, size_t count, const int32_t *indices);
/* Same, enqueued on the context stream without waiting.  Device buffers only: indices,
 * vertex attribute 0, render targets and log buffers in host memory fail with
 * SWR_ERR_INVALID_ARG.  The context's arenas must already be large enough (swr_reserve, or
 * a previous synchronous draw of the same scene); otherwise the draw leaves the targets
 * untouched and the next swr_synchronize returns SWR_ERR_CAPACITY, however many draws were
 * queued after it.  On a stream being captured, a draw that would have to grow an arena
 * returns SWR_ERR_CAPACITY at once and enqueues nothing. */
int  swr_draw_elements_async(swr_context *ctx, int mode, size_t count, const int32_t *indices);
int  swr_reserve(swr_context *ctx, size_t primitives, size_t tile_pairs);

/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * triangle draw of swr_draw_elements_async (same arguments, same results, same
 * device-buffers-only rule, failures reported by swr_synchronize) whose working memory is
 * carved from one fixed array per GPU of SWR_NOALLOC_SCRATCH_BYTES instead of the context's
 * growable arenas: it calls no hipMalloc, hipFree or memset.  Draws whose input-sized
 * arenas (about 64 bytes per triangle) and records (one per input triangle, plus an eighth)
 * do not fit fail at once with SWR_ERR_CAPACITY; tile pairs take what remains, and a draw
 * that emits more fails at the next swr_synchronize with SWR_ERR_CAPACITY (nothing drawn).
 * Every call on one GPU uses that GPU's one array: calls that overlap on one GPU (two
 * streams, two contexts, two threads) race and must be ordered by the caller. */
#define SWR_NOALLOC_SCRATCH_BYTES (64 << 20)
int  swr_draw_elements_noalloc(swr_context *ctx, int mode, size_t count, const int32_t *indices);

/* ---- batched draws --------------------------------------------------------------------
 * No counterpart in the reference: a batch stands for a sequence of independent
 * drawElements(Triangle) calls (VertexProcessor.cpp:78-120), each preceded by its own
 * setViewport / setDepthRange / setCullMode (VertexProcessor.cpp:43-65), setRasterMode /
 * setScissorRect (Rasterizer.h:75-87), setVertexAttribPointer(0, ...) (VertexProcessor.cpp:
 * 67-76) and render targets, all run as one pass of the pipeline (one set of launches for
 * the whole batch).  The vertex program, its counts, the pixel program and its counts are
 * the context's at draw time and shared by every scene.  Every scene's result is the one
 * its own draw would give (bit for bit).  Device buffers only; SWR_PS_LOG is not batched. */
typedef struct swr_scene {
    const int32_t *indices;        /* device; index_count / 3 triangles                   */
    uint64_t index_count;
    const void *attrib0;           /* device vertex data of attribute 0                  */
    int32_t  attrib0_stride;
    int32_t  cull_mode;            /* VertexProcessor::setCullMode                        */
    uint64_t attrib0_bytes;        /* 0: unchecked (as swr_set_vertex_attrib_pointer)     */
    float    mvp[16];              /* the MVP programs' uniform for this scene            */
    int32_t  viewport[4];          /* x, y, w, h (VertexProcessor::setViewport)           */
    float    depth_near, depth_far;/* VertexProcessor::setDepthRange                      */
    int32_t  raster_mode;          /* Rasterizer::setRasterMode                           */
    int32_t  scissor[4];           /* x, y, w, h (Rasterizer::setScissorRect)             */
    int32_t  width, height, pitch; /* render targets, as in swr_pixel_shader_desc         */
    void    *target0;
    void    *target1;
    void    *target2;
} swr_scene;

typedef struct swr_batch swr_batch;
/* Validates the scenes and uploads their table to the device (synchronous).  The scene
 * array is copied; the buffers it points to must stay valid while the batch is drawn. */
int  swr_batch_create(swr_context *ctx, const swr_scene *scenes, size_t count, swr_batch **out);
/* Draws every scene.  async = 0: returns when all targets are final (arenas grow as needed).
 * async = 1: only enqueues kernels and one status copy on the context stream, so the call
 * can be captured into a HIP graph; arenas must already be large enough (a synchronous draw
 * of the same batch sizes them), else swr_synchronize reports SWR_ERR_CAPACITY. */
int  swr_batch_draw(swr_context *ctx, const swr_batch *batch, int async);
int  swr_batch_destroy(swr_batch *batch);

/* ---- differentiable shading over a visibility buffer ---------------------------------
 * No counterpart in the reference.  For V views of one triangle mesh (clip-space positions
 * per view, attributes shared), after a SWR_PS_VISIBILITY draw of every view (viewport
 * (0, 0, width, height), depth cleared to 1, ids to 0):
 *   forward:  bary  = perspective-correct barycentrics (b0, b1) of the visible triangle at
 *                     each pixel centre; color = sum_i b_i * attribute_i (0 on background);
 *             image = color plus the analytic edge-antialiasing blend across silhouette
 *                     edges between 4-neighbour pixels.
 *   backward: given d(loss)/d(image), adds d(loss)/d(clip x, y, w) per view and vertex to
 *             grad_clip ([views][vertex_count][4]; z is left alone) and d(loss)/d(attribute)
 *             summed over views to grad_attributes ([vertex_count][channels]).
 * All buffers are device memory; both calls only enqueue kernels on the context stream
 * (they can be captured into a HIP graph).  The backward pass recomputes everything else
 * from ids, depth, bary and color.
 *
 * Several meshes in one call (scenes > 1): scene s owns views scene_views[s] ..
 * scene_views[s+1]-1, vertices scene_vertices[s] .. scene_vertices[s+1]-1 and triangles
 * scene_triangles[s] .. scene_triangles[s+1]-1 (host arrays of scenes + 1 prefix sums, from
 * 0 to views / vertex_count / triangle_count).  Triangles hold vertex numbers of their own
 * scene in the global numbering, adjacency global triangle numbers; ids hold the scene's own
 * triangle index + 1 (what a visibility draw of the scene's triangle list writes).  clip and
 * grad_clip then hold, scene by scene, each of its views' rows for its own vertices only:
 * sum over s of views(s) * vertices(s) rows of 4 floats.  The tables are uploaded when they
 * change (a synchronous copy: not while the stream is being captured). */
#define SWR_DIFF_CULLED 1          /* every visible triangle faces the same way (the visibility
                                      draw culled one winding): pixels of two triangles that
                                      share a vertex are never separated by a silhouette */
#define SWR_DIFF_L2_TARGET 2       /* swr_diff_backward: grad_image holds a target image T and
                                      the loss is 1/2 |image - T|^2, so dL/dimage = image - T
                                      (formed where it is read, from the forward's image) */
typedef struct swr_diff_desc {
    int32_t views, width, height;
    int32_t vertex_count, triangle_count;
    int32_t channels;              /* 1..4 */
    int32_t flags;                 /* SWR_DIFF_CULLED, SWR_DIFF_L2_TARGET                  */
    int32_t scenes;                /* 0 or 1: one mesh seen by every view; > 1: see above  */
    const float    *clip;          /* [views][vertex_count][4] clip-space x, y, z, w       */
    const int32_t  *triangles;     /* [triangle_count][3]                                  */
    const int32_t  *adjacency;     /* [triangle_count][3]: the triangle across the edge
                                      opposite vertex e, -1 if none (silhouette test)      */
    const float    *attributes;    /* [vertex_count][channels]                             */
    const uint32_t *ids;           /* [views][height][width]: triangle + 1, 0 = background */
    const float    *depth;         /* [views][height][width]                               */
    float *bary;                   /* [views][height][width][2]: forward out, backward in  */
    float *color;                  /* optional [views][height][width][channels]: forward out,
                                      the colour before antialiasing (NULL: not written; the
                                      kernels recompute the few they need)                 */
    float *image;                  /* [views][height][width][channels]: forward out        */
    const int32_t *scene_views;    /* scenes > 1: host [scenes + 1]                        */
    const int32_t *scene_vertices; /* scenes > 1: host [scenes + 1]                        */
    const int32_t *scene_triangles;/* scenes > 1: host [scenes + 1]                        */
} swr_diff_desc;
int  swr_diff_forward(swr_context *ctx, const swr_diff_desc *d);
int  swr_diff_backward(swr_context *ctx, const swr_diff_desc *d, const float *grad_image,
                       float *grad_clip, float *grad_attributes);

/* ---- tuning ---------------------------------------------------------------------------
 * No counterpart in the reference.  Launch shapes of the draw path that change speed, never
 * results (every setting draws the same bits); softwarerenderer_amd/autotune.py measures
 * them on a workload and picks the fastest.  0 selects the built-in default. */
typedef struct swr_tuning {
    int32_t raster_waves;    /* 8x8 tiles (one wave64 each) per raster workgroup: 1, 2, 4, 8 (4) */
    int32_t raster_group;    /* primitives per row-walk group in the raster kernel of the
                                depth-checker and visibility programs: 1, 2, 4, 8 (program's) */
    int32_t xcd_order;       /* 1: raster tile blocks grouped per XCD (default), -1: linear */
    int32_t expand_chunk;    /* tile pairs per expansion workgroup: 256 .. 4096 (1024) */
    int32_t vertex_pass;     /* 1: shade every vertex of attribute 0 once into post-transform
                                vertices, -1: shade each triangle corner where it is used,
                                0: automatic (the pass when the buffer's vertex count is known
                                and each vertex is referenced at least twice on average) */
    int32_t bin_size;        /* screen bins of 8, 16 or 32 pixels (8) */
    int32_t waves_per_bin;   /* bins of 16 or 32 pixels: waves sharing a bin's 8x8 blocks:
                                1, 2, 4 (4) */
    int32_t raster_queue;    /* 1: persistent raster workgroups (as many as the GPU holds)
                                whose waves take bins from per-XCD queues, stealing from the
                                other XCDs' when their own is empty; -1: one workgroup per bin
                                (default) */
} swr_tuning;
int  swr_set_tuning(swr_context *ctx, const swr_tuning *t);
int  swr_get_tuning(swr_context *ctx, swr_tuning *t);

/* ---- textures (Texture.h:13-33, 237-296) ---------------------------------------------
 * A texture with its mip chain: level 0 is width x height texels 0x00RRGGBB (the top byte
 * is ignored) read with a row pitch in texels from host or device memory; each further level
 * halves the size (at least 1) with the 2x2 box filter of generateMipmaps (a missing
 * neighbour repeats the top-left texel, channels averaged with >> 2), down to 1x1.
 * max_anisotropy is clamped to 1..16 (Texture.h:15-16).  A texture is used by the context
 * that created it; destroy it before that context. */
typedef struct swr_texture swr_texture;
int  swr_texture_create(swr_context *ctx, const uint32_t *texels, int width, int height,
                        int pitch, int max_anisotropy, swr_texture **out);
int  swr_texture_destroy(swr_texture *texture);

/* ---- statistics of the last completed draw ------------------------------------------- */
typedef struct swr_draw_stats {
    uint64_t input_primitives;   /* triangles / lines / points submitted          */
    uint64_t raster_primitives;  /* primitives after clip, cull and setup          */
    uint64_t tile_pairs;         /* (primitive, 8x8 tile) pairs binned             */
    uint64_t tiles;              /* tiles of the render target                     */
    float    stage_ms[8];        /* per-stage GPU time when profiling is enabled:   */
                                 /* 0 geometry, 1 scans, 2 emit, 3 sort, 4 ranges, 5 raster */
    uint32_t stage_launches[8];
    uint64_t pairs_emitted;      /* pairs written by binning, before dropping the (tile, block)
                                    pairs that cannot shade: what the pair arena must hold,
                                    the figure to give swr_reserve                   */
} swr_draw_stats;
int  swr_get_draw_stats(swr_context *ctx, swr_draw_stats *out);
/* Record HIP events around each stage (on the stream the stage runs on). */
int  swr_set_profiling(swr_context *ctx, int enable);

#ifdef __cplusplus
}
#endif
#endif /* SWR_HIP_H */
