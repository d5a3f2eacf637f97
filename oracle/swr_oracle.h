/*
 * swr_oracle.h — TEST INFRASTRUCTURE ONLY (parity checker, CPU baseline).
 *
 * A plain-C restatement of trenki2/SoftwareRenderer's draw path.  Only tests/, the
 * smoke check and bench.py's cpu_baseline leg may load it; the product library never
 * links it.  oracle/ref_driver.cpp exports the same entry points (prefix swr_ref_) on top
 * of the reference sources themselves, so the two can be compared call for call.
 *
 * All pointers are host pointers.  Pixel programs, vertex programs, keys and the record
 * layout of SWR_PS_LOG are those documented in include/swr_hip.h.
 */
#ifndef SWR_ORACLE_H
#define SWR_ORACLE_H

#include "../include/swr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The whole pipeline state of one draw: what Rasterizer + VertexProcessor hold. */
typedef struct swr_pipeline_state {
    int32_t raster_mode;           /* Rasterizer::setRasterMode        */
    int32_t scissor[4];            /* Rasterizer::setScissorRect x,y,w,h */
    int32_t viewport[4];           /* VertexProcessor::setViewport     */
    float   depth_near, depth_far; /* VertexProcessor::setDepthRange   */
    int32_t cull_mode;             /* VertexProcessor::setCullMode     */
    int32_t reserved;
    swr_vertex_shader_desc vs;
    swr_pixel_shader_desc  ps;
    const void *attrib[SWR_MAX_VERTEX_ATTRIBS];
    int32_t stride[SWR_MAX_VERTEX_ATTRIBS];
} swr_pipeline_state;

/* Oracle restatement.  Return SWR_OK or a negative SWR_ERR_* code. */
int swr_oracle_draw_elements(const swr_pipeline_state *st, int mode, size_t count,
                             const int32_t *indices);
int swr_oracle_draw_triangle_list(const swr_pipeline_state *st, const swr_rasterizer_vertex *v,
                                  size_t vertex_count, const int32_t *indices, size_t index_count);
int swr_oracle_draw_line_list(const swr_pipeline_state *st, const swr_rasterizer_vertex *v,
                              size_t vertex_count, const int32_t *indices, size_t index_count);
int swr_oracle_draw_point_list(const swr_pipeline_state *st, const swr_rasterizer_vertex *v,
                               size_t vertex_count, const int32_t *indices, size_t index_count);

/* Triangle setup of TriangleEquations (TriangleEquations.h:47-71) for three window-space
 * vertices, written as floats: area2, e0.abc, e1.abc, e2.abc, z.abc, invw.abc,
 * avar[na].abc, pvar[np].abc.  Returns the number of floats written (0 when area2 <= 0
 * stops the setup early, in which case only area2 and the edges are valid). */
int swr_oracle_triangle_setup(const swr_rasterizer_vertex *v0, const swr_rasterizer_vertex *v1,
                              const swr_rasterizer_vertex *v2, int na, int np, float *out);

/* Sutherland-Hodgman clip of one clip-space triangle against the planes selected by the
 * OR of its clip masks (VertexProcessor.cpp:217-263, PolyClipper.cpp:45-81).  Writes up to
 * max_out vertices of the resulting polygon; returns its vertex count (< 3: fully clipped). */
int swr_oracle_clip_triangle(const swr_rasterizer_vertex *v0, const swr_rasterizer_vertex *v1,
                             const swr_rasterizer_vertex *v2, int na, int np,
                             swr_rasterizer_vertex *out, int max_out);

/* A texture for SWR_PS_TEXTURE (its pointer goes in ps.texture): level 0 copied from
 * `texels` (0x00RRGGBB, row pitch in texels) and the mip chain of Texture::generateMipmaps
 * (Texture.h:237-296); max_anisotropy clamped to 1..16. */
typedef struct swr_oracle_texture swr_oracle_texture;
swr_oracle_texture *swr_oracle_texture_create(const uint32_t *texels, int width, int height,
                                              int pitch, int max_anisotropy);
void swr_oracle_texture_destroy(swr_oracle_texture *t);
/* Texture::sample of the texture (Texture.h:35-118), the level-of-detail log2, and mip level
 * `level` (its texels stay owned by the texture). */
uint32_t swr_oracle_texture_sample(const swr_oracle_texture *t, float u, float v, float dudx,
                                   float dvdx, float dudy, float dvdy);
float swr_oracle_log2(float x);
/* 1: the sampler's level of detail uses the C library's log2f (the reference's call) instead
 * of log2 rounded once (the device's); process-wide, for measuring the difference. */
void swr_oracle_set_log2_libm(int on);
int swr_oracle_texture_level(const swr_oracle_texture *t, int level, const uint32_t **texels,
                             int *width, int *height);

#ifdef __cplusplus
}
#endif
#endif
