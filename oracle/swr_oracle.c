/*
 * swr_oracle.c — TEST INFRASTRUCTURE ONLY: the parity checker for the gfx950 draw path.
 *
 * A plain-C restatement of trenki2/SoftwareRenderer's draw path
 * (VertexProcessor::drawElements -> ... -> PixelShader::drawPixel).  Each function cites
 * the reference file:line it follows (paths relative to src/renderer/).  It is compiled
 * with -O2 -ffp-contract=off on x86-64 (SSE scalar float, IEEE round-to-nearest, denormals
 * kept), so every float operation rounds once, in the order written.
 *
 * Where the reference is undefined the oracle fixes a behaviour, and the GPU path does
 * the same:
 *   - float -> int conversions follow x86-64 cvttss2si (out of range / NaN -> INT_MIN);
 *   - pixels outside the render target are never written (the reference writes out of
 *     bounds there);
 *   - triangles with a non-finite window-space x or y, or a non-finite area, are skipped
 *     (the reference loops over billions of scanlines / blocks there);
 *   - vertices created by the clipper carry 0 in the varyings the vertex shader does not
 *     declare (the reference leaves them uninitialised, PolyClipper.h:36);
 *   - Block mode with PVarCount > 0 and InterpolateW == false chains 1/w as in Span mode
 *     (the reference reads an uninitialised invw there, PixelShaderBase.h:122-136);
 *   - negative indices are rejected (the reference reads before the attribute buffer).
 *
 * What pins each part (tests/): the triangle, line and point paths against the reference
 * compiled from its own sources (oracle/_ref) and the golden fixtures it wrote; OBJ loading
 * against the reference's ObjData.cpp; the texture sampler (Texture.h needs SDL, so the
 * reference's class cannot be built) by closed forms, and its level-of-detail log2 against
 * float64 and against the C library's log2f on textured scenes.  Parity unpinned: the
 * varyings of Block / Adaptive draws with PVarCount > 0 and InterpolateW == false (the
 * reference reads uninitialised memory there; coverage is pinned), and textured pixels
 * whose trilinear blend lands on a rounding boundary where log2f and the log2 rounded once
 * disagree (none on the test scenes).
 */
#include "swr_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef swr_rasterizer_vertex vtx;

/* ---------------------------------------------------------------- scalar helpers */

/* (int)f as executed by x86-64 cvttss2si. */
static int f2i(float f)
{
    if (f > -2147483904.0f && f < 2147483648.0f)
        return (int)f;
    return INT_MIN;
}

/* std::min / std::max argument order (they matter for NaN). */
static float min_f(float a, float b) { return (b < a) ? b : a; }
static float max_f(float a, float b) { return (a < b) ? b : a; }
static int min_i(int a, int b) { return (b < a) ? b : a; }
static int max_i(int a, int b) { return (a < b) ? b : a; }

static uint32_t fbits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* ---------------------------------------------------------------- setup math */

/* EdgeEquation::init / evaluate / test, EdgeEquation.h:37-61 */
typedef struct { float a, b, c; int tie; } edge_t;

static void edge_init(edge_t *e, const vtx *v0, const vtx *v1)
{
    e->a = v0->y - v1->y;
    e->b = v1->x - v0->x;
    e->c = -(e->a * (v0->x + v1->x) + e->b * (v0->y + v1->y)) / 2.0f;
    e->tie = e->a != 0.0f ? e->a > 0.0f : e->b > 0.0f;
}

static float edge_eval(const edge_t *e, float x, float y) { return e->a * x + e->b * y + e->c; }
static int edge_test(const edge_t *e, float v) { return v > 0.0f || (v == 0.0f && e->tie); }

/* ParameterEquation::init, ParameterEquation.h:36-48 */
typedef struct { float a, b, c; } param_t;

static void param_init(param_t *p, float p0, float p1, float p2, const edge_t *e0,
                       const edge_t *e1, const edge_t *e2, float f)
{
    p->a = f * (p0 * e0->a + p1 * e1->a + p2 * e2->a);
    p->b = f * (p0 * e0->b + p1 * e1->b + p2 * e2->b);
    p->c = f * (p0 * e0->c + p1 * e1->c + p2 * e2->c);
}

static float param_eval(const param_t *p, float x, float y) { return p->a * x + p->b * y + p->c; }

/* TriangleEquations constructor, TriangleEquations.h:47-71 */
typedef struct {
    float area2;
    edge_t e0, e1, e2;
    param_t z, invw, avar[SWR_MAX_AVARS], pvar[SWR_MAX_PVARS];
} tri_t;

static void tri_setup(tri_t *q, const vtx *v0, const vtx *v1, const vtx *v2, int na, int np)
{
    edge_init(&q->e0, v1, v2);
    edge_init(&q->e1, v2, v0);
    edge_init(&q->e2, v0, v1);
    q->area2 = q->e0.c + q->e1.c + q->e2.c;
    if (q->area2 <= 0.0f)
        return;
    float f = 1.0f / q->area2;
    param_init(&q->z, v0->z, v1->z, v2->z, &q->e0, &q->e1, &q->e2, f);
    float iw0 = 1.0f / v0->w, iw1 = 1.0f / v1->w, iw2 = 1.0f / v2->w;
    param_init(&q->invw, iw0, iw1, iw2, &q->e0, &q->e1, &q->e2, f);
    for (int i = 0; i < na; ++i)
        param_init(&q->avar[i], v0->avar[i], v1->avar[i], v2->avar[i], &q->e0, &q->e1, &q->e2, f);
    for (int i = 0; i < np; ++i)
        param_init(&q->pvar[i], v0->pvar[i] * iw0, v1->pvar[i] * iw1, v2->pvar[i] * iw2,
                   &q->e0, &q->e1, &q->e2, f);
}

/* ---------------------------------------------------------------- draw context */

typedef struct {
    const swr_pipeline_state *st;
    int minX, maxX, minY, maxY;      /* Rasterizer.h:81-87 */
    int na, np, iz, iw;              /* pixel shader statics */
    uint32_t key;                    /* primitive in draw order */
    uint32_t src;                    /* input triangle the primitive comes from */
} rctx;

static void rctx_init(rctx *r, const swr_pipeline_state *st)
{
    r->st = st;
    r->minX = st->scissor[0];
    r->minY = st->scissor[1];
    r->maxX = st->scissor[0] + st->scissor[2];
    r->maxY = st->scissor[1] + st->scissor[3];
    r->na = st->ps.avar_count;
    r->np = st->ps.pvar_count;
    r->iz = st->ps.interpolate_z != 0;
    r->iw = st->ps.interpolate_w != 0;
    r->key = 0;
    r->src = 0;
}

/* PixelData, PixelData.h:38-125 */
typedef struct {
    int x, y;
    float z, w, invw;
    float avar[SWR_MAX_AVARS], pvar[SWR_MAX_PVARS], pvt[SWR_MAX_PVARS];
    const tri_t *eq;  /* the triangle's equations (PixelData.h:55-56); NULL for points / lines */
} pix_t;

static void pix_init(pix_t *p, const rctx *r, const tri_t *q, float x, float y)
{
    p->eq = q;
    if (r->iz)
        p->z = param_eval(&q->z, x, y);
    if (r->iw || r->np > 0) {
        p->invw = param_eval(&q->invw, x, y);
        p->w = 1.0f / p->invw;
    }
    for (int i = 0; i < r->na; ++i)
        p->avar[i] = param_eval(&q->avar[i], x, y);
    for (int i = 0; i < r->np; ++i) {
        p->pvt[i] = param_eval(&q->pvar[i], x, y);
        p->pvar[i] = p->pvt[i] * p->w;
    }
}

/* stepX (dir 0) / stepY (dir 1), PixelData.h:84-125 */
static void pix_step(pix_t *p, const rctx *r, const tri_t *q, int dir)
{
#define STEP(P) ((dir) ? (P).b : (P).a)
    if (r->iz)
        p->z = p->z + STEP(q->z);
    if (r->iw || r->np > 0) {
        p->invw = p->invw + STEP(q->invw);
        p->w = 1.0f / p->invw;
    }
    for (int i = 0; i < r->na; ++i)
        p->avar[i] = p->avar[i] + STEP(q->avar[i]);
    for (int i = 0; i < r->np; ++i) {
        p->pvt[i] = p->pvt[i] + STEP(q->pvar[i]);
        p->pvar[i] = p->pvt[i] * p->w;
    }
#undef STEP
}

/* ---------------------------------------------------------------- pixel programs */

static float clamp01(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }
static uint32_t mix32(uint32_t h, uint32_t x) { return (h ^ x) * 16777619u; }

static uint32_t checker_rgba8(const pix_t *p, float cells)
{
    int cu = f2i(floorf(p->pvar[0] * cells));
    int cv = f2i(floorf(p->pvar[1] * cells));
    float k = (((uint32_t)cu + (uint32_t)cv) & 1u) ? 1.0f : 0.5f;
    uint32_t r = (uint32_t)(clamp01(p->avar[0] * k) * 255.0f);
    uint32_t g = (uint32_t)(clamp01(p->avar[1] * k) * 255.0f);
    uint32_t b = (uint32_t)(clamp01(p->avar[2] * k) * 255.0f);
    return r | (g << 8) | (b << 16) | 0xFF000000u;
}

/* ---------------------------------------------------------------- textures */

/* Texture (Texture.h:11-33) with its mip chain (generateMipmaps, Texture.h:237-296). */
struct swr_oracle_texture {
    int levels, max_aniso;
    int w[32], h[32];
    uint32_t *px[32];
};

static uint32_t tex_r(uint32_t c) { return (c >> 16) & 0xFFu; }
static uint32_t tex_g(uint32_t c) { return (c >> 8) & 0xFFu; }
static uint32_t tex_b(uint32_t c) { return c & 0xFFu; }
static uint32_t tex_pack(uint32_t r, uint32_t g, uint32_t b)
{
    return ((r & 0xFFu) << 16) | ((g & 0xFFu) << 8) | (b & 0xFFu);
}

swr_oracle_texture *swr_oracle_texture_create(const uint32_t *texels, int width, int height,
                                              int pitch, int max_anisotropy)
{
    if (!texels || width < 1 || height < 1 || pitch < width)
        return NULL;
    swr_oracle_texture *t = (swr_oracle_texture *)calloc(1, sizeof *t);
    if (!t)
        return NULL;
    /* std::clamp(maxAnisotropy, 1, MAX_ANISOTROPY), Texture.h:15-16 */
    t->max_aniso = max_anisotropy < 1 ? 1 : (max_anisotropy > 16 ? 16 : max_anisotropy);
    int w = width, h = height, lv = 0;
    t->px[0] = (uint32_t *)malloc((size_t)w * h * 4);
    if (!t->px[0]) {
        free(t);
        return NULL;
    }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            t->px[0][(size_t)y * w + x] = texels[(size_t)y * pitch + x] & 0xFFFFFFu;
    t->w[0] = w;
    t->h[0] = h;
    while (w > 1 || h > 1) {
        int nw = w / 2 > 1 ? w / 2 : 1, nh = h / 2 > 1 ? h / 2 : 1;
        const uint32_t *src = t->px[lv];
        uint32_t *dst = (uint32_t *)malloc((size_t)nw * nh * 4);
        if (!dst)
            break;
        for (int y = 0; y < nh; ++y)
            for (int x = 0; x < nw; ++x) {
                uint32_t p00 = src[(size_t)(y * 2) * w + (x * 2)];
                uint32_t p10 = x * 2 + 1 < w ? src[(size_t)(y * 2) * w + (x * 2 + 1)] : p00;
                uint32_t p01 = y * 2 + 1 < h ? src[(size_t)(y * 2 + 1) * w + (x * 2)] : p00;
                uint32_t p11 = (x * 2 + 1 < w && y * 2 + 1 < h)
                                   ? src[(size_t)(y * 2 + 1) * w + (x * 2 + 1)]
                                   : p00;
                dst[(size_t)y * nw + x] =
                    tex_pack((tex_r(p00) + tex_r(p10) + tex_r(p01) + tex_r(p11)) >> 2,
                             (tex_g(p00) + tex_g(p10) + tex_g(p01) + tex_g(p11)) >> 2,
                             (tex_b(p00) + tex_b(p10) + tex_b(p01) + tex_b(p11)) >> 2);
            }
        ++lv;
        t->px[lv] = dst;
        t->w[lv] = nw;
        t->h[lv] = nh;
        w = nw;
        h = nh;
    }
    t->levels = lv + 1;
    return t;
}

void swr_oracle_texture_destroy(swr_oracle_texture *t)
{
    if (!t)
        return;
    for (int i = 0; i < t->levels; ++i)
        free(t->px[i]);
    free(t);
}

/* log2 with one rounding to float: the C library's log2f (which Texture.h:124 calls through
 * std::log2) may round the other way by one ulp, so the level of detail is evaluated here as
 * a double series, then rounded; the device evaluates the same operations.  x = m 2^e with
 * m in [sqrt(1/2), sqrt(2)); log(m) = 2 atanh(s) = 2 (s + s^3/3 + ...), s = (m-1)/(m+1). */
static float log2_rounded(float xf)
{
    if (!(xf > 0.0f))
        return xf == 0.0f ? -INFINITY : NAN;
    if (isinf(xf))
        return xf;
    int e;
    double m = frexp((double)xf, &e);
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double q = 1.0 / 25.0;
    q = q * s2 + 1.0 / 23.0;
    q = q * s2 + 1.0 / 21.0;
    q = q * s2 + 1.0 / 19.0;
    q = q * s2 + 1.0 / 17.0;
    q = q * s2 + 1.0 / 15.0;
    q = q * s2 + 1.0 / 13.0;
    q = q * s2 + 1.0 / 11.0;
    q = q * s2 + 1.0 / 9.0;
    q = q * s2 + 1.0 / 7.0;
    q = q * s2 + 1.0 / 5.0;
    q = q * s2 + 1.0 / 3.0;
    q = q * s2 + 1.0;
    const double ln_m = 2.0 * s * q;
    return (float)((double)e + ln_m * 1.4426950408889634074);
}

/* std::clamp of the reference (argument order matters for NaN) */
static float clamp_f(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

/* Texture::sampleBilinear / bilinearLerpColors, Texture.h:148-228 */
static uint32_t tex_bilinear(const swr_oracle_texture *t, int lv, float u, float v)
{
    const int w = t->w[lv], h = t->h[lv];
    float px = u * (float)(w - 1);
    float py = v * (float)(h - 1);
    int x0 = max_i(0, f2i(floorf(px)));
    int y0 = max_i(0, f2i(floorf(py)));
    int x1 = min_i(w - 1, x0 + 1);
    int y1 = min_i(h - 1, y0 + 1);
    float fx = px - (float)x0;
    float fy = py - (float)y0;
    const uint32_t *p = t->px[lv];
    uint32_t c00 = p[(size_t)y0 * w + x0], c10 = p[(size_t)y0 * w + x1];
    uint32_t c01 = p[(size_t)y1 * w + x0], c11 = p[(size_t)y1 * w + x1];
    uint32_t (*ch[3])(uint32_t) = {tex_r, tex_g, tex_b};
    uint32_t out[3];
    for (int k = 0; k < 3; ++k) {
        float s = (float)ch[k](c00) * (1.0f - fx) * (1.0f - fy) + (float)ch[k](c10) * fx * (1.0f - fy) +
                  (float)ch[k](c01) * (1.0f - fx) * fy + (float)ch[k](c11) * fx * fy;
        out[k] = (uint32_t)(uint8_t)f2i(s);
    }
    return tex_pack(out[0], out[1], out[2]);
}

/* Texture::lerpColors, Texture.h:200-205: the channel difference times t converted to
 * Uint8 (two's complement wrap for a negative difference, as x86-64 converts), added */
static uint32_t tex_lerp(uint32_t c1, uint32_t c2, float t)
{
    uint32_t (*ch[3])(uint32_t) = {tex_r, tex_g, tex_b};
    uint32_t out[3];
    for (int k = 0; k < 3; ++k) {
        int d = (int)ch[k](c2) - (int)ch[k](c1);
        uint8_t step = (uint8_t)f2i((float)d * t);
        out[k] = (uint8_t)(ch[k](c1) + step);
    }
    return tex_pack(out[0], out[1], out[2]);
}

/* Which log2 the level of detail uses: 0 (default) log2_rounded, as the device does; 1 the
 * C library's log2f, as the reference's std::log2 does (swr_oracle_set_log2_libm).  Tests
 * use mode 1 only to measure how far the device's choice moves the images. */
static int g_log2_libm = 0;
void swr_oracle_set_log2_libm(int on) { g_log2_libm = on != 0; }

/* Texture::sampleTrilinear, Texture.h:121-146 */
static uint32_t tex_trilinear(const swr_oracle_texture *t, float u, float v, float rho)
{
    const float r = rho < 1e-6f ? 1e-6f : rho;
    float lod = g_log2_libm ? log2f(r) : log2_rounded(r);
    lod = clamp_f(lod, 0.0f, (float)(t->levels - 1));
    int base = f2i(floorf(lod));
    int next = min_i(base + 1, t->levels - 1);
    float frac = clamp_f(lod - (float)base, 0.0f, 1.0f);
    uint32_t cb = tex_bilinear(t, base, u, v);
    if (base != next)
        return tex_lerp(cb, tex_bilinear(t, next, u, v), frac);
    return cb;
}

static float wrap01(float u)
{
    u = fmodf(u, 1.0f);
    if (u < 0)
        u += 1.0f;
    return u;
}

/* Texture::sample, Texture.h:35-118 */
static uint32_t tex_sample(const swr_oracle_texture *t, float u, float v, float dudx, float dvdx,
                           float dudy, float dvdy)
{
    u = wrap01(u);
    v = wrap01(v);
    const float W = (float)t->w[0], H = (float)t->h[0];
    float dudx_s = dudx * W, dvdx_s = dvdx * H, dudy_s = dudy * W, dvdy_s = dvdy * H;
    float dx_len = sqrtf(dudx_s * dudx_s + dvdx_s * dvdx_s);
    float dy_len = sqrtf(dudy_s * dudy_s + dvdy_s * dvdy_s);
    dx_len = max_f(dx_len, 1e-6f);
    dy_len = max_f(dy_len, 1e-6f);
    float major = max_f(dx_len, dy_len), minor = min_f(dx_len, dy_len);
    float ratio = min_f(major / minor, (float)t->max_aniso);
    int n = max_i(1, f2i(ceilf(ratio)));
    if (n <= 1)
        return tex_trilinear(t, u, v, major);
    float mdu, mdv;
    if (dx_len > dy_len) {
        mdu = dudx / dx_len;
        mdv = dvdx / dx_len;
    } else {
        mdu = dudy / dy_len;
        mdv = dvdy / dy_len;
    }
    float r = 0, g = 0, b = 0;
    float step = 1.0f / (float)n;
    for (int i = 0; i < n; ++i) {
        float tt = ((float)i + 0.5f) * step - 0.5f;
        float su = wrap01(u + mdu * major * tt);
        float sv = wrap01(v + mdv * major * tt);
        uint32_t c = tex_trilinear(t, su, sv, minor);
        r += (float)tex_r(c);
        g += (float)tex_g(c);
        b += (float)tex_b(c);
    }
    r = min_f(255.0f, r / (float)n);
    g = min_f(255.0f, g / (float)n);
    b = min_f(255.0f, b / (float)n);
    return tex_pack((uint8_t)f2i(r), (uint8_t)f2i(g), (uint8_t)f2i(b));
}

/* exported probes of the sampler (tests pin them by closed forms) */
uint32_t swr_oracle_texture_sample(const swr_oracle_texture *t, float u, float v, float dudx,
                                   float dvdx, float dudy, float dvdy)
{
    return tex_sample(t, u, v, dudx, dvdx, dudy, dvdy);
}

float swr_oracle_log2(float x) { return log2_rounded(x); }

int swr_oracle_texture_level(const swr_oracle_texture *t, int level, const uint32_t **texels,
                             int *width, int *height)
{
    if (!t || level < 0 || level >= t->levels)
        return SWR_ERR_INVALID_ARG;
    *texels = t->px[level];
    *width = t->w[level];
    *height = t->h[level];
    return SWR_OK;
}

/* PixelData::computePerspectiveDerivatives, PixelData.h:127-145 */
static void persp_derivs(const pix_t *p, int k, float *ddx, float *ddy)
{
    const tri_t *q = p->eq;
    float xf = (float)p->x + 0.5f, yf = (float)p->y + 0.5f;
    float var = param_eval(&q->pvar[k], xf, yf);
    float iw = param_eval(&q->invw, xf, yf);
    *ddx = (iw * q->pvar[k].a - var * q->invw.a) / (iw * iw);
    *ddy = (iw * q->pvar[k].b - var * q->invw.b) / (iw * iw);
}

/* The user drawPixel of PixelShaderBase.h:114-119, for the built-in programs. */
static void draw_pixel(const rctx *r, const pix_t *p)
{
    const swr_pixel_shader_desc *ps = &r->st->ps;
    if (p->x < 0 || p->y < 0 || p->x >= ps->width || p->y >= ps->height)
        return;
    size_t o = (size_t)p->y * (size_t)ps->pitch + (size_t)p->x;
    switch (ps->program) {
    case SWR_PS_DEPTH_CHECKER: {
        float *depth = (float *)ps->target1;
        uint32_t *color = (uint32_t *)ps->target0;
        if (p->z < depth[o]) {
            depth[o] = p->z;
            color[o] = checker_rgba8(p, ps->checker_cells);
        }
        break;
    }
    case SWR_PS_RGB: {
        uint32_t ri = (uint32_t)f2i(p->avar[0] * 255.0f);
        uint32_t gi = (uint32_t)f2i(p->avar[1] * 255.0f);
        uint32_t bi = (uint32_t)f2i(p->avar[2] * 255.0f);
        ((uint32_t *)ps->target0)[o] = (ri << 16) | (gi << 8) | bi;
        break;
    }
    case SWR_PS_COVERAGE:
        ((uint32_t *)ps->target0)[o] = 1u;
        break;
    case SWR_PS_TEXTURE: {
        /* Box.cpp:39-53 */
        float dudx, dudy, dvdx, dvdy;
        persp_derivs(p, 0, &dudx, &dudy);
        persp_derivs(p, 1, &dvdx, &dvdy);
        ((uint32_t *)ps->target0)[o] = tex_sample((const swr_oracle_texture *)ps->texture,
                                                  p->pvar[0], p->pvar[1], dudx, dvdx, dudy, dvdy);
        break;
    }
    case SWR_PS_VISIBILITY: {
        float *depth = (float *)ps->target1;
        if (p->z < depth[o]) {
            depth[o] = p->z;
            ((uint32_t *)ps->target0)[o] = r->src + 1u;
        }
        break;
    }
    case SWR_PS_HASH: {
        uint32_t *hp = (uint32_t *)ps->target0;
        uint32_t h = mix32(hp[o], r->key);
        if (r->iz)
            h = mix32(h, fbits(p->z));
        if (r->iw) {
            h = mix32(h, fbits(p->w));
            h = mix32(h, fbits(p->invw));
        }
        for (int i = 0; i < r->na; ++i)
            h = mix32(h, fbits(p->avar[i]));
        for (int i = 0; i < r->np; ++i)
            h = mix32(h, fbits(p->pvar[i]));
        hp[o] = h;
        ((uint32_t *)ps->target1)[o] += 1u;
        break;
    }
    case SWR_PS_LOG: {
        uint32_t *cnt = (uint32_t *)ps->target0;
        uint32_t seq = cnt[o]++;
        uint64_t n = (*ps->log_count)++;
        if (n < ps->log_capacity) {
            uint32_t *rec = ps->log + n * SWR_LOG_WORDS;
            memset(rec, 0, SWR_LOG_WORDS * 4);
            rec[0] = (uint32_t)p->x;
            rec[1] = (uint32_t)p->y;
            rec[2] = r->key;
            rec[3] = seq;
            if (r->iz)
                rec[4] = fbits(p->z);
            if (r->iw) {
                rec[5] = fbits(p->w);
                rec[6] = fbits(p->invw);
            }
            for (int i = 0; i < r->na; ++i)
                rec[7 + i] = fbits(p->avar[i]);
            for (int i = 0; i < r->np; ++i)
                rec[7 + SWR_MAX_AVARS + i] = fbits(p->pvar[i]);
        }
        break;
    }
    default:
        break;
    }
}

/* ---------------------------------------------------------------- Block traversal */

/* PixelShaderBase::drawBlock<TestEdges>, PixelShaderBase.h:55-94 (copyPixelData :122-136) */
static void draw_block(const rctx *r, const tri_t *q, int x, int y, int test_edges)
{
    const swr_pixel_shader_desc *ps = &r->st->ps;
    float xf = (float)x + 0.5f, yf = (float)y + 0.5f;
    pix_t po, pi;
    pix_init(&po, r, q, xf, yf);
    float eo0 = 0, eo1 = 0, eo2 = 0;
    if (test_edges) {
        eo0 = edge_eval(&q->e0, xf, yf);
        eo1 = edge_eval(&q->e1, xf, yf);
        eo2 = edge_eval(&q->e2, xf, yf);
    }
    for (int yy = y; yy < y + SWR_BLOCK_SIZE; yy++) {
        pi = po;
        float ei0 = eo0, ei1 = eo1, ei2 = eo2;
        for (int xx = x; xx < x + SWR_BLOCK_SIZE; xx++) {
            if (!test_edges || (edge_test(&q->e0, ei0) && edge_test(&q->e1, ei1) &&
                                edge_test(&q->e2, ei2))) {
                pi.x = xx;
                pi.y = yy;
                draw_pixel(r, &pi);
            }
            pix_step(&pi, r, q, 0);
            if (test_edges) {
                ei0 = ei0 + q->e0.a;
                ei1 = ei1 + q->e1.a;
                ei2 = ei2 + q->e2.a;
            }
        }
        pix_step(&po, r, q, 1);
        if (test_edges) {
            eo0 = eo0 + q->e0.b;
            eo1 = eo1 + q->e1.b;
            eo2 = eo2 + q->e2.b;
        }
        if (yy >= ps->height)
            break; /* later rows are outside the target */
    }
}

/* Corner classification of one block, Rasterizer.h:267-304.
 * Returns 0 (skip), 1 (drawBlock<true>) or 2 (drawBlock<false>). */
static int classify_block(const tri_t *q, int x, int y)
{
    const float s = (float)(SWR_BLOCK_SIZE - 1);
    float xf = (float)x + 0.5f, yf = (float)y + 0.5f;
    const edge_t *E[3] = {&q->e0, &q->e1, &q->e2};
    int t[4][3];
    for (int k = 0; k < 3; ++k) {
        float v00 = edge_eval(E[k], xf, yf);
        float v01 = v00 + E[k]->b * s;
        float v10 = v00 + E[k]->a * s;
        float v11 = v01 + E[k]->a * s;
        t[0][k] = edge_test(E[k], v00);
        t[1][k] = edge_test(E[k], v01);
        t[2][k] = edge_test(E[k], v10);
        t[3][k] = edge_test(E[k], v11);
    }
    int result = 0, all_same = 1;
    for (int c = 0; c < 4; ++c) {
        result += t[c][0] && t[c][1] && t[c][2];
        /* `a == b == c` parses as (a == b) == c: true iff an odd number of edges pass */
        all_same = all_same && ((t[c][0] == t[c][1]) == t[c][2]);
    }
    if (result == 0)
        return all_same ? 0 : 1;
    return result == 4 ? 2 : 1;
}

/* Rasterizer::drawTriangleBlockTemplate, Rasterizer.h:224-306 */
static void raster_block(const rctx *r, const vtx *v0, const vtx *v1, const vtx *v2)
{
    tri_t q;
    tri_setup(&q, v0, v1, v2, r->na, r->np);
    if (q.area2 <= 0.0f)
        return;
    long long minX = f2i(min_f(min_f(v0->x, v1->x), v2->x));
    long long maxX = f2i(max_f(max_f(v0->x, v1->x), v2->x));
    long long minY = f2i(min_f(min_f(v0->y, v1->y), v2->y));
    long long maxY = f2i(max_f(max_f(v0->y, v1->y), v2->y));
    minX = minX > r->minX ? minX : r->minX;
    maxX = maxX < r->maxX ? maxX : r->maxX;
    minY = minY > r->minY ? minY : r->minY;
    maxY = maxY < r->maxY ? maxY : r->maxY;
    minX &= ~(long long)(SWR_BLOCK_SIZE - 1);
    maxX &= ~(long long)(SWR_BLOCK_SIZE - 1);
    minY &= ~(long long)(SWR_BLOCK_SIZE - 1);
    maxY &= ~(long long)(SWR_BLOCK_SIZE - 1);
    /* C division truncates toward zero, as in the reference's int arithmetic */
    long long stepsX = (maxX - minX) / SWR_BLOCK_SIZE + 1;
    long long stepsY = (maxY - minY) / SWR_BLOCK_SIZE + 1;
    /* Block i covers (minX + (i % stepsX)*8, minY + (i / stepsX)*8) for i < stepsX*stepsY.
     * Both counts negative gives a positive product: the blocks then run right from minX
     * and up from minY (Rasterizer.h:254-265). */
    long long bx0, bx1, by0, by1;
    if (stepsX > 0 && stepsY > 0) {
        bx0 = minX; bx1 = minX + 8 * (stepsX - 1);
        by0 = minY; by1 = minY + 8 * (stepsY - 1);
    } else if (stepsX < 0 && stepsY < 0) {
        bx0 = minX; bx1 = minX + 8 * (-stepsX - 1);
        by0 = minY - 8 * (-stepsY - 1); by1 = minY;
    } else {
        return;
    }
    /* blocks entirely outside the render target draw nothing observable */
    const swr_pixel_shader_desc *ps = &r->st->ps;
    if (bx0 < -8) bx0 += ((-8 - bx0) / 8) * 8;
    if (by0 < -8) by0 += ((-8 - by0) / 8) * 8;
    if (bx1 > ps->width) bx1 -= ((bx1 - ps->width) / 8) * 8;
    if (by1 > ps->height) by1 -= ((by1 - ps->height) / 8) * 8;
    for (long long by = by0; by <= by1; by += 8) {
        for (long long bx = bx0; bx <= bx1; bx += 8) {
            int cls = classify_block(&q, (int)bx, (int)by);
            if (cls)
                draw_block(r, &q, (int)bx, (int)by, cls == 1);
        }
    }
}

/* ---------------------------------------------------------------- Span traversal */

/* PixelShaderBase::drawSpan, PixelShaderBase.h:96-112 */
static void draw_span(const rctx *r, const tri_t *q, int x, int y, int x2)
{
    const swr_pixel_shader_desc *ps = &r->st->ps;
    if (y < 0 || y >= ps->height)
        return;
    pix_t p;
    p.y = y;
    pix_init(&p, r, q, (float)x + 0.5f, (float)y + 0.5f);
    if (x2 > ps->width)
        x2 = ps->width; /* later pixels are outside the target */
    while (x < x2) {
        p.x = x;
        draw_pixel(r, &p);
        pix_step(&p, r, q, 0);
        x++;
    }
}

/* Rasterizer::drawBottomFlatTriangle, Rasterizer.h:360-385 */
static void bottom_flat(const rctx *r, const tri_t *q, const vtx *v0, const vtx *v1, const vtx *v2)
{
    float is1 = (v1->x - v0->x) / (v1->y - v0->y);
    float is2 = (v2->x - v0->x) / (v2->y - v0->y);
    int lo = f2i(v0->y + 0.5f), hi = f2i(v1->y + 0.5f);
    lo = max_i(lo, -1);
    hi = min_i(hi, r->st->ps.height);
    for (int sy = lo; sy < hi; sy++) {
        float dy = ((float)sy - v0->y) + 0.5f;
        float cx1 = v0->x + is1 * dy + 0.5f;
        float cx2 = v0->x + is2 * dy + 0.5f;
        int xl = max_i(r->minX, f2i(cx1));
        int xr = min_i(r->maxX, f2i(cx2));
        draw_span(r, q, xl, sy, xr);
    }
}

/* Rasterizer::drawTopFlatTriangle, Rasterizer.h:387-411 (scanlines run bottom-up) */
static void top_flat(const rctx *r, const tri_t *q, const vtx *v0, const vtx *v1, const vtx *v2)
{
    float is1 = (v2->x - v0->x) / (v2->y - v0->y);
    float is2 = (v2->x - v1->x) / (v2->y - v1->y);
    int hi = f2i(v2->y - 0.5f), lo = f2i(v0->y - 0.5f);
    hi = min_i(hi, r->st->ps.height);
    lo = max_i(lo, -2);
    for (int sy = hi; sy > lo; sy--) {
        float dy = ((float)sy - v2->y) + 0.5f;
        float cx1 = v2->x + is1 * dy + 0.5f;
        float cx2 = v2->x + is2 * dy + 0.5f;
        int xl = max_i(r->minX, f2i(cx1));
        int xr = min_i(r->maxX, f2i(cx2));
        draw_span(r, q, xl, sy, xr);
    }
}

/* Rasterizer::drawTriangleSpanTemplate, Rasterizer.h:308-358 */
static void raster_span(const rctx *r, const vtx *v0, const vtx *v1, const vtx *v2)
{
    tri_t q;
    tri_setup(&q, v0, v1, v2, r->na, r->np);
    if (q.area2 <= 0.0f)
        return;
    const vtx *t = v0, *m = v1, *b = v2, *tmp;
    if (t->y > m->y) { tmp = t; t = m; m = tmp; }
    if (m->y > b->y) { tmp = m; m = b; b = tmp; }
    if (t->y > m->y) { tmp = t; t = m; m = tmp; }
    float dy = b->y - t->y;
    float iy = m->y - t->y;
    if (m->y == t->y) {
        const vtx *l = m, *rr = t;
        if (l->x > rr->x) { tmp = l; l = rr; rr = tmp; }
        top_flat(r, &q, l, rr, b);
    } else if (m->y == b->y) {
        const vtx *l = m, *rr = b;
        if (l->x > rr->x) { tmp = l; l = rr; rr = tmp; }
        bottom_flat(r, &q, t, l, rr);
    } else {
        vtx v4; /* only x and y of the split vertex are read (Rasterizer.h:344-351) */
        v4.y = m->y;
        v4.x = t->x + ((b->x - t->x) / dy) * iy;
        const vtx *l = m, *rr = &v4;
        if (l->x > rr->x) { tmp = l; l = rr; rr = tmp; }
        bottom_flat(r, &q, t, l, rr);
        top_flat(r, &q, l, rr, b);
    }
}

static int window_finite(const vtx *v0, const vtx *v1, const vtx *v2)
{
    return isfinite(v0->x) && isfinite(v0->y) && isfinite(v1->x) && isfinite(v1->y) &&
           isfinite(v2->x) && isfinite(v2->y);
}

/* Rasterizer::drawTriangleModeTemplate / drawTriangleAdaptiveTemplate, Rasterizer.h:413-445 */
static void draw_triangle(const rctx *r, const vtx *v0, const vtx *v1, const vtx *v2)
{
    if (!window_finite(v0, v1, v2))
        return;
    {
        edge_t e0, e1, e2;
        edge_init(&e0, v1, v2);
        edge_init(&e1, v2, v0);
        edge_init(&e2, v0, v1);
        if (!isfinite(e0.c + e1.c + e2.c))
            return;
    }
    switch (r->st->raster_mode) {
    case SWR_RASTER_SPAN:
        raster_span(r, v0, v1, v2);
        break;
    case SWR_RASTER_BLOCK:
        raster_block(r, v0, v1, v2);
        break;
    case SWR_RASTER_ADAPTIVE: {
        float minX = min_f(min_f(v0->x, v1->x), v2->x);
        float maxX = max_f(max_f(v0->x, v1->x), v2->x);
        float minY = min_f(min_f(v0->y, v1->y), v2->y);
        float maxY = max_f(max_f(v0->y, v1->y), v2->y);
        float orient = (maxX - minX) / (maxY - minY);
        /* the thresholds are double literals: the float is promoted */
        if ((double)orient > 0.4 && (double)orient < 1.6)
            raster_block(r, v0, v1, v2);
        else
            raster_span(r, v0, v1, v2);
        break;
    }
    default:
        break;
    }
}

/* ---------------------------------------------------------------- points and lines */

static int scissor_test(const rctx *r, float x, float y)
{
    return x >= (float)r->minX && x < (float)r->maxX && y >= (float)r->minY && y < (float)r->maxY;
}

/* Rasterizer::pixelDataFromVertex, Rasterizer.h:160-173 */
static void pix_from_vertex(const rctx *r, const vtx *v, pix_t *p)
{
    p->eq = NULL;
    p->x = f2i(v->x);
    p->y = f2i(v->y);
    if (r->iz)
        p->z = v->z;
    if (r->iw) {
        p->w = v->w;
        p->invw = 1.0f / v->w;
    }
    for (int i = 0; i < r->na; ++i)
        p->avar[i] = v->avar[i];
    for (int i = 0; i < r->np; ++i)
        p->pvar[i] = v->pvar[i];
}

/* Rasterizer::drawPointTemplate, Rasterizer.h:149-158 */
static void draw_point(const rctx *r, const vtx *v)
{
    if (!scissor_test(r, v->x, v->y))
        return;
    pix_t p;
    pix_from_vertex(r, v, &p);
    draw_pixel(r, &p);
}

/* Rasterizer::drawLineTemplate / computeVertexStep / stepVertex, Rasterizer.h:175-222 */
static void draw_line(const rctx *r, const vtx *v0, const vtx *v1)
{
    int adx = abs(f2i(v1->x) - f2i(v0->x));
    int ady = abs(f2i(v1->y) - f2i(v0->y));
    int steps = max_i(adx, ady);
    float fs = (float)steps;
    vtx step, v = *v0;
    memset(&step, 0, sizeof step);
    step.x = (v1->x - v0->x) / fs;
    step.y = (v1->y - v0->y) / fs;
    if (r->iz)
        step.z = (v1->z - v0->z) / fs;
    if (r->iw)
        step.w = (v1->w - v0->w) / fs;
    for (int i = 0; i < r->na; ++i)
        step.avar[i] = (v1->avar[i] - v0->avar[i]) / fs;
    for (int i = 0; i < r->np; ++i)
        step.pvar[i] = (v1->pvar[i] - v0->pvar[i]) / fs;
    while (steps-- > 0) {
        pix_t p;
        pix_from_vertex(r, &v, &p);
        if (scissor_test(r, v.x, v.y))
            draw_pixel(r, &p);
        v.x += step.x;
        v.y += step.y;
        if (r->iz)
            v.z += step.z;
        if (r->iw)
            v.w += step.w;
        for (int i = 0; i < r->na; ++i)
            v.avar[i] += step.avar[i];
        for (int i = 0; i < r->np; ++i)
            v.pvar[i] += step.pvar[i];
    }
}

/* ---------------------------------------------------------------- IRasterizer lists */

static uint32_t make_key(uint32_t batch, size_t slot) { return (batch << 15) + (uint32_t)slot; }

/* Rasterizer::drawTriangleList, Rasterizer.h:134-141.  src: batch-local input triangle of
 * each slot (NULL: the slot itself). */
static void draw_triangle_list(rctx *r, uint32_t batch, const vtx *v, const int32_t *idx, size_t n,
                               const uint32_t *src)
{
    for (size_t i = 0; i + 3 <= n; i += 3) {
        if (idx[i] == -1)
            continue;
        r->key = make_key(batch, i / 3);
        r->src = src ? batch * (uint32_t)SWR_BATCH_PRIMITIVES + src[i / 3] : (uint32_t)(i / 3);
        draw_triangle(r, &v[idx[i]], &v[idx[i + 1]], &v[idx[i + 2]]);
    }
}

/* Rasterizer::drawLineList, Rasterizer.h:125-132 */
static void draw_line_list(rctx *r, uint32_t batch, const vtx *v, const int32_t *idx, size_t n)
{
    for (size_t i = 0; i + 2 <= n; i += 2) {
        if (idx[i] == -1)
            continue;
        r->key = make_key(batch, i / 2);
        r->src = batch * (uint32_t)SWR_BATCH_PRIMITIVES + (uint32_t)(i / 2);
        draw_line(r, &v[idx[i]], &v[idx[i + 1]]);
    }
}

/* Rasterizer::drawPointList, Rasterizer.h:116-123 */
static void draw_point_list(rctx *r, uint32_t batch, const vtx *v, const int32_t *idx, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        if (idx[i] == -1)
            continue;
        r->key = make_key(batch, i);
        r->src = batch * (uint32_t)SWR_BATCH_PRIMITIVES + (uint32_t)i;
        draw_point(r, &v[idx[i]]);
    }
}

/* The texture program needs its texture and a triangle's planes (points and lines have
 * none: their PixelData::equations is null, PixelData.h:58). */
static int check_texture(const swr_pipeline_state *st, int mode)
{
    if (st->ps.program != SWR_PS_TEXTURE)
        return SWR_OK;
    if (!st->ps.texture || st->ps.pvar_count < 2)
        return SWR_ERR_INVALID_ARG;
    return mode == SWR_DRAW_TRIANGLE ? SWR_OK : SWR_ERR_UNSUPPORTED;
}

static int check_list(const swr_pipeline_state *st, const vtx *v, size_t nv, const int32_t *idx,
                      size_t n)
{
    if (!st || (n && (!v || !idx)))
        return SWR_ERR_INVALID_ARG;
    for (size_t i = 0; i < n; ++i)
        if (idx[i] != -1 && (idx[i] < 0 || (size_t)idx[i] >= nv))
            return SWR_ERR_INDEX_RANGE;
    return SWR_OK;
}

int swr_oracle_draw_triangle_list(const swr_pipeline_state *st, const vtx *v, size_t nv,
                                  const int32_t *idx, size_t n)
{
    int rc = check_list(st, v, nv, idx, n);
    if (!rc)
        rc = check_texture(st, SWR_DRAW_TRIANGLE);
    if (rc)
        return rc;
    rctx r;
    rctx_init(&r, st);
    draw_triangle_list(&r, 0, v, idx, n, NULL);
    return SWR_OK;
}

int swr_oracle_draw_line_list(const swr_pipeline_state *st, const vtx *v, size_t nv,
                              const int32_t *idx, size_t n)
{
    int rc = check_list(st, v, nv, idx, n);
    if (!rc)
        rc = check_texture(st, SWR_DRAW_LINE);
    if (rc)
        return rc;
    rctx r;
    rctx_init(&r, st);
    draw_line_list(&r, 0, v, idx, n);
    return SWR_OK;
}

int swr_oracle_draw_point_list(const swr_pipeline_state *st, const vtx *v, size_t nv,
                               const int32_t *idx, size_t n)
{
    int rc = check_list(st, v, nv, idx, n);
    if (!rc)
        rc = check_texture(st, SWR_DRAW_POINT);
    if (rc)
        return rc;
    rctx r;
    rctx_init(&r, st);
    draw_point_list(&r, 0, v, idx, n);
    return SWR_OK;
}

/* ---------------------------------------------------------------- vertex programs */

static void mvp_apply(const float *m, float x, float y, float z, vtx *o)
{
    float r[4];
    for (int i = 0; i < 4; ++i) {
        /* vector_math.h:55-60: result = 0; result += a[i]*b[i] ... */
        float s = 0.0f;
        s += m[4 * i + 0] * x;
        s += m[4 * i + 1] * y;
        s += m[4 * i + 2] * z;
        s += m[4 * i + 3] * 1.0f;
        r[i] = s;
    }
    o->x = r[0];
    o->y = r[1];
    o->z = r[2];
    o->w = r[3];
}

/* VertexProcessor::initVertexInput / processVertex, VertexProcessor.cpp:134-150, with the
 * built-in vertex programs.  The output starts zeroed, as std::vector::resize leaves it. */
static void run_vertex_program(const swr_pipeline_state *st, int index, vtx *out)
{
    memset(out, 0, sizeof *out);
    const swr_vertex_shader_desc *vs = &st->vs;
    if (vs->program == SWR_VS_DUMMY)
        return;
    const float *a =
        (const float *)((const char *)st->attrib[0] + (size_t)st->stride[0] * (size_t)index);
    switch (vs->program) {
    case SWR_VS_RAW:
        out->x = a[0];
        out->y = a[1];
        out->z = a[2];
        out->w = a[3];
        for (int i = 0; i < vs->avar_count; ++i)
            out->avar[i] = a[4 + i];
        for (int i = 0; i < vs->pvar_count; ++i)
            out->pvar[i] = a[4 + vs->avar_count + i];
        break;
    case SWR_VS_POS_RGB:
        out->x = a[0];
        out->y = a[1];
        out->z = a[2];
        out->w = 1.0f;
        out->avar[0] = a[3];
        out->avar[1] = a[4];
        out->avar[2] = a[5];
        break;
    case SWR_VS_MVP_POS_RGB_UV:
        mvp_apply(vs->mvp, a[0], a[1], a[2], out);
        out->avar[0] = a[3];
        out->avar[1] = a[4];
        out->avar[2] = a[5];
        out->pvar[0] = a[6];
        out->pvar[1] = a[7];
        break;
    case SWR_VS_MVP_POS_UV:
        mvp_apply(vs->mvp, a[0], a[1], a[2], out);
        out->pvar[0] = a[6];
        out->pvar[1] = a[7];
        break;
    default:
        break;
    }
}

/* ---------------------------------------------------------------- VertexProcessor */

/* growable batch arrays (VertexProcessor.h:146-151) */
typedef struct {
    vtx *v;
    size_t nv, cv;
    int32_t *idx;
    size_t ni, ci;
    uint32_t *src;   /* after clipping: the batch-local input triangle of each slot */
} batch_t;

static int push_vertex(batch_t *b, const vtx *v)
{
    if (b->nv == b->cv) {
        size_t nc = b->cv ? 2 * b->cv : 4096;
        vtx *p = (vtx *)realloc(b->v, nc * sizeof(vtx));
        if (!p)
            return -1;
        b->v = p;
        b->cv = nc;
    }
    b->v[b->nv++] = *v;
    return (int)(b->nv - 1);
}

static int push_index(batch_t *b, int32_t i)
{
    if (b->ni == b->ci) {
        size_t nc = b->ci ? 2 * b->ci : 4096;
        int32_t *p = (int32_t *)realloc(b->idx, nc * sizeof(int32_t));
        if (!p)
            return -1;
        b->idx = p;
        b->ci = nc;
    }
    b->idx[b->ni++] = i;
    return 0;
}

/* VertexProcessor::clipMask, VertexProcessor.cpp:122-132 */
static int clip_mask(const vtx *v)
{
    int mask = 0;
    if (v->w - v->x < 0) mask |= 0x01;
    if (v->x + v->w < 0) mask |= 0x02;
    if (v->w - v->y < 0) mask |= 0x04;
    if (v->y + v->w < 0) mask |= 0x08;
    if (v->w - v->z < 0) mask |= 0x10;
    if (v->z + v->w < 0) mask |= 0x20;
    return mask;
}

/* The six clip planes in the order VertexProcessor.cpp:237-242 applies them. */
static const float k_planes[6][4] = {
    {-1, 0, 0, 1}, {1, 0, 0, 1}, {0, -1, 0, 1}, {0, 1, 0, 1}, {0, 0, -1, 1}, {0, 0, 1, 1}};

/* Helper::interpolateVertex, PolyClipper.h:34-48 */
static void lerp_vertex(const vtx *v0, const vtx *v1, float t, int na, int np, vtx *out)
{
    memset(out, 0, sizeof *out);
    out->x = v0->x * (1.0f - t) + v1->x * t;
    out->y = v0->y * (1.0f - t) + v1->y * t;
    out->z = v0->z * (1.0f - t) + v1->z * t;
    out->w = v0->w * (1.0f - t) + v1->w * t;
    for (int i = 0; i < na; ++i)
        out->avar[i] = v0->avar[i] * (1.0f - t) + v1->avar[i] * t;
    for (int i = 0; i < np; ++i)
        out->pvar[i] = v0->pvar[i] * (1.0f - t) + v1->pvar[i] * t;
}

static float plane_dot(const float *pl, const vtx *v)
{
    return pl[0] * v->x + pl[1] * v->y + pl[2] * v->z + pl[3] * v->w;
}

static int sgn(float v) { return (0.0f < v) - (v < 0.0f); }

/* PolyClipper::clipToPlane, PolyClipper.cpp:45-81.  `in` holds n indices into the batch
 * vertices plus room for one more; returns the new count written to `out`. */
static int clip_poly_plane(batch_t *b, const float *pl, const int32_t *in, int n, int32_t *out,
                           int max_out, int na, int np)
{
    if (n < 3)
        return n;
    int m = 0;
    int32_t iprev = in[0];
    float dprev = plane_dot(pl, &b->v[iprev]);
    for (int i = 1; i <= n; ++i) {
        int32_t ic = (i == n) ? in[0] : in[i];
        float dp = plane_dot(pl, &b->v[ic]);
        if (dprev >= 0) {
            if (m >= max_out)
                return -1;
            out[m++] = iprev;
        }
        if (sgn(dp) != sgn(dprev)) {
            float t = dp < 0 ? dprev / (dprev - dp) : -dprev / (dp - dprev);
            vtx nv;
            lerp_vertex(&b->v[iprev], &b->v[ic], t, na, np, &nv);
            int k = push_vertex(b, &nv);
            if (k < 0 || m >= max_out)
                return -1;
            out[m++] = k;
        }
        iprev = ic;
        dprev = dp;
    }
    return m;
}

#define POLY_MAX 256

/* VertexProcessor::clipTriangles, VertexProcessor.cpp:217-263 */
static int clip_triangles(batch_t *b, int na, int np)
{
    size_t nvert = b->nv;
    int *mask = (int *)malloc((nvert ? nvert : 1) * sizeof(int));
    if (!mask)
        return SWR_ERR_CAPACITY;
    for (size_t i = 0; i < nvert; ++i)
        mask[i] = clip_mask(&b->v[i]);
    size_t n = b->ni;
    int32_t pa[POLY_MAX], pb[POLY_MAX];
    /* source triangle of every slot: slot j < n/3 is triangle j; fans append theirs */
    size_t cap_src = n / 3 + 64;
    uint32_t *src = (uint32_t *)realloc(b->src, cap_src * sizeof(uint32_t));
    if (!src) {
        free(mask);
        return SWR_ERR_CAPACITY;
    }
    b->src = src;
    for (size_t j = 0; j < n / 3; ++j)
        b->src[j] = (uint32_t)j;
    size_t nsrc = n / 3;
    for (size_t i = 0; i < n; i += 3) {
        int32_t i0 = b->idx[i], i1 = b->idx[i + 1], i2 = b->idx[i + 2];
        int cm = mask[i0] | mask[i1] | mask[i2];
        int32_t *in = pa, *out = pb, *sw;
        int cnt = 3;
        in[0] = i0;
        in[1] = i1;
        in[2] = i2;
        for (int p = 0; p < 6; ++p) {
            if (!(cm & (1 << p)))
                continue;
            int m = clip_poly_plane(b, k_planes[p], in, cnt, out, POLY_MAX, na, np);
            if (m < 0) {
                free(mask);
                return SWR_ERR_CLIP_OVERFLOW;
            }
            if (cnt < 3)
                continue; /* fullyClipped: nothing changes */
            cnt = m;
            sw = in;
            in = out;
            out = sw;
        }
        if (cnt < 3) {
            b->idx[i] = b->idx[i + 1] = b->idx[i + 2] = -1;
            continue;
        }
        b->idx[i] = in[0];
        b->idx[i + 1] = in[1];
        b->idx[i + 2] = in[2];
        for (int k = 3; k < cnt; ++k) {
            if (push_index(b, in[0]) || push_index(b, in[k - 1]) || push_index(b, in[k])) {
                free(mask);
                return SWR_ERR_CAPACITY;
            }
            if (nsrc == cap_src) {
                cap_src *= 2;
                uint32_t *q = (uint32_t *)realloc(b->src, cap_src * sizeof(uint32_t));
                if (!q) {
                    free(mask);
                    return SWR_ERR_CAPACITY;
                }
                b->src = q;
            }
            b->src[nsrc++] = (uint32_t)(i / 3);
        }
    }
    free(mask);
    return SWR_OK;
}

/* LineClipper::clipToPlane, LineClipper.cpp:30-56 */
static void clip_line_plane(const float *pl, const vtx *v0, const vtx *v1, float *t0, float *t1,
                            int *clipped)
{
    if (*clipped)
        return;
    float dp0 = plane_dot(pl, v0);
    float dp1 = plane_dot(pl, v1);
    int n0 = dp0 < 0, n1 = dp1 < 0;
    if (n0 && n1) {
        *clipped = 1;
        return;
    }
    if (n0) {
        float t = -dp0 / (dp1 - dp0);
        *t0 = max_f(*t0, t);
    } else {
        float t = dp0 / (dp0 - dp1);
        *t1 = min_f(*t1, t);
    }
}

/* VertexProcessor::clipLines, VertexProcessor.cpp:167-215 */
static int clip_lines(batch_t *b, int na, int np)
{
    size_t nvert = b->nv;
    int *mask = (int *)malloc((nvert ? nvert : 1) * sizeof(int));
    if (!mask)
        return SWR_ERR_CAPACITY;
    for (size_t i = 0; i < nvert; ++i)
        mask[i] = clip_mask(&b->v[i]);
    for (size_t i = 0; i + 1 < b->ni; i += 2) {
        int32_t i0 = b->idx[i], i1 = b->idx[i + 1];
        vtx v0 = b->v[i0], v1 = b->v[i1];
        int cm = mask[i0] | mask[i1];
        float t0 = 0.0f, t1 = 1.0f;
        int clipped = 0;
        for (int p = 0; p < 6; ++p)
            if (cm & (1 << p))
                clip_line_plane(k_planes[p], &v0, &v1, &t0, &t1, &clipped);
        if (clipped) {
            b->idx[i] = b->idx[i + 1] = -1;
            continue;
        }
        if (mask[i0]) {
            vtx nv;
            lerp_vertex(&v0, &v1, t0, na, np, &nv);
            int k = push_vertex(b, &nv);
            if (k < 0) {
                free(mask);
                return SWR_ERR_CAPACITY;
            }
            b->idx[i] = k;
        }
        if (mask[i1]) {
            vtx nv;
            lerp_vertex(&v0, &v1, t1, na, np, &nv);
            int k = push_vertex(b, &nv);
            if (k < 0) {
                free(mask);
                return SWR_ERR_CAPACITY;
            }
            b->idx[i + 1] = k;
        }
    }
    free(mask);
    return SWR_OK;
}

/* VertexProcessor::clipPoints, VertexProcessor.cpp:152-165 */
static void clip_points(batch_t *b)
{
    for (size_t i = 0; i < b->ni; ++i)
        if (clip_mask(&b->v[b->idx[i]]))
            b->idx[i] = -1;
}

/* VertexProcessor::transformVertices, VertexProcessor.cpp:347-377 */
static int transform_vertices(batch_t *b, const swr_pipeline_state *st)
{
    unsigned char *done = (unsigned char *)calloc(b->nv ? b->nv : 1, 1);
    if (!done)
        return SWR_ERR_CAPACITY;
    float px = (float)st->viewport[2] / 2.0f, py = (float)st->viewport[3] / 2.0f;
    float ox = (float)st->viewport[0] + px, oy = (float)st->viewport[1] + py;
    float n = st->depth_near, f = st->depth_far;
    for (size_t i = 0; i < b->ni; ++i) {
        int32_t k = b->idx[i];
        if (k == -1 || done[k])
            continue;
        vtx *v = &b->v[k];
        float invW = 1.0f / v->w;
        v->x *= invW;
        v->y *= invW;
        v->z *= invW;
        v->x = (px * v->x + ox);
        v->y = (py * -v->y + oy);
        v->z = 0.5f * (f - n) * v->z + 0.5f * (n + f);
        done[k] = 1;
    }
    free(done);
    return SWR_OK;
}

/* VertexProcessor::cullTriangles, VertexProcessor.cpp:319-345 */
static void cull_triangles(batch_t *b, int cull_mode)
{
    for (size_t i = 0; i + 3 <= b->ni; i += 3) {
        if (b->idx[i] == -1)
            continue;
        const vtx *v0 = &b->v[b->idx[i]], *v1 = &b->v[b->idx[i + 1]], *v2 = &b->v[b->idx[i + 2]];
        float facing = (v0->x - v1->x) * (v2->y - v1->y) - (v2->x - v1->x) * (v0->y - v1->y);
        if (facing < 0) {
            if (cull_mode == SWR_CULL_CW)
                b->idx[i] = b->idx[i + 1] = b->idx[i + 2] = -1;
        } else {
            if (cull_mode == SWR_CULL_CCW) {
                b->idx[i] = b->idx[i + 1] = b->idx[i + 2] = -1;
            } else {
                int32_t s = b->idx[i];
                b->idx[i] = b->idx[i + 2];
                b->idx[i + 2] = s;
            }
        }
    }
}

/* VertexProcessor::processPrimitives, VertexProcessor.cpp:281-317 */
static int process_primitives(batch_t *b, rctx *r, int mode, uint32_t batch)
{
    const swr_pipeline_state *st = r->st;
    int rc = SWR_OK;
    if (mode == SWR_DRAW_TRIANGLE)
        rc = clip_triangles(b, st->vs.avar_count, st->vs.pvar_count);
    else if (mode == SWR_DRAW_LINE)
        rc = clip_lines(b, st->vs.avar_count, st->vs.pvar_count);
    else
        clip_points(b);
    if (rc)
        return rc;
    rc = transform_vertices(b, st);
    if (rc)
        return rc;
    if (mode == SWR_DRAW_TRIANGLE) {
        cull_triangles(b, st->cull_mode);
        draw_triangle_list(r, batch, b->v, b->idx, b->ni, b->src);
    } else if (mode == SWR_DRAW_LINE) {
        draw_line_list(r, batch, b->v, b->idx, b->ni);
    } else {
        draw_point_list(r, batch, b->v, b->idx, b->ni);
    }
    return SWR_OK;
}

/* VertexCache, VertexCache.h:29-63: 16-entry direct-mapped index -> output vertex map.
 * The vertex program is pure, so the cache changes no result; it is kept for fidelity. */
typedef struct { int in[16]; int out[16]; } vcache_t;
static void vcache_clear(vcache_t *c) { for (int i = 0; i < 16; ++i) c->in[i] = -1; }

/* VertexProcessor::drawElements, VertexProcessor.cpp:78-120 */
int swr_oracle_draw_elements(const swr_pipeline_state *st, int mode, size_t count,
                             const int32_t *indices)
{
    if (!st || (count && !indices) || mode < SWR_DRAW_POINT || mode > SWR_DRAW_TRIANGLE)
        return SWR_ERR_INVALID_ARG;
    if (st->vs.program != SWR_VS_DUMMY && !st->attrib[0])
        return SWR_ERR_INVALID_ARG;
    int rc0 = check_texture(st, mode);
    if (rc0)
        return rc0;
    for (size_t i = 0; i < count; ++i)
        if (indices[i] < 0)
            return SWR_ERR_INDEX_RANGE;
    int factor = mode == SWR_DRAW_TRIANGLE ? 3 : (mode == SWR_DRAW_LINE ? 2 : 1);
    size_t usable = count - count % (size_t)factor; /* a partial last primitive is dropped */
    rctx r;
    rctx_init(&r, st);
    batch_t b;
    memset(&b, 0, sizeof b);
    vcache_t cache;
    vcache_clear(&cache);
    uint32_t batch = 0;
    int rc = SWR_OK;
    for (size_t i = 0; i < usable && rc == SWR_OK; ++i) {
        int index = indices[i];
        int slot = index % 16;
        if (cache.in[slot] == index) {
            rc = push_index(&b, cache.out[slot]) ? SWR_ERR_CAPACITY : SWR_OK;
        } else {
            vtx v;
            run_vertex_program(st, index, &v);
            int k = push_vertex(&b, &v);
            if (k < 0 || push_index(&b, k)) {
                rc = SWR_ERR_CAPACITY;
                break;
            }
            cache.in[slot] = index;
            cache.out[slot] = k;
        }
        if (rc == SWR_OK && b.ni / (size_t)factor >= SWR_BATCH_PRIMITIVES) {
            rc = process_primitives(&b, &r, mode, batch++);
            b.nv = b.ni = 0;
            vcache_clear(&cache);
        }
    }
    if (rc == SWR_OK)
        rc = process_primitives(&b, &r, mode, batch);
    free(b.v);
    free(b.idx);
    free(b.src);
    return rc;
}

/* ---------------------------------------------------------------- exported probes */

int swr_oracle_triangle_setup(const vtx *v0, const vtx *v1, const vtx *v2, int na, int np,
                              float *out)
{
    tri_t q;
    memset(&q, 0, sizeof q);
    tri_setup(&q, v0, v1, v2, na, np);
    int k = 0;
    out[k++] = q.area2;
    const edge_t *E[3] = {&q.e0, &q.e1, &q.e2};
    for (int e = 0; e < 3; ++e) {
        out[k++] = E[e]->a;
        out[k++] = E[e]->b;
        out[k++] = E[e]->c;
    }
    if (q.area2 <= 0.0f)
        return 0;
    const param_t *P[2] = {&q.z, &q.invw};
    for (int p = 0; p < 2; ++p) {
        out[k++] = P[p]->a;
        out[k++] = P[p]->b;
        out[k++] = P[p]->c;
    }
    for (int i = 0; i < na; ++i) {
        out[k++] = q.avar[i].a;
        out[k++] = q.avar[i].b;
        out[k++] = q.avar[i].c;
    }
    for (int i = 0; i < np; ++i) {
        out[k++] = q.pvar[i].a;
        out[k++] = q.pvar[i].b;
        out[k++] = q.pvar[i].c;
    }
    return k;
}

int swr_oracle_clip_triangle(const vtx *v0, const vtx *v1, const vtx *v2, int na, int np, vtx *out,
                             int max_out)
{
    batch_t b;
    memset(&b, 0, sizeof b);
    if (push_vertex(&b, v0) < 0 || push_vertex(&b, v1) < 0 || push_vertex(&b, v2) < 0 ||
        push_index(&b, 0) || push_index(&b, 1) || push_index(&b, 2)) {
        free(b.v);
        free(b.idx);
        return SWR_ERR_CAPACITY;
    }
    /* rerun the per-triangle part of clipTriangles keeping the whole polygon */
    int cm = clip_mask(v0) | clip_mask(v1) | clip_mask(v2);
    int32_t pa[POLY_MAX], pb[POLY_MAX];
    int32_t *in = pa, *o = pb, *sw;
    int cnt = 3;
    in[0] = 0;
    in[1] = 1;
    in[2] = 2;
    for (int p = 0; p < 6; ++p) {
        if (!(cm & (1 << p)))
            continue;
        int m = clip_poly_plane(&b, k_planes[p], in, cnt, o, POLY_MAX, na, np);
        if (m < 0) {
            free(b.v);
            free(b.idx);
            return SWR_ERR_CLIP_OVERFLOW;
        }
        if (cnt < 3)
            continue;
        cnt = m;
        sw = in;
        in = o;
        o = sw;
    }
    int n = cnt < max_out ? cnt : max_out;
    for (int k = 0; k < n; ++k)
        out[k] = b.v[in[k]];
    free(b.v);
    free(b.idx);
    return cnt;
}
