/*
 * swr_diff_ref.c — TEST INFRASTRUCTURE ONLY: the FP64 CPU reference of the differentiable
 * shading pass (include/swr_hip.h, swr_diff_forward / swr_diff_backward).  Only tests/,
 * the smoke check and the bench's CPU baseline use it, as the checker.
 *
 * The reference renderer has no differentiable path, so this is the specification restated
 * in double precision: barycentrics D_i = r . (h_j x h_k), b_i = D_i / sum D; colour
 * sum b_i A_i; the silhouette antialiasing blend between 4-neighbour pixels (the walk to
 * the first silhouette edge along the segment between the two centres, axis weights from
 * the edge direction, the colour at the crossing for the pixel beyond the edge); and the exact
 * derivatives of all three.  The discrete choices of the antialiasing pass (which pixel is
 * in front, which edge is crossed first, whether it is a silhouette, which half of the
 * segment holds the crossing) are made with the same float operations as the device code,
 * so that both sides differentiate the same piecewise-smooth function; every value that
 * enters an image or a gradient is computed in double.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

typedef struct { double x, y, w; } h3d;
typedef struct { float x, y; } f2;
typedef struct { double x, y; } d2;

typedef struct {
    int V, W, H, N, T, C;
    int culled;
    const float *clip;
    const int32_t *tris, *adj;
    const float *attr;
    const uint32_t *ids;
    const float *depth;
} dctx;

static h3d hd(const dctx *d, int view, int v)
{
    const float *c = d->clip + ((size_t)view * d->N + v) * 4;
    h3d r = {c[0], c[1], c[3]};
    return r;
}

static h3d cross3(h3d a, h3d b)
{
    h3d r = {a.y * b.w - a.w * b.y, a.w * b.x - a.x * b.w, a.x * b.y - a.y * b.x};
    return r;
}

static double dot3(h3d a, h3d b) { return a.x * b.x + a.y * b.y + a.w * b.w; }

static h3d ray(const dctx *d, int px, int py)
{
    h3d r = {2.0 * (px + 0.5) / d->W - 1.0, 1.0 - 2.0 * (py + 0.5) / d->H, 1.0};
    return r;
}

/* float window position, as the device computes it (decisions only) */
static f2 win_f(const dctx *d, int view, int v)
{
    const float *c = d->clip + ((size_t)view * d->N + v) * 4;
    float sx = (float)d->W / 2.0f, sy = (float)d->H / 2.0f;
    f2 r;
    r.x = sx * (c[0] / c[3]) + sx;
    r.y = -sy * (c[1] / c[3]) + sy;
    return r;
}

static d2 win_d(const dctx *d, int view, int v)
{
    h3d h = hd(d, view, v);
    double sx = d->W / 2.0, sy = d->H / 2.0;
    d2 r = {sx * (h.x / h.w) + sx, -sy * (h.y / h.w) + sy};
    return r;
}

static float edge_f(f2 j, f2 k, f2 p) { return (k.x - j.x) * (p.y - j.y) - (k.y - j.y) * (p.x - j.x); }
static double edge_d(d2 j, d2 k, d2 p) { return (k.x - j.x) * (p.y - j.y) - (k.y - j.y) * (p.x - j.x); }

typedef struct {
    int active, p_blended, front_is_p, vj, vk, horiz, tri;
    float tf;   /* float crossing (decisions) */
    double s;
} pair_t;

static int wpos(const dctx *d, int view, int v) { return d->clip[((size_t)view * d->N + v) * 4 + 3] > 0.0f; }

static int tri_win(const dctx *d, int view, int t, f2 *v, int *vid)
{
    for (int i = 0; i < 3; ++i) {
        vid[i] = d->tris[3 * t + i];
        if (!wpos(d, view, vid[i]))
            return 0;
        v[i] = win_f(d, view, vid[i]);
    }
    return 1;
}

#define MARCH_STEPS 8

/* Walk from the front pixel's triangle along the segment F -> Q across interior edges to
 * the first silhouette edge. */
static pair_t eval_pair(const dctx *d, int view, int px, int py, int qx, int qy)
{
    pair_t r;
    memset(&r, 0, sizeof r);
    size_t base = (size_t)view * d->W * d->H;
    size_t ip = base + (size_t)py * d->W + px, iq = base + (size_t)qy * d->W + qx;
    uint32_t a = d->ids[ip], b = d->ids[iq];
    if (a == b)
        return r;
    int pf;
    if (a == 0)
        pf = 0;
    else if (b == 0)
        pf = 1;
    else {
        float dp = d->depth[ip], dq = d->depth[iq];
        pf = dp < dq || (dp == dq && ip < iq);
    }
    int cur = (int)(pf ? a : b) - 1;
    if (cur < 0 || cur >= d->T)
        return r;
    if (d->culled && a != 0 && b != 0) {
        /* two visible triangles of a culled draw that share a vertex: no silhouette taken
         * to separate them */
        int other = (int)(pf ? b : a) - 1;
        if (other >= 0 && other < d->T) {
            const int32_t *c = d->tris + 3 * cur, *o = d->tris + 3 * other;
            int share = 0;
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y)
                    share |= c[x] == o[y];
            if (share)
                return r;
        }
    }
    int horiz = py == qy;
    f2 F = {(float)(pf ? px : qx) + 0.5f, (float)(pf ? py : qy) + 0.5f};
    f2 Q = {(float)(pf ? qx : px) + 0.5f, (float)(pf ? qy : py) + 0.5f};
    f2 v[3];
    int vid[3];
    if (!tri_win(d, view, cur, v, vid))
        return r;
    float area = edge_f(v[0], v[1], v[2]);
    if (!(area != 0.0f))
        return r;
    int from = -1;
    for (int step = 0; step < MARCH_STEPS; ++step) {
        float s = area > 0.0f ? 1.0f : -1.0f;
        float best = 2.0f;
        int be = -1;
        for (int e = 0; e < 3; ++e) {
            if (e == from)
                continue;
            int j = (e + 1) % 3, k = (e + 2) % 3;
            float ef = s * edge_f(v[j], v[k], F), eq = s * edge_f(v[j], v[k], Q);
            if (eq < 0.0f && ef > eq) {
                float tt = ef / (ef - eq);
                if (tt < best) {
                    best = tt;
                    be = e;
                }
            }
        }
        if (be < 0)
            return r;
        int n = d->adj[3 * cur + be];
        f2 nv[3];
        int nvid[3];
        float an = 0.0f;
        int interior = 0;
        if (n >= 0 && n < d->T && tri_win(d, view, n, nv, nvid)) {
            an = edge_f(nv[0], nv[1], nv[2]);
            interior = an != 0.0f && ((an > 0.0f) == (area > 0.0f));
        }
        if (!interior) {
            int j = (be + 1) % 3, k = (be + 2) % 3;
            if (v[k].x == v[j].x && v[k].y == v[j].y)
                return r;
            r.active = 1;
            r.horiz = horiz;
            r.tri = cur;
            r.tf = best;
            r.s = s;
            r.vj = vid[j];
            r.vk = vid[k];
            r.front_is_p = pf;
            r.p_blended = (best < 0.5f) == pf;
            return r;
        }
        int fe = -1;
        for (int e = 0; e < 3; ++e)
            if (d->adj[3 * n + e] == cur)
                fe = e;
        if (fe < 0)
            return r;
        cur = n;
        from = fe;
        area = an;
        for (int i = 0; i < 3; ++i) {
            v[i] = nv[i];
            vid[i] = nvid[i];
        }
    }
    return r;
}

/* t and the axis weight a of an active pair in double, with d/d(window j), d/d(window k):
 * a = ey^2 / (ex^2 + ey^2) for a horizontal pair, ex^2 / (ex^2 + ey^2) for a vertical one */
static double pair_t_d(const dctx *d, int view, const pair_t *pr, int px, int py, int qx, int qy,
                       double *gj, double *gk, double *a, double *aj, double *ak)
{
    d2 j = win_d(d, view, pr->vj), k = win_d(d, view, pr->vk);
    d2 P = {px + 0.5, py + 0.5}, Qp = {qx + 0.5, qy + 0.5};
    d2 F = pr->front_is_p ? P : Qp, Q = pr->front_is_p ? Qp : P;
    double ef = pr->s * edge_d(j, k, F), eq = pr->s * edge_d(j, k, Q);
    double den = ef - eq;
    double dtf = -eq / (den * den), dtq = ef / (den * den);
    gj[0] = pr->s * (dtf * (k.y - F.y) + dtq * (k.y - Q.y));
    gj[1] = pr->s * (dtf * (F.x - k.x) + dtq * (Q.x - k.x));
    gk[0] = pr->s * (dtf * (F.y - j.y) + dtq * (Q.y - j.y));
    gk[1] = pr->s * (dtf * (j.x - F.x) + dtq * (j.x - Q.x));
    double ex = k.x - j.x, ey = k.y - j.y, L2 = ex * ex + ey * ey;
    double wx = ey * ey / L2;
    double dex = -2.0 * ex * ey * ey / (L2 * L2), dey = 2.0 * ey * ex * ex / (L2 * L2);
    double sg = pr->horiz ? 1.0 : -1.0;
    *a = pr->horiz ? wx : 1.0 - wx;
    ak[0] = sg * dex;
    ak[1] = sg * dey;
    aj[0] = -sg * dex;
    aj[1] = -sg * dey;
    return ef / den;
}

static int nb(const dctx *d, int px, int py, int k, int *qx, int *qy)
{
    *qx = px + (k == 0 ? 1 : (k == 1 ? -1 : 0));
    *qy = py + (k == 2 ? 1 : (k == 3 ? -1 : 0));
    return *qx >= 0 && *qx < d->W && *qy >= 0 && *qy < d->H;
}

/* perspective-correct barycentrics of triangle t at pixel (px, py) and the colour they
 * give, in double; 0 when S == 0 */
typedef struct {
    int v[3];
    h3d h[3], r;
    double b[3], S;
} bary_t;

static int bary_at_ray(const dctx *d, int view, int t, h3d r, bary_t *B)
{
    for (int m = 0; m < 3; ++m) {
        B->v[m] = d->tris[3 * t + m];
        B->h[m] = hd(d, view, B->v[m]);
    }
    B->r = r;
    double D[3] = {dot3(B->r, cross3(B->h[1], B->h[2])), dot3(B->r, cross3(B->h[2], B->h[0])),
                   dot3(B->r, cross3(B->h[0], B->h[1]))};
    B->S = D[0] + D[1] + D[2];
    if (B->S == 0.0)
        return 0;
    for (int m = 0; m < 3; ++m)
        B->b[m] = D[m] / B->S;
    return 1;
}

static int bary_at(const dctx *d, int view, int t, int px, int py, bary_t *B)
{
    return bary_at_ray(d, view, t, ray(d, px, py), B);
}

/* the segment's crossing of the silhouette edge as a ray, and d(ray)/dt */
static h3d crossing_ray(const dctx *d, const pair_t *pr, double t, int px, int py, int qx, int qy,
                        h3d *drdt)
{
    double fx = (pr->front_is_p ? px : qx) + 0.5, fy = (pr->front_is_p ? py : qy) + 0.5;
    double ox = (pr->front_is_p ? qx : px) + 0.5, oy = (pr->front_is_p ? qy : py) + 0.5;
    double wx = fx + t * (ox - fx), wy = fy + t * (oy - fy);
    h3d dr = {2.0 * (ox - fx) / d->W, -2.0 * (oy - fy) / d->H, 0.0};
    *drdt = dr;
    h3d r = {2.0 * wx / d->W - 1.0, 1.0 - 2.0 * wy / d->H, 1.0};
    return r;
}

/* d(colour at ray r)/dr . dr per channel */
static void bary_color_dr(const dctx *d, const bary_t *B, h3d dr, double *out)
{
    h3d c0 = cross3(B->h[1], B->h[2]), c1 = cross3(B->h[2], B->h[0]), c2 = cross3(B->h[0], B->h[1]);
    double e[3] = {dot3(c0, dr), dot3(c1, dr), dot3(c2, dr)};
    double es = e[0] + e[1] + e[2];
    double db[3];
    for (int m = 0; m < 3; ++m)
        db[m] = (e[m] - B->b[m] * es) / B->S;
    for (int c = 0; c < d->C; ++c)
        out[c] = db[0] * d->attr[(size_t)B->v[0] * d->C + c] +
                 db[1] * d->attr[(size_t)B->v[1] * d->C + c] +
                 db[2] * d->attr[(size_t)B->v[2] * d->C + c];
}

static void bary_color(const dctx *d, const bary_t *B, double *col)
{
    for (int c = 0; c < d->C; ++c)
        col[c] = B->b[0] * d->attr[(size_t)B->v[0] * d->C + c] +
                 B->b[1] * d->attr[(size_t)B->v[1] * d->C + c] +
                 B->b[2] * d->attr[(size_t)B->v[2] * d->C + c];
}

/* colour of pixel (px, py) before antialiasing (0 on the background) */
static void pixel_color(const dctx *d, int view, int px, int py, double *col)
{
    size_t i = (size_t)view * d->W * d->H + (size_t)py * d->W + px;
    for (int c = 0; c < d->C; ++c)
        col[c] = 0.0;
    uint32_t id = d->ids[i];
    bary_t B;
    if (id == 0 || (int)id > d->T || !bary_at(d, view, (int)id - 1, px, py, &B))
        return;
    bary_color(d, &B, col);
}

/* backward of col = sum b_m A_m, given dL/dcol */
static void bary_backward(const dctx *d, int view, const bary_t *B, const double *gcol,
                          double *gclip, double *gattr)
{
    const int C = d->C;
    double u[3];
    for (int m = 0; m < 3; ++m) {
        u[m] = 0.0;
        for (int c = 0; c < C; ++c) {
            u[m] += gcol[c] * d->attr[(size_t)B->v[m] * C + c];
            gattr[(size_t)B->v[m] * C + c] += B->b[m] * gcol[c];
        }
    }
    double ub = u[0] * B->b[0] + u[1] * B->b[1] + u[2] * B->b[2];
    for (int m = 0; m < 3; ++m) {
        double gD = (u[m] - ub) / B->S;
        int j = (m + 1) % 3, k = (m + 2) % 3;
        h3d dj = cross3(B->h[k], B->r), dk = cross3(B->r, B->h[j]);
        double *g1 = gclip + ((size_t)view * d->N + B->v[j]) * 4;
        double *g2 = gclip + ((size_t)view * d->N + B->v[k]) * 4;
        g1[0] += gD * dj.x;
        g1[1] += gD * dj.y;
        g1[3] += gD * dj.w;
        g2[0] += gD * dk.x;
        g2[1] += gD * dk.y;
        g2[3] += gD * dk.w;
    }
}

static void add_window_grad(const dctx *d, int view, int v, double gx, double gy, double *gclip)
{
    h3d h = hd(d, view, v);
    double sx = d->W / 2.0, sy = d->H / 2.0;
    double *g = gclip + ((size_t)view * d->N + v) * 4;
    g[0] += gx * sx / h.w;
    g[1] += -gy * sy / h.w;
    g[3] += (-gx * sx * h.x + gy * sy * h.y) / (h.w * h.w);
}

/* Forward image (double) and, when grad_image is given, the gradients (added to gclip /
 * gattr, double).  A pixel blended by a pair receives, as the front pixel (t < 1/2),
 * a (1/2 - t) (C_q - C_p); as the other pixel, a (t - 1/2) (E - C_p) with E the silhouette
 * triangle's colour where the segment crosses the edge.  Returns 0, or -1 for channels outside 1..4. */
int swr_oracle_diff(int views, int width, int height, int nverts, int ntris, int channels,
                    int culled, const float *clip, const int32_t *tris, const int32_t *adj, const float *attr,
                    const uint32_t *ids, const float *depth, double *image,
                    const float *grad_image, double *gclip, double *gattr)
{
    if (channels < 1 || channels > 4)
        return -1;
    dctx d = {views, width, height, nverts, ntris, channels, culled, clip, tris, adj, attr, ids,
              depth};
    const int C = channels;
    for (int view = 0; view < views; ++view) {
        for (int py = 0; py < height; ++py) {
            for (int px = 0; px < width; ++px) {
                size_t i = ((size_t)view * height + py) * width + px;
                double col[4];
                pixel_color(&d, view, px, py, col);
                double out[4];
                for (int c = 0; c < C; ++c)
                    out[c] = col[c];
                double gc[4] = {0, 0, 0, 0};
                const float *gp = grad_image ? grad_image + i * C : NULL;
                if (gp)
                    for (int c = 0; c < C; ++c)
                        gc[c] = gp[c];
                for (int k = 0; k < 4; ++k) {
                    int qx, qy;
                    if (!nb(&d, px, py, k, &qx, &qy))
                        continue;
                    pair_t pr = eval_pair(&d, view, px, py, qx, qy);
                    if (!pr.active)
                        continue;
                    double gj[2], gk[2], a, aj[2], ak[2];
                    double t = pair_t_d(&d, view, &pr, px, py, qx, qy, gj, gk, &a, aj, ak);
                    int t_low = pr.tf < 0.5f;
                    double wt = t_low ? 0.5 - t : t - 0.5;
                    double w = a * wt;
                    if (!pr.p_blended) {
                        if (gp && t_low) {  /* q (front) reads C_p */
                            const float *gq =
                                grad_image + (((size_t)view * height + qy) * width + qx) * C;
                            for (int c = 0; c < C; ++c)
                                gc[c] += w * gq[c];
                        }
                        continue;
                    }
                    double src[4];
                    double dedt = 0.0;
                    bary_t B;
                    if (t_low) {
                        pixel_color(&d, view, qx, qy, src);
                    } else {
                        h3d drdt;
                        h3d rx = crossing_ray(&d, &pr, t, px, py, qx, qy, &drdt);
                        if (!bary_at_ray(&d, view, pr.tri, rx, &B))
                            continue;
                        bary_color(&d, &B, src);
                        if (gp) {
                            double gsrc[4], de[4];
                            for (int c = 0; c < C; ++c)
                                gsrc[c] = w * gp[c];
                            bary_backward(&d, view, &B, gsrc, gclip, gattr);
                            bary_color_dr(&d, &B, drdt, de);
                            for (int c = 0; c < C; ++c)
                                dedt += gsrc[c] * de[c];
                        }
                    }
                    double dot = 0.0;
                    for (int c = 0; c < C; ++c) {
                        out[c] += w * (src[c] - col[c]);
                        if (gp) {
                            gc[c] -= w * gp[c];
                            dot += gp[c] * (src[c] - col[c]);
                        }
                    }
                    if (gp) {
                        double dldt = a * (t_low ? -dot : dot) + dedt, dlda = wt * dot;
                        add_window_grad(&d, view, pr.vj, dldt * gj[0] + dlda * aj[0],
                                        dldt * gj[1] + dlda * aj[1], gclip);
                        add_window_grad(&d, view, pr.vk, dldt * gk[0] + dlda * ak[0],
                                        dldt * gk[1] + dlda * ak[1], gclip);
                    }
                }
                if (image)
                    for (int c = 0; c < C; ++c)
                        image[i * C + c] = out[c];
                bary_t B;
                if (!gp || ids[i] == 0 || (int)ids[i] > ntris ||
                    !bary_at(&d, view, (int)ids[i] - 1, px, py, &B))
                    continue;
                bary_backward(&d, view, &B, gc, gclip, gattr);
            }
        }
    }
    return 0;
}
