// swr_diff.hip — differentiable shading over a visibility buffer, forward and backward.
//
// The visibility draw (SWR_PS_VISIBILITY, one batched draw over every view) leaves per
// pixel the nearest input triangle and its depth.  Everything else is recomputed from it:
//
//   barycentrics   r = (X, Y, 1), the pixel centre in NDC; h_i = (x_i, y_i, w_i) the clip
//                  coordinates of the triangle's vertices.  D_i = r . (h_j x h_k) over the
//                  cyclic (i, j, k); b_i = D_i / (D_0 + D_1 + D_2): the perspective-correct
//                  weights (sum_i b_i h_i is parallel to r).
//   colour         C = sum_i b_i A_i (per channel), 0 on the background.
//   edge antialiasing
//                  for every pair of 4-neighbour pixels (p, q) showing different triangles:
//                  the front pixel f is the one with the smaller depth (background is
//                  farthest; ties go to the lower pixel index) and T its triangle.  Among
//                  T's edges that separate the two pixel centres (edge function >= 0 at f,
//                  < 0 at the other pixel o) take the first one crossed from f, at
//                  t = E(f) / (E(f) - E(o)) in [0, 1], and walk on across interior edges
//                  to the first silhouette edge the segment crosses (no neighbouring
//                  triangle across it, or the neighbour faces the other way).  The pixel
//                  whose half of the segment holds the crossing is blended with the other:
//                    t < 1/2: f += a (1/2 - t) (C_o - C_f);  else o += a (t - 1/2) (E - C_o)
//                  where E is the silhouette triangle's colour at the crossing point,
//                  with the axis weight a = ey^2 / (ex^2 + ey^2) for horizontal pairs and
//                  ex^2 / (ex^2 + ey^2) for vertical pairs, (ex, ey) the edge's direction on
//                  screen.  Each axis alone keeps the image continuous as an edge sweeps
//                  over a pixel centre, and the weights sum to one, so the sum does too.
//                  Blends use the colours before antialiasing and add up.
//
// Passes.  Forward: k_diff_interp (barycentrics, colour, image = colour, and the cheap part
// of every pair's decision; pairs that need the silhouette walk go to a queue),
// k_diff_march (one lane per queued pair: the walk, the stored decision, per-tile flags),
// k_diff_blend (only tiles with an active pair: the antialiasing added to the image).
// Backward: k_diff_backward<INTERIOR> (every pixel: its own triangle's interpolation under
// dL/dimage) and k_diff_backward<EDGES> (flagged tiles only: what the antialiasing adds),
// both reading the stored decisions; gradients reach the vertex attributes through b_i,
// the clip coordinates through D_i (barycentrics) and through t (the silhouette crossing),
// accumulated per tile in LDS and flushed with float atomics.  A backward whose buffers are
// not those of the last forward on the context takes decisions again itself
// (k_diff_backward<ALL>).  The walk is queued because few pairs need it (about 1% on a
// 41k-vertex sphere at 512 x 512) and a wave that holds one waits on its chain of dependent
// loads with all 64 lanes.
//
// Precision.  D_i is a cancellation-prone quantity: for a small or grazing triangle the
// three clip vectors are nearly parallel, and fp32 loses up to 1e-4 of D.  The values
// (D, b, t and every derivative) are therefore evaluated in fp64 from the fp32 inputs; the
// kernels are bound by memory and atomics, not by the fp64 rate.  The discrete choices of
// the antialiasing pass use fp32, so they are reproducible by a CPU restatement.
#include <cstring>
#include <vector>

#include "swr_common.h"
#include "swr_internal.h"

namespace swr {

struct DiffArgs {
    int V, W, H, N, T;
    int culled;    // every visible triangle faces the same way (the visibility draw culled)
    int tiles_x, tiles_per_view;  // 16x16-pixel tiles, one per workgroup
    float sx, sy;  // W/2, H/2: NDC -> window scale (viewport (0, 0, W, H))
    const float4 *clip;
    const int32_t *tris;
    const int32_t *adj;
    const float *attr;
    const uint32_t *ids;
    const float *depth;
    float *bary;
    float *color;
    float *image;
    const float *gimage;  // dL/dimage, or the target image T when l2_target
    int l2_target;        // dL/dimage = image - T
    float *gclip;
    float *gattr;  // per-view partial attribute gradients [V][N][C], summed afterwards
    float *gattr_out;
    // antialiasing decisions of the forward pass, two pairs per pixel (to the right, below):
    // flags (PairLds::flags | edge << 4) and {crossing t, triangle}; reused by the backward
    uint8_t *pflag;
    int2 *pdata;
    uint32_t *tflag;   // per (view, tile): bit 0 some pair of the tile's pixels is active,
                       // bit 1 some pair to the right of its last column, bit 2 some pair
                       // below its last row
    // per (view, tile) b: the pairs that need the silhouette walk (pair index 2 * pixel +
    // dir) at queue[b * 512 ..) with qcnt[b] entries, and the pixels with an active pair (the
    // blend pass lists them) at elist[b * 256 ..) with ecnt[b] entries
    uint32_t *queue, *qcnt;
    uint32_t *elist, *ecnt;
    int pairs;     // 0: evaluate, 1: evaluate and store, 2: load the stored decisions
    // several meshes (null for one): per view, the clip row of global vertex g is
    // vrow[view] + g and visibility id k names triangle vtri[view].x + k - 1 (k <= .y);
    // per scene (S of them), its first vertex, first view and first row
    const long long *vrow;
    const int2 *vtri;
    const int *svert, *sview;
    int S;
};

// dL/d(image channel c of pixel i): the caller's image gradient, or image - target for the
// fused L2 loss (SWR_DIFF_L2_TARGET), formed in the same fp32 operation torch would use.
__device__ inline float grad_image(const DiffArgs &d, size_t i, int C, int c)
{
    const float g = d.gimage[i * C + c];
    return d.l2_target ? d.image[i * C + c] - g : g;
}

// Clip / gradient row of vertex v (global number) as seen by `view`.
__device__ inline size_t crow(const DiffArgs &d, int view, int v)
{
    return d.vrow ? (size_t)(d.vrow[view] + v) : (size_t)view * d.N + v;
}

// Triangle named by visibility id `id` in `view`; -1 for the background (0) or an id past
// the view's triangles.
__device__ inline int id_tri(const DiffArgs &d, int view, uint32_t id)
{
    if (!d.vtri)
        return (id == 0 || id > (uint32_t)d.T) ? -1 : (int)id - 1;
    const int2 t = d.vtri[view];
    return (id == 0 || id > (uint32_t)t.y) ? -1 : t.x + (int)id - 1;
}

// Pixel of this thread: a workgroup of 256 threads covers one 16x16 tile of one view, so a
// wave (16x4 pixels) reads its neighbours, triangles and vertices from a small area.
struct PixelIdx {
    int view, px, py;
    int tile;  // (view, tile) index of the workgroup
    size_t i;
};

// Workgroup b runs on XCD b % 8 (round-robin dispatch).  Renumber so that each XCD works
// through one contiguous run of (view, tile) blocks: the views an XCD is working on at any
// moment have their vertex data and gradient slices in that XCD's L2, where the float
// atomics of the backward pass resolve.
__device__ inline int xcd_order(unsigned b, unsigned n)
{
    const unsigned q = n / 8u, r = n % 8u, x = b % 8u, i = b / 8u;
    return (int)(x * q + (x < r ? x : r) + i);
}

__device__ inline bool pixel_of_thread(const DiffArgs &d, PixelIdx &q)
{
    const int b = xcd_order(blockIdx.x, gridDim.x);
    q.tile = b;
    q.view = b / d.tiles_per_view;
    const int t = b - q.view * d.tiles_per_view;
    const int ty = t / d.tiles_x, tx = t - ty * d.tiles_x;
    q.px = tx * 16 + (int)(threadIdx.x & 15);
    q.py = ty * 16 + (int)(threadIdx.x >> 4);
    if (q.px >= d.W || q.py >= d.H)
        return false;
    q.i = ((size_t)q.view * d.H + q.py) * d.W + q.px;
    return true;
}

// clip (x, y, w) of a vertex, in fp64
struct H3 {
    double x, y, w;
};

__device__ inline H3 cross3(const H3 &a, const H3 &b)
{
    return {a.y * b.w - a.w * b.y, a.w * b.x - a.x * b.w, a.x * b.y - a.y * b.x};
}

__device__ inline double dot3(const H3 &a, const H3 &b) { return a.x * b.x + a.y * b.y + a.w * b.w; }

__device__ inline H3 load_h(const DiffArgs &d, int view, int v)
{
    const float4 c = d.clip[crow(d, view, v)];
    return {(double)c.x, (double)c.y, (double)c.w};
}

__device__ inline H3 pixel_ray(const DiffArgs &d, int px, int py)
{
    const double X = 2.0 * ((double)px + 0.5) / (double)d.W - 1.0;
    const double Y = 1.0 - 2.0 * ((double)py + 0.5) / (double)d.H;
    return {X, Y, 1.0};
}

// Window position of a clip-space vertex (w > 0): fp32 for decisions, fp64 for values.
__device__ inline float2 to_win_f(const DiffArgs &d, int view, int v)
{
    const float4 h = d.clip[crow(d, view, v)];
    return make_float2(d.sx * (h.x / h.w) + d.sx, -d.sy * (h.y / h.w) + d.sy);
}

__device__ inline double2 to_win_d(const DiffArgs &d, const H3 &h)
{
    const double sx = 0.5 * d.W, sy = 0.5 * d.H;
    return make_double2(sx * (h.x / h.w) + sx, -sy * (h.y / h.w) + sy);
}

__device__ inline float edge_fn(float2 j, float2 k, float2 p)
{
    return (k.x - j.x) * (p.y - j.y) - (k.y - j.y) * (p.x - j.x);
}

__device__ inline double edge_d(double2 j, double2 k, double2 p)
{
    return (k.x - j.x) * (p.y - j.y) - (k.y - j.y) * (p.x - j.x);
}

__device__ inline bool w_positive(const DiffArgs &d, int view, int v)
{
    return d.clip[crow(d, view, v)].w > 0.0f;
}

// One neighbour pair as the antialiasing pass sees it.
struct Pair {
    bool active;      // a silhouette edge crosses the segment between the centres
    bool p_blended;   // the blended pixel is p (else q)
    bool horiz;       // horizontal pair
    float tf, s;      // fp32 crossing (decisions), orientation of the triangle
    int tri, vj, vk;  // triangle and the edge's vertices
    int edge;         // the edge's index in the triangle (opposite vertex)
    int front_is_p;
};

// Values of an active pair in fp64: the crossing t and the axis weight a, with their
// derivatives with respect to the window positions of the edge's vertices j and k.
struct PairVal {
    double t, a;
    double2 tj, tk, aj, ak;
};

__device__ inline PairVal pair_val(const DiffArgs &d, int view, const Pair &pr, int px, int py,
                                   int qx, int qy)
{
    PairVal r;
    const double2 j = to_win_d(d, load_h(d, view, pr.vj)), k = to_win_d(d, load_h(d, view, pr.vk));
    const double2 P = make_double2(px + 0.5, py + 0.5), Qp = make_double2(qx + 0.5, qy + 0.5);
    const double2 F = pr.front_is_p ? P : Qp, Q = pr.front_is_p ? Qp : P;
    const double s = pr.s;
    const double ef = s * edge_d(j, k, F), eq = s * edge_d(j, k, Q);
    const double den = ef - eq;
    r.t = ef / den;
    const double dtf = -eq / (den * den), dtq = ef / (den * den);
    // edge(j, k, P) partials: d/dj = (k.y - P.y, P.x - k.x), d/dk = (P.y - j.y, j.x - P.x)
    r.tj = make_double2(s * (dtf * (k.y - F.y) + dtq * (k.y - Q.y)),
                        s * (dtf * (F.x - k.x) + dtq * (Q.x - k.x)));
    r.tk = make_double2(s * (dtf * (F.y - j.y) + dtq * (Q.y - j.y)),
                        s * (dtf * (j.x - F.x) + dtq * (j.x - Q.x)));
    // a = ey^2 / L2 (horizontal pair) or ex^2 / L2, L2 = ex^2 + ey^2, (ex, ey) = k - j
    const double ex = k.x - j.x, ey = k.y - j.y, L2 = ex * ex + ey * ey;
    const double wx = ey * ey / L2;
    const double dex = -2.0 * ex * ey * ey / (L2 * L2), dey = 2.0 * ey * ex * ex / (L2 * L2);
    const double sg = pr.horiz ? 1.0 : -1.0;  // d(1 - wx) = -d(wx)
    r.a = pr.horiz ? wx : 1.0 - wx;
    r.ak = make_double2(sg * dex, sg * dey);
    r.aj = make_double2(-sg * dex, -sg * dey);
    return r;
}

// Window positions of triangle t's vertices (fp32); false when a vertex has w <= 0.
__device__ inline bool tri_win(const DiffArgs &d, int view, int t, float2 *v, int *vid)
{
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        vid[i] = d.tris[3 * t + i];
        if (!w_positive(d, view, vid[i]))
            return false;
        v[i] = to_win_f(d, view, vid[i]);
    }
    return true;
}

constexpr int MARCH_STEPS = 8;

// The pair (p, q): start in the front pixel's triangle and walk along the segment from the
// front centre F to the other centre Q across interior edges (adjacent triangles facing the
// same way) until the segment leaves the mesh through a silhouette edge, ends inside a
// triangle (no crossing: inactive), or MARCH_STEPS triangles were visited (inactive).  The
// exit edge of a triangle is the one the segment crosses first while leaving its half-plane
// (edge value < 0 at Q and falling from F to Q).
// The cheap part of a pair's decision: false when no silhouette can separate p and q (same
// triangle, no front triangle, or two visible triangles of a culled draw sharing a vertex);
// otherwise the front pixel and its triangle, where the walk starts.
__device__ inline bool pair_start(const DiffArgs &d, int view, int px, int py, int qx, int qy,
                                  bool &pf_out, int &cur_out)
{
    const size_t base = (size_t)view * d.W * d.H;
    const size_t ip = base + (size_t)py * d.W + px, iq = base + (size_t)qy * d.W + qx;
    const uint32_t a = d.ids[ip], b = d.ids[iq];
    if (a == b)
        return false;
    bool pf;
    if (a == 0)
        pf = false;
    else if (b == 0)
        pf = true;
    else {
        const float dp = d.depth[ip], dq = d.depth[iq];
        pf = dp < dq || (dp == dq && ip < iq);
    }
    const int cur = id_tri(d, view, pf ? a : b);
    if (cur < 0)
        return false;
    if (d.culled && a != 0 && b != 0) {
        // two visible triangles of a culled draw that share a vertex both face the viewer:
        // no silhouette is taken to separate their pixels
        const int other = id_tri(d, view, pf ? b : a);
        if (other >= 0) {
            const int c0 = d.tris[3 * cur], c1 = d.tris[3 * cur + 1], c2 = d.tris[3 * cur + 2];
            const int o0 = d.tris[3 * other], o1 = d.tris[3 * other + 1],
                      o2 = d.tris[3 * other + 2];
            const bool share = c0 == o0 || c0 == o1 || c0 == o2 || c1 == o0 || c1 == o1 ||
                               c1 == o2 || c2 == o0 || c2 == o1 || c2 == o2;
            if (share)
                return false;
        }
    }
    pf_out = pf;
    cur_out = cur;
    return true;
}

__device__ inline Pair eval_pair(const DiffArgs &d, int view, int px, int py, int qx, int qy)
{
    Pair r;
    r.active = r.p_blended = r.horiz = false;
    r.tf = 0.0f;
    r.s = 1.0f;
    r.tri = r.vj = r.vk = r.edge = r.front_is_p = 0;
    bool pf;
    int cur;
    if (!pair_start(d, view, px, py, qx, qy, pf, cur))
        return r;
    const bool horiz = py == qy;
    const float2 F = pf ? make_float2(px + 0.5f, py + 0.5f) : make_float2(qx + 0.5f, qy + 0.5f);
    const float2 Q = pf ? make_float2(qx + 0.5f, qy + 0.5f) : make_float2(px + 0.5f, py + 0.5f);
    float2 v[3];
    int vid[3];
    if (!tri_win(d, view, cur, v, vid))
        return r;
    float area = edge_fn(v[0], v[1], v[2]);
    if (!(area != 0.0f))
        return r;
    int from = -1;
    for (int step = 0; step < MARCH_STEPS; ++step) {
        const float s = area > 0.0f ? 1.0f : -1.0f;
        float best = 2.0f;
        int be = -1;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            if (e == from)
                continue;
            const int j = (e + 1) % 3, k = (e + 2) % 3;
            const float ef = s * edge_fn(v[j], v[k], F), eq = s * edge_fn(v[j], v[k], Q);
            if (eq < 0.0f && ef > eq) {
                const float tt = ef / (ef - eq);
                if (tt < best) {
                    best = tt;
                    be = e;
                }
            }
        }
        if (be < 0)
            return r;  // the segment ends inside this triangle
        const int n = d.adj[3 * cur + be];
        float2 nv[3];
        int nvid[3];
        float an = 0.0f;
        bool interior = false;
        if (n >= 0 && n < d.T && tri_win(d, view, n, nv, nvid)) {
            an = edge_fn(nv[0], nv[1], nv[2]);
            interior = an != 0.0f && ((an > 0.0f) == (area > 0.0f));
        }
        if (!interior) {
            const int j = (be + 1) % 3, k = (be + 2) % 3;
            if (v[k].x == v[j].x && v[k].y == v[j].y)
                return r;  // a degenerate edge has no direction
            r.active = true;
            r.horiz = horiz;
            r.tf = best;
            r.s = s;
            r.tri = cur;
            r.edge = be;
            r.vj = vid[j];
            r.vk = vid[k];
            r.front_is_p = pf;
            r.p_blended = (best < 0.5f) == pf;
            return r;
        }
        int fe = -1;
#pragma unroll
        for (int e = 0; e < 3; ++e)
            if (d.adj[3 * n + e] == cur)
                fe = e;
        if (fe < 0)
            return r;
        cur = n;
        from = fe;
        area = an;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            v[i] = nv[i];
            vid[i] = nvid[i];
        }
    }
    return r;
}

// Perspective-correct barycentrics of triangle t at the centre of pixel (px, py), in fp64
// (the pixel may lie outside the triangle: the antialiasing extrapolates).  False when the
// three clip vectors are linearly dependent there (S == 0).
struct Bary {
    int v[3];
    H3 h[3];
    H3 r;
    double b[3], S, inv;  // inv = 1 / S: one fp64 division per pixel, products after it
};

__device__ inline bool bary_at_ray(const DiffArgs &d, int view, int t, const H3 &r, Bary &B)
{
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        B.v[m] = d.tris[3 * t + m];
        B.h[m] = load_h(d, view, B.v[m]);
    }
    B.r = r;
    const double D0 = dot3(B.r, cross3(B.h[1], B.h[2])), D1 = dot3(B.r, cross3(B.h[2], B.h[0])),
                 D2 = dot3(B.r, cross3(B.h[0], B.h[1]));
    B.S = D0 + D1 + D2;
    if (B.S == 0.0)
        return false;
    B.inv = 1.0 / B.S;
    B.b[0] = D0 * B.inv;
    B.b[1] = D1 * B.inv;
    B.b[2] = D2 * B.inv;
    return true;
}

__device__ inline bool bary_at(const DiffArgs &d, int view, int t, int px, int py, Bary &B)
{
    return bary_at_ray(d, view, t, pixel_ray(d, px, py), B);
}

// The point where the pair's segment crosses the silhouette edge, as a ray (NDC, 1), and
// its derivative with respect to t.
__device__ inline H3 crossing_ray(const DiffArgs &d, const Pair &pr, double t, int px, int py,
                                  int qx, int qy, H3 &drdt)
{
    const double fx = (pr.front_is_p ? px : qx) + 0.5, fy = (pr.front_is_p ? py : qy) + 0.5;
    const double ox = (pr.front_is_p ? qx : px) + 0.5, oy = (pr.front_is_p ? qy : py) + 0.5;
    const double wx = fx + t * (ox - fx), wy = fy + t * (oy - fy);
    drdt = {2.0 * (ox - fx) / d.W, -2.0 * (oy - fy) / d.H, 0.0};
    return {2.0 * wx / d.W - 1.0, 1.0 - 2.0 * wy / d.H, 1.0};
}

// d(colour at ray r)/d(r) . dr, per channel, for barycentrics B: b_m = D_m / S with
// dD_m/dr = h_j x h_k.
template <int C>
__device__ inline void bary_color_dr(const DiffArgs &d, const Bary &B, const H3 &dr, double *out)
{
    const H3 c0 = cross3(B.h[1], B.h[2]), c1 = cross3(B.h[2], B.h[0]), c2 = cross3(B.h[0], B.h[1]);
    const double e0 = dot3(c0, dr), e1 = dot3(c1, dr), e2 = dot3(c2, dr);
    const double es = e0 + e1 + e2;
    const double db[3] = {(e0 - B.b[0] * es) * B.inv, (e1 - B.b[1] * es) * B.inv,
                          (e2 - B.b[2] * es) * B.inv};
#pragma unroll
    for (int c = 0; c < C; ++c)
        out[c] = db[0] * d.attr[(size_t)B.v[0] * C + c] + db[1] * d.attr[(size_t)B.v[1] * C + c] +
                 db[2] * d.attr[(size_t)B.v[2] * C + c];
}

template <int C>
__device__ inline void bary_color(const DiffArgs &d, const Bary &B, double *col)
{
#pragma unroll
    for (int c = 0; c < C; ++c)
        col[c] = B.b[0] * d.attr[(size_t)B.v[0] * C + c] + B.b[1] * d.attr[(size_t)B.v[1] * C + c] +
                 B.b[2] * d.attr[(size_t)B.v[2] * C + c];
}

// The colour interp_pixel gives pixel (px, py) of `view` before antialiasing (the attributes
// at its barycentrics, 0 on the background, rounded to fp32): read from the colour buffer
// when the caller keeps one, else computed again by the same operations (same bits).
template <int C>
__device__ inline void pixel_color(const DiffArgs &d, int view, size_t i, int px, int py,
                                   double *out)
{
    if (d.color) {
#pragma unroll
        for (int c = 0; c < C; ++c)
            out[c] = d.color[i * C + c];
        return;
    }
    double col[C];
#pragma unroll
    for (int c = 0; c < C; ++c)
        col[c] = 0.0;
    Bary B;
    const int tri = id_tri(d, view, d.ids[i]);
    if (tri >= 0 && bary_at(d, view, tri, px, py, B))
        bary_color<C>(d, B, col);
#pragma unroll
    for (int c = 0; c < C; ++c)
        out[c] = (double)(float)col[c];
}

// Gradient accumulation of one workgroup (one 16x16 tile of one view): an open-addressing
// table in LDS keyed by vertex, flushed to global memory once per distinct vertex.  The
// pixels of a tile touch a few dozen vertices, so this replaces most global float atomics
// (which serialise on shared vertices) by LDS atomics.  A vertex that finds no slot after
// `probes` probes goes straight to global memory.  Slots are never freed, so a vertex that
// got a slot is found again within the same probes.
constexpr int ACC_SLOTS = 256;   // power of two
constexpr int ACC_PROBES = 32;   // one 16x16 tile: a few dozen vertices, the table never fills
// the edge backward's table serves a group of 64 tiles, whose silhouette pixels can touch
// more vertices than it holds: a full table then costs every new vertex all its probes
constexpr int ACC_PROBES_SPARSE = 8;
constexpr int ACC_VALS = 8;      // at most: clip x, y, w, attributes 0..4
// values per slot with C attribute channels: clip x, y, w and the C attributes
template <int C> constexpr int acc_vals() { return 3 + C; }
// Each value is kept in four copies, picked by the pixel's (x, y) parity: the pixels of one
// triangle are neighbours, and four copies cut the lanes of a wave that add into the same
// LDS word (which serialise) about four times.  The flush sums the copies.
constexpr int ACC_COPIES = 4;

__device__ inline int acc_copy() { return (int)((threadIdx.x & 1u) | ((threadIdx.x >> 3) & 2u)); }

struct VertexAcc {
    int *key;
    float *val;
    int probes;
};

__device__ inline int acc_slot(const VertexAcc &A, int v)
{
    if (!A.key)
        return -1;  // no table: straight to global memory
    const unsigned h = ((unsigned)v * 2654435761u) >> 24;  // 8 bits: ACC_SLOTS
#pragma unroll 1
    for (int probe = 0; probe < A.probes; ++probe) {
        const int sl = (int)((h + (unsigned)probe) & (ACC_SLOTS - 1));
        const int k = __atomic_load_n(&A.key[sl], __ATOMIC_RELAXED);
        if (k == v)
            return sl;
        if (k == -1) {
            const int old = atomicCAS(&A.key[sl], -1, v);
            if (old == -1 || old == v)
                return sl;
        }
    }
    return -1;
}

template <int C>
__device__ inline void acc_clip(const DiffArgs &d, const VertexAcc &A, int view, int v, float gx,
                                float gy, float gw)
{
    const int sl = acc_slot(A, v);
    if (sl >= 0) {
        float *x = &A.val[sl * acc_vals<C>() * ACC_COPIES + acc_copy()];
        atomicAdd(x + 0 * ACC_COPIES, gx);
        atomicAdd(x + 1 * ACC_COPIES, gy);
        atomicAdd(x + 2 * ACC_COPIES, gw);
    } else {
        float *g = d.gclip + crow(d, view, v) * 4;
        atomicAdd(g + 0, gx);
        atomicAdd(g + 1, gy);
        atomicAdd(g + 3, gw);
    }
}

// clip and attribute gradients of one vertex with a single table lookup
template <int C>
__device__ inline void acc_vertex(const DiffArgs &d, const VertexAcc &A, int view, int v,
                                  float gx, float gy, float gw, const float *ga)
{
    const int sl = acc_slot(A, v);
    if (sl >= 0) {
        float *x = &A.val[sl * acc_vals<C>() * ACC_COPIES + acc_copy()];
        atomicAdd(x + 0 * ACC_COPIES, gx);
        atomicAdd(x + 1 * ACC_COPIES, gy);
        atomicAdd(x + 2 * ACC_COPIES, gw);
#pragma unroll
        for (int c = 0; c < C; ++c)
            atomicAdd(x + (3 + c) * ACC_COPIES, ga[c]);
    } else {
        float *g = d.gclip + crow(d, view, v) * 4;
        atomicAdd(g + 0, gx);
        atomicAdd(g + 1, gy);
        atomicAdd(g + 3, gw);
#pragma unroll
        for (int c = 0; c < C; ++c)
            atomicAdd(d.gattr + crow(d, view, v) * C + c, ga[c]);
    }
}

// The same with the three slots already found (acc_slot; -1: straight to global memory).
template <int C>
__device__ inline void acc_triangle_at(const DiffArgs &d, const VertexAcc &A, int view,
                                       const int (&v)[3], const int (&sl)[3],
                                       const float (&g)[3][3 + C])
{
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        if (sl[m] >= 0) {
            float *x = &A.val[sl[m] * acc_vals<C>() * ACC_COPIES + acc_copy()];
#pragma unroll
            for (int k = 0; k < 3 + C; ++k)
                atomicAdd(x + k * ACC_COPIES, g[m][k]);
        } else {
            float *gc = d.gclip + crow(d, view, v[m]) * 4;
            atomicAdd(gc + 0, g[m][0]);
            atomicAdd(gc + 1, g[m][1]);
            atomicAdd(gc + 3, g[m][2]);
#pragma unroll
            for (int c = 0; c < C; ++c)
                atomicAdd(d.gattr + crow(d, view, v[m]) * C + c, g[m][3 + c]);
        }
    }
}

// The slots of the three vertices of one pixel's triangle: the first probes of the three
// lookups are issued together (one LDS round trip instead of three in the common case that
// the vertex already has its slot), the rest probe on.
__device__ inline void acc_find3(const VertexAcc &A, const int (&v)[3], int (&sl)[3])
{
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const int h = (int)((((unsigned)v[m] * 2654435761u) >> 24) & (ACC_SLOTS - 1));
        sl[m] = __atomic_load_n(&A.key[h], __ATOMIC_RELAXED) == v[m] ? h : -1;
    }
#pragma unroll
    for (int m = 0; m < 3; ++m)
        if (sl[m] < 0)
            sl[m] = acc_slot(A, v[m]);
}

template <int C>
__device__ inline void acc_triangle(const DiffArgs &d, const VertexAcc &A, int view,
                                   const int (&v)[3], const float (&g)[3][3 + C])
{
    int sl[3];
    acc_find3(A, v, sl);
    acc_triangle_at<C>(d, A, view, v, sl, g);
}

template <int C>
__device__ inline void acc_flush(const DiffArgs &d, const VertexAcc &A, int view)
{
    for (int sl = threadIdx.x; sl < ACC_SLOTS; sl += blockDim.x) {
        const int v = A.key[sl];
        if (v < 0)
            continue;
        float x[ACC_VALS];
#pragma unroll
        for (int k = 0; k < acc_vals<C>(); ++k) {
            const float4 q = *reinterpret_cast<const float4 *>(&A.val[(sl * acc_vals<C>() + k) * ACC_COPIES]);
            x[k] = (q.x + q.y) + (q.z + q.w);
        }
        float *g = d.gclip + crow(d, view, v) * 4;
        if (x[0] != 0.0f) atomicAdd(g + 0, x[0]);
        if (x[1] != 0.0f) atomicAdd(g + 1, x[1]);
        if (x[2] != 0.0f) atomicAdd(g + 3, x[2]);
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (x[3 + c] != 0.0f)
                atomicAdd(d.gattr + crow(d, view, v) * C + c, x[3 + c]);
    }
}

// Backward of col = sum_m b_m A_m given dL/dcol: adds to the attributes and, through
// b_m = D_m / S with D_m = r . (h_j x h_k), to the clip coordinates of the three vertices.
// The gradients bary_backward accumulates, per vertex m: {clip x, y, w, attribute 0..C).
template <int C>
__device__ inline void bary_grads(const DiffArgs &d, const Bary &B, const double *gcol,
                                  float (&out)[3][3 + C])
{
    double u[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        u[m] = 0.0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            u[m] += gcol[c] * d.attr[(size_t)B.v[m] * C + c];
            out[m][3 + c] = (float)(B.b[m] * gcol[c]);
        }
    }
    const double ub = u[0] * B.b[0] + u[1] * B.b[1] + u[2] * B.b[2];
    H3 g[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const double gD = (u[m] - ub) * B.inv;
        const int j = (m + 1) % 3, k = (m + 2) % 3;
        // dD_m/dh_j = h_k x r, dD_m/dh_k = r x h_j
        const H3 dj = cross3(B.h[k], B.r), dk = cross3(B.r, B.h[j]);
        g[j].x += gD * dj.x;
        g[j].y += gD * dj.y;
        g[j].w += gD * dj.w;
        g[k].x += gD * dk.x;
        g[k].y += gD * dk.y;
        g[k].w += gD * dk.w;
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        out[m][0] = (float)g[m].x;
        out[m][1] = (float)g[m].y;
        out[m][2] = (float)g[m].w;
    }
}

template <int C>
__device__ inline void bary_backward(const DiffArgs &d, const VertexAcc &A, int view,
                                     const Bary &B, const double *gcol)
{
    float g[3][3 + C];
    bary_grads<C>(d, B, gcol, g);
#pragma unroll
    for (int m = 0; m < 3; ++m)
        acc_vertex<C>(d, A, view, B.v[m], g[m][0], g[m][1], g[m][2], &g[m][3]);
}

// Up to two values per thread into the workgroup's own segment `seg`; thread 0 writes the
// segment's length to *cnt.  Every thread of the workgroup calls it.  (A list shared by all
// workgroups needs a global counter, whose atomics serialise.)
__device__ inline void tile_append2(uint32_t *seg, uint32_t *cnt, bool t0, uint32_t v0, bool t1,
                                    uint32_t v1)
{
    __shared__ uint32_t s_n;
    const unsigned lane = threadIdx.x & 63u;
    if (threadIdx.x == 0)
        s_n = 0;
    __syncthreads();
    const unsigned long long m0 = __ballot(t0), m1 = __ballot(t1);
    const uint32_t n = (uint32_t)(__popcll(m0) + __popcll(m1));
    uint32_t base = 0;
    if (lane == 0 && n)
        base = atomicAdd(&s_n, n);
    base = __shfl(base, 0, 64);
    const unsigned long long below = (1ull << lane) - 1ull;
    if (t0)
        seg[base + (uint32_t)__popcll(m0 & below)] = v0;
    if (t1)
        seg[base + (uint32_t)__popcll(m0) + (uint32_t)__popcll(m1 & below)] = v1;
    __syncthreads();
    if (threadIdx.x == 0)
        *cnt = s_n;
}

template <int C>
__device__ inline void interp_pixel(const DiffArgs &d, const PixelIdx &pi, bool *walk)
{
    const size_t i = pi.i;
    const int view = pi.view, px = pi.px, py = pi.py;
    const uint32_t id = d.ids[i];
    double col[C];
#pragma unroll
    for (int c = 0; c < C; ++c)
        col[c] = 0.0;
    float2 bb = make_float2(0.0f, 0.0f);
    Bary B;
    const int tri = id_tri(d, view, id);
    if (tri >= 0 && bary_at(d, view, tri, px, py, B)) {
        bary_color<C>(d, B, col);
        bb = make_float2((float)B.b[0], (float)B.b[1]);
    }
    reinterpret_cast<float2 *>(d.bary)[i] = bb;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (d.color)
            d.color[i * C + c] = (float)col[c];
        d.image[i * C + c] = (float)col[c];  // the blend pass adds the antialiasing
    }
    if (threadIdx.x == 0) {
        d.tflag[pi.tile] = 0u;
        d.ecnt[pi.tile] = 0u;  // the blend pass fills it for tiles with an active pair
    }
    // this pixel's two pairs (to the right, below): decided here unless a silhouette walk
    // is needed
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
        const int qx = px + (dir == 0), qy = py + (dir == 1);
        bool pf;
        int cur;
        walk[dir] = qx < d.W && qy < d.H && pair_start(d, view, px, py, qx, qy, pf, cur);
        d.pflag[2 * i + dir] = 0;
    }
}

template <int C>
__global__ void __launch_bounds__(256) k_diff_interp(DiffArgs d)
{
    PixelIdx pi;
    bool walk[2] = {false, false};
    if (pixel_of_thread(d, pi))
        interp_pixel<C>(d, pi, walk);
    // the pairs that need a silhouette walk, queued for k_diff_march
    const uint32_t k0 = (walk[0] || walk[1]) ? (uint32_t)(2 * pi.i) : 0u;
    tile_append2(d.queue + (size_t)pi.tile * 512, d.qcnt + pi.tile, walk[0], k0, walk[1], k0 + 1u);
}


__device__ inline bool neighbour(const DiffArgs &d, int px, int py, int k, int &qx, int &qy)
{
    qx = px + (k == 0 ? 1 : (k == 1 ? -1 : 0));
    qy = py + (k == 2 ? 1 : (k == 3 ? -1 : 0));
    return qx >= 0 && qx < d.W && qy >= 0 && qy < d.H;
}

// ---- pair decisions shared by a tile -------------------------------------------------------
// Each pair of neighbouring pixels is evaluated once per 16x16 tile: thread (lx, ly) takes the
// pair to its right and the pair below it, the tile's first column and row also the pairs
// across the tile's left and top border.  h[ly][lx + 1] holds the pair (lx, ly)-(lx + 1, ly),
// v[ly + 1][lx] the pair (lx, ly)-(lx, ly + 1), both seen from the left / upper pixel; the
// other pixel reads it with the roles swapped.
struct PairLds {
    int flags;  // 1 active, 2 first pixel blended, 4 first pixel in front, 8 s < 0,
                // bits 4-5 the edge index
    float tf;
    int tri, vj, vk;
};

// Stored form of a pair: the flags byte, and {t, triangle} when active; the edge's vertices
// are re-read from the triangle on load.
__device__ inline void store_pair(const DiffArgs &d, size_t k, const PairLds &x)
{
    d.pflag[k] = (uint8_t)x.flags;
    if (x.flags & 1)
        d.pdata[k] = make_int2(__float_as_int(x.tf), x.tri);
}

__device__ inline PairLds load_pair(const DiffArgs &d, size_t k)
{
    PairLds x;
    const int f = d.pflag[k];
    x.flags = f;
    if (f & 1) {
        const int2 q = d.pdata[k];
        const int e = f >> 4;
        x.tf = __int_as_float(q.x);
        x.tri = q.y;
        x.vj = d.tris[3 * x.tri + (e + 1) % 3];
        x.vk = d.tris[3 * x.tri + (e + 2) % 3];
    }
    return x;
}

__device__ inline PairLds pack_pair(const Pair &p)
{
    PairLds x;
    x.flags = (p.active ? 1 | (p.edge << 4) : 0) | (p.p_blended ? 2 : 0) |
              (p.front_is_p ? 4 : 0) | (p.s < 0.0f ? 8 : 0);
    x.tf = p.tf;
    x.tri = p.tri;
    x.vj = p.vj;
    x.vk = p.vk;
    return x;
}

__device__ inline Pair unpack_pair(const PairLds &x, bool swap, bool horiz)
{
    Pair p;
    p.active = x.flags & 1;
    p.p_blended = ((x.flags & 2) != 0) != swap;
    p.front_is_p = ((x.flags & 4) != 0) != swap;
    p.s = (x.flags & 8) ? -1.0f : 1.0f;
    p.tf = x.tf;
    p.tri = x.tri;
    p.vj = x.vj;
    p.vk = x.vk;
    p.horiz = horiz;
    return p;
}



// A workgroup's share of per-tile segments: TILES_PER_GROUP consecutive tiles, their
// lengths scanned into s_off (s_off[k] = entries before tile k, s_off[64] = all), so the
// workgroup's threads take the entries densely, one each.
constexpr int TILES_PER_GROUP = 64;

__device__ inline void group_segments(const uint32_t *cnt, uint32_t tiles, uint32_t t0,
                                      uint32_t *s_off)
{
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x;
        uint32_t v = t0 + lane < tiles ? cnt[t0 + lane] : 0u;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(v, o, 64);
            if ((int)lane >= o)
                v += u;
        }
        s_off[lane + 1] = v;
        if (lane == 0)
            s_off[0] = 0;
    }
    __syncthreads();
}

// The tile (within the group) of entry j: the largest k with s_off[k] <= j.
__device__ inline uint32_t group_tile(const uint32_t *s_off, uint32_t j)
{
    uint32_t lo = 0, hi = TILES_PER_GROUP;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_off[mid] <= j)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// The queued pairs, one lane each: the silhouette walk, the stored decision, and the tile
// flags of active pairs.  A walk is a chain of dependent loads (adjacency, then the next
// triangle's vertices), so the kernel is latency-bound: 1024 lanes per group of 64 tiles
// take a group's queue in one pass where 256 lanes needed several in turn.
constexpr int MARCH_THREADS = 1024;

__global__ void __launch_bounds__(MARCH_THREADS) k_diff_march(DiffArgs d)
{
    __shared__ uint32_t s_off[TILES_PER_GROUP + 1];
    const uint32_t tiles = (uint32_t)(d.tiles_per_view * d.V), t0 = blockIdx.x * TILES_PER_GROUP;
    group_segments(d.qcnt, tiles, t0, s_off);
    const uint32_t n = s_off[TILES_PER_GROUP];
    const size_t HW = (size_t)d.W * d.H;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        const uint32_t t = group_tile(s_off, j);
        const uint32_t k = d.queue[(size_t)(t0 + t) * 512 + (j - s_off[t])];
        const size_t i = k >> 1;
        const int dir = (int)(k & 1u);
        const int view = (int)(i / HW);
        const size_t rem = i - (size_t)view * HW;
        const int py = (int)(rem / (size_t)d.W), px = (int)(rem - (size_t)py * d.W);
        const Pair pr = eval_pair(d, view, px, py, px + (dir == 0), py + (dir == 1));
        store_pair(d, k, pack_pair(pr));
        if (pr.active) {
            const int tx = px >> 4, ty = py >> 4;
            const uint32_t bits = 1u | (dir == 0 && (px & 15) == 15 ? 2u : 0u) |
                                  (dir == 1 && (py & 15) == 15 ? 4u : 0u);
            atomicOr(&d.tflag[(size_t)view * d.tiles_per_view + ty * d.tiles_x + tx], bits);
        }
    }
}

// Whether (view, tile) b has an active pair: one of its pixels' own pairs, or a pair across
// its left or top border (owned by the neighbouring tile's last column or row).
__device__ inline bool tile_has_pairs(const DiffArgs &d, int b)
{
    const int t = b % d.tiles_per_view;
    const int tx = t % d.tiles_x, ty = t / d.tiles_x;
    uint32_t f = d.tflag[b] & 1u;
    if (tx > 0)
        f |= d.tflag[b - 1] & 2u;
    if (ty > 0)
        f |= d.tflag[b - d.tiles_x] & 4u;
    return f != 0;
}

constexpr int TP_H = 16 * 17, TP_V = 17 * 16;

__device__ inline PairLds eval_packed(const DiffArgs &d, int view, int ax, int ay, int bx, int by)
{
    if (ax < 0 || ay < 0 || bx >= d.W || by >= d.H || ax >= d.W || ay >= d.H) {
        PairLds x;
        x.flags = 0;
        return x;
    }
    return pack_pair(eval_pair(d, view, ax, ay, bx, by));
}

// Fills the tile's pair table; every thread of the workgroup calls it, then a barrier.
// d.pairs: 0 evaluate, 1 evaluate and store each pixel's own two pairs, 2 load them.
template <bool LOAD = false>
__device__ inline void tile_pairs(const DiffArgs &d, const PixelIdx &pi, bool valid, PairLds *h,
                                  PairLds *v)
{
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int px = pi.px, py = pi.py, view = pi.view;
    if (LOAD) {
        PairLds none;
        none.flags = 0;
        const size_t i = ((size_t)view * d.H + py) * d.W + px;  // valid pixels only
        h[ly * 17 + lx + 1] = valid ? load_pair(d, 2 * i) : none;
        v[(ly + 1) * 16 + lx] = valid ? load_pair(d, 2 * i + 1) : none;
        if (lx == 0)
            h[ly * 17] = valid && px > 0 ? load_pair(d, 2 * (i - 1)) : none;
        if (ly == 0)
            v[lx] = valid && py > 0 ? load_pair(d, 2 * (i - d.W) + 1) : none;
        return;
    }
    const PairLds ph = eval_packed(d, view, px, py, px + 1, py);
    const PairLds pv = eval_packed(d, view, px, py, px, py + 1);
    h[ly * 17 + lx + 1] = ph;
    v[(ly + 1) * 16 + lx] = pv;
    if (d.pairs == 1 && valid) {
        const size_t i = ((size_t)view * d.H + py) * d.W + px;
        store_pair(d, 2 * i, ph);
        store_pair(d, 2 * i + 1, pv);
    }
    if (lx == 0)
        h[ly * 17] = eval_packed(d, view, px - 1, py, px, py);
    if (ly == 0)
        v[lx] = eval_packed(d, view, px, py - 1, px, py);
}

// The pair between this thread's pixel and neighbour k (0 right, 1 left, 2 below, 3 above).
__device__ inline Pair tile_pair(const PairLds *h, const PairLds *v, int k);

// Where a pixel's four pairs come from: the workgroup's LDS table, or the stored decisions.
struct TilePairs {
    const PairLds *h, *v;
    __device__ Pair get(int k) const { return tile_pair(h, v, k); }
};
struct StoredPairs {
    PairLds p[4];  // right, left, below, above, each seen from its left / upper pixel
    __device__ StoredPairs(const DiffArgs &d, size_t i, int px, int py)
    {
        PairLds none;
        none.flags = 0;
        p[0] = px + 1 < d.W ? load_pair(d, 2 * i) : none;
        p[1] = px > 0 ? load_pair(d, 2 * (i - 1)) : none;
        p[2] = py + 1 < d.H ? load_pair(d, 2 * i + 1) : none;
        p[3] = py > 0 ? load_pair(d, 2 * (i - d.W) + 1) : none;
    }
    __device__ Pair get(int k) const { return unpack_pair(p[k], k == 1 || k == 3, k < 2); }
};

__device__ inline Pair tile_pair(const PairLds *h, const PairLds *v, int k)
{
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    switch (k) {
    case 0: return unpack_pair(h[ly * 17 + lx + 1], false, true);
    case 1: return unpack_pair(h[ly * 17 + lx], true, true);
    case 2: return unpack_pair(v[(ly + 1) * 16 + lx], false, false);
    default: return unpack_pair(v[ly * 16 + lx], true, false);
    }
}

// The blend a pixel receives from one active pair where it is the blended pixel:
//   front pixel (t < 1/2):  + a (1/2 - t) (C_q - C_p)
//   other pixel (t >= 1/2): + a (t - 1/2) (E - C_p), E = the silhouette triangle's colour
//                           where the segment crosses the edge.  As the edge sweeps onto
//                           this pixel's centre E becomes the colour the pixel then shows,
//                           so the image stays continuous; on the edge, E is a convex mix of
//                           the edge's endpoint colours, so it stays bounded.
template <int C>
__global__ void __launch_bounds__(256) k_diff_blend(DiffArgs d)
{
    __shared__ PairLds s_h[TP_H], s_v[TP_V];
    const size_t HW = (size_t)d.W * d.H;
    PixelIdx pi;
    if (!tile_has_pairs(d, xcd_order(blockIdx.x, gridDim.x)))
        return;  // the image is the colour here (uniform over the workgroup)
    const bool valid = pixel_of_thread(d, pi);
    tile_pairs<true>(d, pi, valid, s_h, s_v);
    __syncthreads();
    // the backward pass's antialiasing work is per pixel with an active pair: list them
    bool any = false;
    for (int k = 0; k < 4 && valid; ++k) {
        int qx, qy;
        any = any || (neighbour(d, pi.px, pi.py, k, qx, qy) && tile_pair(s_h, s_v, k).active);
    }
    const int b = xcd_order(blockIdx.x, gridDim.x);
    tile_append2(d.elist + (size_t)b * 256, d.ecnt + b, any, any ? (uint32_t)pi.i : 0u, false, 0u);
    if (!valid)
        return;
    const size_t i = pi.i;
    const int view = pi.view, px = pi.px, py = pi.py;
    double cp[C], out[C];
#pragma unroll
    for (int c = 0; c < C; ++c)  // the image holds the colour until this thread blends it
        out[c] = cp[c] = d.image[i * C + c];
    for (int k = 0; k < 4; ++k) {
        int qx, qy;
        if (!neighbour(d, px, py, k, qx, qy))
            continue;
        const Pair pr = tile_pair(s_h, s_v, k);
        if (!pr.active || !pr.p_blended)
            continue;
        const PairVal pv = pair_val(d, view, pr, px, py, qx, qy);
        double src[C];
        if (pr.front_is_p) {
            const size_t iq = (size_t)view * HW + (size_t)qy * d.W + qx;
            pixel_color<C>(d, view, iq, qx, qy, src);
        } else {
            Bary B;
            H3 drdt;
            if (!bary_at_ray(d, view, pr.tri, crossing_ray(d, pr, pv.t, px, py, qx, qy, drdt), B))
                continue;
            bary_color<C>(d, B, src);
        }
        const double w = pv.a * (pr.front_is_p ? 0.5 - pv.t : pv.t - 0.5);
#pragma unroll
        for (int c = 0; c < C; ++c)
            out[c] += w * (src[c] - cp[c]);
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
        d.image[i * C + c] = (float)out[c];
}

// d(window x, y) of a clip-space vertex -> d(clip x, y, w), accumulated.
template <int C>
__device__ inline void scatter_window_grad(const DiffArgs &d, const VertexAcc &A, int view, int v,
                                           double gx, double gy)
{
    const H3 h = load_h(d, view, v);
    const double iw = 1.0 / h.w, sx = 0.5 * d.W, sy = 0.5 * d.H;
    acc_clip<C>(d, A, view, v, (float)(gx * sx * iw), (float)(-gy * sy * iw),
             (float)((-gx * sx * h.x + gy * sy * h.y) * iw * iw));
}

// Which part of a pixel's gradient backward_pixel computes: everything, only the
// interpolation of its own triangle under dL/dimage, or only what the antialiasing adds
// (the interpolation backward is linear in the colour gradient, so the two parts add up).
enum BwdPart { BWD_ALL, BWD_INTERIOR, BWD_EDGES };

template <int C, int PART, class PS>
__device__ inline void backward_pixel(const DiffArgs &d, const VertexAcc &A, const PixelIdx &pi,
                                      const PS &pairs)
{
    const size_t HW = (size_t)d.W * d.H;
    const size_t i = pi.i;
    const int view = pi.view, px = pi.px, py = pi.py;
    double gp[C], gc[C], cp[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        gp[c] = grad_image(d, i, C, c);
        gc[c] = PART == BWD_EDGES ? 0.0 : gp[c];
        cp[c] = 0.0;
    }
    if (PART != BWD_INTERIOR)
        pixel_color<C>(d, view, i, px, py, cp);
    bool touched = PART != BWD_EDGES;

    // antialiasing: colour gradients (gather) and silhouette crossings (scatter)
    for (int k = 0; k < 4 && PART != BWD_INTERIOR; ++k) {
        int qx, qy;
        if (!neighbour(d, px, py, k, qx, qy))
            continue;
        const Pair pr = pairs.get(k);
        if (!pr.active)
            continue;
        const size_t iq = (size_t)view * HW + (size_t)qy * d.W + qx;
        const bool t_low = pr.tf < 0.5f;  // the front pixel is the blended one
        if (!pr.p_blended) {
            // q is blended; its blend reads C_p only when q is the front pixel
            if (t_low) {
                const PairVal pv = pair_val(d, view, pr, px, py, qx, qy);
                const double w = pv.a * (0.5 - pv.t);
#pragma unroll
                for (int c = 0; c < C; ++c)
                    gc[c] += w * grad_image(d, iq, C, c);
                touched = true;
            }
            continue;
        }
        const PairVal pv = pair_val(d, view, pr, px, py, qx, qy);
        const double wt = t_low ? 0.5 - pv.t : pv.t - 0.5;
        const double w = pv.a * wt;
        double src[C];
        double dedt = 0.0;  // d(w gp . E)/dt through the crossing point
        Bary B;
        if (t_low) {
            pixel_color<C>(d, view, iq, qx, qy, src);
        } else {
            H3 drdt;
            if (!bary_at_ray(d, view, pr.tri, crossing_ray(d, pr, pv.t, px, py, qx, qy, drdt), B))
                continue;
            bary_color<C>(d, B, src);
            double gsrc[C], de[C];
#pragma unroll
            for (int c = 0; c < C; ++c)
                gsrc[c] = w * gp[c];
            bary_backward<C>(d, A, view, B, gsrc);
            bary_color_dr<C>(d, B, drdt, de);
#pragma unroll
            for (int c = 0; c < C; ++c)
                dedt += gsrc[c] * de[c];
        }
        double dot = 0.0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            gc[c] -= w * gp[c];
            dot += gp[c] * (src[c] - cp[c]);
        }
        touched = true;
        const double dldt = pv.a * (t_low ? -dot : dot) + dedt, dlda = wt * dot;
        scatter_window_grad<C>(d, A, view, pr.vj, dldt * pv.tj.x + dlda * pv.aj.x,
                            dldt * pv.tj.y + dlda * pv.aj.y);
        scatter_window_grad<C>(d, A, view, pr.vk, dldt * pv.tk.x + dlda * pv.ak.x,
                            dldt * pv.tk.y + dlda * pv.ak.y);
    }

    // barycentric interpolation of the pixel's own triangle
    if (!touched)
        return;
    const uint32_t id = d.ids[i];
    Bary B;
    const int tri = id_tri(d, view, id);
    if (tri < 0 || !bary_at(d, view, tri, px, py, B))
        return;
    bary_backward<C>(d, A, view, B, gc);
}

// The whole backward of one 16x16 tile, for a backward whose buffers the last forward did
// not store decisions for: pair decisions evaluated into LDS, vertex gradients accumulated
// in LDS and flushed once per vertex.
template <int C>
__global__ void __launch_bounds__(256) k_diff_backward(DiffArgs d)
{
    __shared__ int s_key[ACC_SLOTS];
    __shared__ __align__(16) float s_val[ACC_SLOTS * acc_vals<C>() * ACC_COPIES];
    __shared__ PairLds s_h[TP_H], s_v[TP_V];
    for (int sl = threadIdx.x; sl < ACC_SLOTS; sl += blockDim.x)
        s_key[sl] = -1;
    for (int k = threadIdx.x; k < ACC_SLOTS * acc_vals<C>() * ACC_COPIES / 4; k += blockDim.x)
        reinterpret_cast<float4 *>(s_val)[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    PixelIdx pi;
    const bool valid = pixel_of_thread(d, pi);
    tile_pairs<false>(d, pi, valid, s_h, s_v);
    __syncthreads();
    const VertexAcc A{s_key, s_val, ACC_PROBES};
    if (valid)  // every thread reaches the barrier below
        backward_pixel<C, BWD_ALL>(d, A, pi, TilePairs{s_h, s_v});
    __syncthreads();
    acc_flush<C>(d, A, pi.view);
}

// DPP lane shifts inside rows of 16 lanes (one pixel row of the 16x16 tile): lane i reads
// lane i + n; a read past the row end yields `edge`.
template <int N>
__device__ inline int row_next(int v, int edge)
{
    return __builtin_amdgcn_update_dpp(edge, v, 0x100 + N, 0xF, 0xF, false);
}
template <int N>
__device__ inline float row_next(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + N, 0xF, 0xF,
                                                      false));
}
template <int N>
__device__ inline int row_prev(int v, int edge)
{
    return __builtin_amdgcn_update_dpp(edge, v, 0x110 + N, 0xF, 0xF, false);
}

// Every pixel's interpolation under dL/dimage.  Neighbouring pixels mostly show the same
// triangle (82% of horizontal pairs on the 41k-vertex sphere), so before the LDS adds the
// lanes of a pixel row merge their 18 gradient values pairwise, then in fours and eights
// (a lane keeps its values when its partner shows another triangle): about half the LDS
// atomics, which serialise on the shared vertices, go away.
template <int C>
__global__ void __launch_bounds__(256) k_diff_bwd_interior(DiffArgs d)
{
    constexpr int NV = 3 + C;
    __shared__ int s_key[ACC_SLOTS];
    __shared__ __align__(16) float s_val[ACC_SLOTS * acc_vals<C>() * ACC_COPIES];
    for (int sl = threadIdx.x; sl < ACC_SLOTS; sl += blockDim.x)
        s_key[sl] = -1;
    for (int k = threadIdx.x; k < ACC_SLOTS * acc_vals<C>() * ACC_COPIES / 4; k += blockDim.x)
        reinterpret_cast<float4 *>(s_val)[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    PixelIdx pi;
    const bool valid = pixel_of_thread(d, pi);
    int tri = -1, vid[3] = {0, 0, 0};
    float g[3][NV];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            g[m][k] = 0.0f;
    if (valid) {
        const uint32_t id = d.ids[pi.i];
        Bary B;
        const int t = id_tri(d, pi.view, id);
        if (t >= 0 && bary_at(d, pi.view, t, pi.px, pi.py, B)) {
            double gc[C];
#pragma unroll
            for (int c = 0; c < C; ++c)
                gc[c] = grad_image(d, pi.i, C, c);
            bary_grads<C>(d, B, gc, g);
            tri = t;
#pragma unroll
            for (int m = 0; m < 3; ++m)
                vid[m] = B.v[m];
        }
    }
    // merge along the row: at stride s, lane i (i % 2s == 0) takes lane i + s's sums when
    // both show the same triangle and lane i + s still holds its own (it merged nothing away)
    const int lx = (int)(threadIdx.x & 15u);
    bool keep = tri >= 0;
#define SWR_MERGE(S)                                                                            \
    {                                                                                          \
        const int t_n = row_next<S>(keep ? tri : -2, -3);                                      \
        const int t_p = row_prev<S>(keep ? tri : -2, -3);                                      \
        const bool take = keep && (lx % (2 * S)) == 0 && t_n == tri;                           \
        const bool give = keep && (lx % (2 * S)) == S && t_p == tri;                           \
        _Pragma("unroll") for (int m = 0; m < 3; ++m)                                          \
            _Pragma("unroll") for (int k = 0; k < NV; ++k)                                     \
        {                                                                                      \
            const float o = row_next<S>(g[m][k]);                                              \
            g[m][k] = take ? g[m][k] + o : g[m][k];                                            \
        }                                                                                      \
        keep = keep && !give;                                                                  \
    }
    SWR_MERGE(1)
    SWR_MERGE(2)
    SWR_MERGE(4)
#undef SWR_MERGE
    __syncthreads();
    const VertexAcc A{s_key, s_val, ACC_PROBES};
    if (keep)
        acc_triangle<C>(d, A, pi.view, vid, g);
    __syncthreads();
    acc_flush<C>(d, A, pi.view);
}

// What the antialiasing adds to the gradients, for the pixels the blend pass listed (one
// lane each; the decisions come from the forward's store).  Vertex gradients are summed in
// an LDS table keyed by vertex when the group's tiles lie in one view (silhouette pixels of
// a ridge share its vertices), else they go straight to global memory.
template <int C>
__global__ void __launch_bounds__(256) k_diff_bwd_edges(DiffArgs d)
{
    __shared__ uint32_t s_off[TILES_PER_GROUP + 1];
    __shared__ int s_key[ACC_SLOTS];
    __shared__ __align__(16) float s_val[ACC_SLOTS * acc_vals<C>() * ACC_COPIES];
    const uint32_t tiles = (uint32_t)(d.tiles_per_view * d.V), t0 = blockIdx.x * TILES_PER_GROUP;
    group_segments(d.ecnt, tiles, t0, s_off);
    const uint32_t n = s_off[TILES_PER_GROUP];
    if (n == 0)
        return;  // n is the same for every thread of the group
    const uint32_t t_last = (t0 + TILES_PER_GROUP < tiles ? t0 + TILES_PER_GROUP : tiles) - 1;
    const int view0 = (int)(t0 / (uint32_t)d.tiles_per_view);
    const bool one_view = (uint32_t)view0 == t_last / (uint32_t)d.tiles_per_view;
    if (one_view) {
        for (int sl = threadIdx.x; sl < ACC_SLOTS; sl += blockDim.x)
            s_key[sl] = -1;
        for (int k = threadIdx.x; k < ACC_SLOTS * acc_vals<C>() * ACC_COPIES / 4; k += blockDim.x)
            reinterpret_cast<float4 *>(s_val)[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __syncthreads();
    }
    const size_t HW = (size_t)d.W * d.H;
    const VertexAcc A = one_view ? VertexAcc{s_key, s_val, ACC_PROBES_SPARSE}
                                   : VertexAcc{nullptr, nullptr, 0};
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        const uint32_t t = group_tile(s_off, j);
        PixelIdx pi;
        pi.i = d.elist[(size_t)(t0 + t) * 256 + (j - s_off[t])];
        pi.view = (int)(pi.i / HW);
        const size_t rem = pi.i - (size_t)pi.view * HW;
        pi.py = (int)(rem / (size_t)d.W);
        pi.px = (int)(rem - (size_t)pi.py * d.W);
        pi.tile = 0;
        backward_pixel<C, BWD_EDGES>(d, A, pi, StoredPairs(d, pi.i, pi.px, pi.py));
    }
    if (one_view) {
        __syncthreads();
        acc_flush<C>(d, A, view0);
    }
}

// grad_attributes += sum over the views that see the vertex (in view order) of the
// partials; the partials are cleared for the next call (they start at zero when the buffer
// is created).
__global__ void __launch_bounds__(256) k_diff_reduce_attr(DiffArgs d, int C)
{
    const size_t n = (size_t)d.N * C;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        const int g = (int)(k / (size_t)C), c = (int)(k - (size_t)g * C);
        int v0 = 0, v1 = d.V;
        if (d.svert) {
            // the scene owning vertex g: the last s with svert[s] <= g (svert[S] = N > g)
            int lo = 0, hi = d.S - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (d.svert[mid] <= g)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            v0 = d.sview[lo];
            v1 = d.sview[lo + 1];
        }
        float s = 0.0f;
        for (int v = v0; v < v1; ++v) {
            float *p = d.gattr + crow(d, v, g) * C + c;
            s += *p;
            *p = 0.0f;
        }
        d.gattr_out[k] += s;
    }
}

// The backward's interior and antialiasing kernels add into the same gradients (commutative
// atomics) and read nothing the other writes, so with a second stream they run side by side:
// the interior kernel is bound by LDS atomics, the edge kernel by its arithmetic.  Below
// SIDE_MIN_TILES tiles (a few 512x512 views) the fork and join cost more than the overlap.
constexpr uint32_t SIDE_MIN_TILES = 8192;
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

template <int C>
static hipError_t launch_diff(const DiffArgs &d, int which, hipStream_t s, const SideStream &side)
{
    const uint32_t blocks = (uint32_t)((size_t)d.tiles_per_view * d.V);
    if (!blocks)
        return hipSuccess;
    const uint32_t groups = (blocks + TILES_PER_GROUP - 1) / TILES_PER_GROUP;
    if (which == 0) {
        hipLaunchKernelGGL(k_diff_interp<C>, dim3(blocks), dim3(256), 0, s, d);
        hipLaunchKernelGGL(k_diff_march, dim3(groups), dim3(MARCH_THREADS), 0, s, d);
        hipLaunchKernelGGL(k_diff_blend<C>, dim3(blocks), dim3(256), 0, s, d);
    } else {
        if (d.pairs == 2) {
            hipStream_t se = s;
            const bool fork = side.s && blocks >= SIDE_MIN_TILES;
            if (fork) {
                hipError_t e = hipEventRecord(side.fork, s);
                if (e == hipSuccess)
                    e = hipStreamWaitEvent(side.s, side.fork, 0);
                if (e != hipSuccess)
                    return e;
                se = side.s;
            }
            hipLaunchKernelGGL(k_diff_bwd_edges<C>, dim3(groups), dim3(256), 0, se, d);
            hipLaunchKernelGGL(k_diff_bwd_interior<C>, dim3(blocks), dim3(256), 0, s, d);
            if (fork) {
                hipError_t e = hipEventRecord(side.join, side.s);
                if (e == hipSuccess)
                    e = hipStreamWaitEvent(s, side.join, 0);
                if (e != hipSuccess)
                    return e;
            }
        } else {
            hipLaunchKernelGGL(k_diff_backward<C>, dim3(blocks), dim3(256), 0, s, d);
        }
        const size_t n = (size_t)d.N * C;
        const uint32_t rb = (uint32_t)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
        if (rb)
            hipLaunchKernelGGL(k_diff_reduce_attr, dim3(rb), dim3(256), 0, s, d, C);
    }
    return hipGetLastError();
}

static int diff_run(swr_context *c, const swr_diff_desc *dd, const float *gimage, float *gclip,
                    float *gattr, int which)
{
    if (!c || !dd)
        return SWR_ERR_INVALID_ARG;
    const swr_diff_desc &q = *dd;
    if (q.views < 0 || q.width < 0 || q.height < 0 || q.vertex_count < 0 || q.triangle_count < 0)
        return context_fail(c, SWR_ERR_INVALID_ARG, "negative size in swr_diff_desc");
    if (q.channels < 1 || q.channels > 4)
        return context_fail(c, SWR_ERR_UNSUPPORTED, "channels must be 1..4");
    const size_t px = (size_t)q.views * q.width * q.height;
    if (px == 0)
        return SWR_OK;
    const void *need[] = {q.clip, q.triangles, q.adjacency, q.attributes, q.ids, q.depth, q.bary};
    for (const void *p : need)
        if (!p || !context_is_device_ptr(p))
            return context_fail(c, SWR_ERR_INVALID_ARG, "swr_diff buffers must be device memory");
    if (q.color && !context_is_device_ptr(q.color))
        return context_fail(c, SWR_ERR_INVALID_ARG, "swr_diff buffers must be device memory");
    if (which == 0 && (!q.image || !context_is_device_ptr(q.image)))
        return context_fail(c, SWR_ERR_INVALID_ARG, "swr_diff_forward needs image");
    if (which == 1 && (!gimage || !gclip || !gattr || !context_is_device_ptr(gimage) ||
                       !context_is_device_ptr(gclip) || !context_is_device_ptr(gattr)))
        return context_fail(c, SWR_ERR_INVALID_ARG, "swr_diff_backward needs device gradients");
    DiffArgs d;
    d.V = q.views;
    d.W = q.width;
    d.H = q.height;
    d.N = q.vertex_count;
    d.T = q.triangle_count;
    d.culled = (q.flags & SWR_DIFF_CULLED) ? 1 : 0;
    d.tiles_x = (q.width + 15) / 16;
    d.tiles_per_view = d.tiles_x * ((q.height + 15) / 16);
    if ((size_t)d.tiles_per_view * d.V > 0x7FFFFFFFull)
        return context_fail(c, SWR_ERR_CAPACITY, "too many pixels for one swr_diff call");
    d.sx = (float)q.width / 2.0f;
    d.sy = (float)q.height / 2.0f;
    d.clip = reinterpret_cast<const float4 *>(q.clip);
    d.tris = q.triangles;
    d.adj = q.adjacency;
    d.attr = q.attributes;
    d.ids = q.ids;
    d.depth = q.depth;
    d.bary = q.bary;
    d.color = q.color;
    d.image = q.image;
    d.gimage = gimage;
    d.l2_target = (q.flags & SWR_DIFF_L2_TARGET) ? 1 : 0;
    if (which == 1 && d.l2_target && (!q.image || !context_is_device_ptr(q.image)))
        return context_fail(c, SWR_ERR_INVALID_ARG, "SWR_DIFF_L2_TARGET needs the forward's image");
    d.gclip = gclip;
    d.gattr = nullptr;
    d.gattr_out = gattr;
    d.vrow = nullptr;
    d.vtri = nullptr;
    d.svert = d.sview = nullptr;
    d.S = 1;
    size_t rows = (size_t)d.V * d.N;  // clip / gradient rows
    uint64_t table_gen = 0;
    if (q.scenes > 1) {
        const int S = q.scenes;
        const int32_t *pv = q.scene_views, *pn = q.scene_vertices, *pt = q.scene_triangles;
        if (!pv || !pn || !pt)
            return context_fail(c, SWR_ERR_INVALID_ARG, "swr_diff scenes > 1 needs the scene tables");
        if (pv[0] != 0 || pn[0] != 0 || pt[0] != 0 || pv[S] != q.views || pn[S] != q.vertex_count ||
            pt[S] != q.triangle_count)
            return context_fail(c, SWR_ERR_INVALID_ARG,
                                "swr_diff scene tables must run from 0 to views / vertex_count / "
                                "triangle_count");
        for (int k = 0; k < S; ++k)
            if (pv[k + 1] < pv[k] || pn[k + 1] < pn[k] || pt[k + 1] < pt[k])
                return context_fail(c, SWR_ERR_INVALID_ARG, "swr_diff scene tables must not decrease");
        // table: vrow [V] i64, vtri [V] int2, svert [S + 1] i32, sview [S + 1] i32
        std::vector<unsigned char> tab((size_t)16 * d.V + (size_t)8 * (S + 1));
        long long *vrow = reinterpret_cast<long long *>(tab.data());
        int2 *vtri = reinterpret_cast<int2 *>(tab.data() + (size_t)8 * d.V);
        int32_t *svert = reinterpret_cast<int32_t *>(tab.data() + (size_t)16 * d.V);
        int32_t *sview = svert + (S + 1);
        long long base = 0;
        for (int k = 0; k < S; ++k) {
            const long long nk = pn[k + 1] - pn[k];
            for (int v = pv[k]; v < pv[k + 1]; ++v) {
                vrow[v] = base + (long long)(v - pv[k]) * nk - pn[k];
                vtri[v] = make_int2(pt[k], pt[k + 1] - pt[k]);
            }
            base += (long long)(pv[k + 1] - pv[k]) * nk;
        }
        std::memcpy(svert, pn, sizeof(int32_t) * (S + 1));
        std::memcpy(sview, pv, sizeof(int32_t) * (S + 1));
        rows = (size_t)base;
        const unsigned char *dev =
            static_cast<const unsigned char *>(context_diff_table(c, tab.data(), tab.size(), &table_gen));
        if (!dev)
            return context_last_code(c);
        d.vrow = reinterpret_cast<const long long *>(dev);
        d.vtri = reinterpret_cast<const int2 *>(dev + (size_t)8 * d.V);
        d.svert = reinterpret_cast<const int *>(dev + (size_t)16 * d.V);
        d.sview = d.svert + (S + 1);
        d.S = S;
    }
    if (which == 1) {
        // per-view partials: views do not contend on one vertex's attribute gradient
        d.gattr = (float *)context_zeroed_scratch(c, 0, rows * q.channels * 4);
        if (!d.gattr)
            return context_last_code(c);
    }
    // The forward stores its pair decisions; a backward over the same buffers and sizes
    // reuses them instead of walking the silhouettes again.
    // layout: flags [2 px] u8, data [2 px] int2, per tile: flags, queue and list lengths
    // u32, queue [tiles][512] u32, list [tiles][256] u32
    const size_t tiles = (size_t)d.tiles_per_view * d.V, off_f = 0,
                 off_d = (off_f + px * 2 + 15) & ~(size_t)15, off_t = off_d + px * 2 * sizeof(int2),
                 off_q = (off_t + tiles * 12 + 15) & ~(size_t)15, off_e = off_q + tiles * 512 * 4;
    if (px * 2 > 0xFFFFFFFFull)
        return context_fail(c, SWR_ERR_CAPACITY, "too many pixels for one swr_diff call");
    uint8_t *pbuf = (uint8_t *)context_zeroed_scratch(c, 1, off_e + tiles * 256 * 4);
    if (!pbuf)
        return context_last_code(c);
    d.qcnt = reinterpret_cast<uint32_t *>(pbuf + off_t) + tiles;
    d.ecnt = reinterpret_cast<uint32_t *>(pbuf + off_t) + 2 * tiles;
    d.pflag = pbuf + off_f;
    d.pdata = reinterpret_cast<int2 *>(pbuf + off_d);
    d.tflag = reinterpret_cast<uint32_t *>(pbuf + off_t);
    d.queue = reinterpret_cast<uint32_t *>(pbuf + off_q);
    d.elist = reinterpret_cast<uint32_t *>(pbuf + off_e);
    const uint64_t tag[9] = {(uint64_t)(uintptr_t)q.clip,      (uint64_t)(uintptr_t)q.triangles,
                             (uint64_t)(uintptr_t)q.adjacency, (uint64_t)(uintptr_t)q.ids,
                             (uint64_t)(uintptr_t)q.depth,     (uint64_t)(uintptr_t)pbuf,
                             ((uint64_t)q.views << 32) | (uint32_t)q.width,
                             ((uint64_t)q.height << 40) | ((uint64_t)(q.flags & SWR_DIFF_CULLED) << 32) |
                                 (uint32_t)q.triangle_count,
                             q.scenes > 1 ? table_gen : 0};
    uint64_t *held = context_diff_tag(c);
    if (which == 0) {
        d.pairs = 1;
        std::memcpy(held, tag, sizeof tag);
    } else {
        d.pairs = std::memcmp(held, tag, sizeof tag) == 0 ? 2 : 0;
    }
    hipStream_t s = context_stream(c);
    SideStream side;
    if (!context_side(c, &side.s, &side.fork, &side.join))
        side = SideStream();
    hipError_t e;
    switch (q.channels) {
    case 1: e = launch_diff<1>(d, which, s, side); break;
    case 2: e = launch_diff<2>(d, which, s, side); break;
    case 3: e = launch_diff<3>(d, which, s, side); break;
    default: e = launch_diff<4>(d, which, s, side); break;
    }
    if (e != hipSuccess)
        return context_fail(c, SWR_ERR_HIP, hipGetErrorString(e));
    return SWR_OK;
}

}  // namespace swr

extern "C" {

int swr_diff_forward(swr_context *c, const swr_diff_desc *d)
{
    return swr::diff_run(c, d, nullptr, nullptr, nullptr, 0);
}

int swr_diff_backward(swr_context *c, const swr_diff_desc *d, const float *grad_image,
                      float *grad_clip, float *grad_attributes)
{
    return swr::diff_run(c, d, grad_image, grad_clip, grad_attributes, 1);
}

}  // extern "C"
