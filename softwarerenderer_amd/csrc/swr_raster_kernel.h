// swr_raster_kernel.h — the raster stage's kernel: one wave64 per 8x8 block.
//
// A block is exactly one of the reference's 8x8 blocks (Rasterizer.h:247-250 aligns blocks
// to multiples of 8 in window space).  Screen bins of 8x8, 16x16 or 32x32 pixels hold their
// primitives in draw order; the wave walks its block's primitives in that order and every
// pixel's target values (colour/depth, hash/count, ...) stay in registers for the whole
// walk, written back once, coalesced.
//
// Interpolants are reproduced bit for bit by replaying the reference's incremental chains:
//   Block (PixelShaderBase.h:55-94): row start R_j = init(x+.5, y+.5) with +b repeated j
//     times (PixelData::stepY per row), then +a per pixel along the row (stepX); edge
//     values alike.  The corner classification (Rasterizer.h:270-304) was done at binning
//     and rides in the pair value.
//   Span (PixelShaderBase.h:96-112): value(x) = init(xl+.5, y+.5) then +a per pixel from
//     xl, with [xl, xr) from the flat-part edge walk (Rasterizer.h:370-405); the additions
//     before the tile's first column are fast-forwarded (chain_add, swr_chain.h), so a
//     tile's cost does not grow with its distance from the span's start.
//   w = 1/invw and pvar = pvarTemp * w at the pixel (PixelData.h:84-103).
//
// Work split.  The chains are independent of draw order, only shading is not.  So the
// wave takes GS primitives at a time.  Phase A: lane = (slot, row) walks one 8-pixel row
// of one primitive — 7 adds per parameter produce the row — and the rows go through LDS
// into pixel-per-lane layout.  Phase B: lane = pixel shades the GS primitives in draw
// order.
//
// The kernel is a template over the pixel program, so that programs compiled outside the
// library (include/swr/pixel_program.h) instantiate it the same way the built-in ones do.
#pragma once

#include <atomic>

#include "swr_common.h"
#include "swr_programs.h"

namespace swr {

// ---- the tile kernel ----------------------------------------------------------------------
// Workgroups are handed to the 8 XCDs round-robin (workgroup b runs on XCD b % 8), and
// each XCD has its own L2.  Renumber so that XCD x works through one contiguous run of
// tile blocks (a band of tile rows): primitive records shared by neighbouring tiles are
// then fetched into one L2 instead of up to eight.  A bijection on [0, n).
constexpr unsigned XCDS = 8;

__device__ inline unsigned xcd_block(unsigned b, unsigned n)
{
    const unsigned q = n / XCDS, r = n % XCDS;
    const unsigned x = b % XCDS, i = b / XCDS;
    return x * q + (x < r ? x : r) + i;
}

// Planes phase A chains for a program: all of them, unless the program declares CHAINED
// (it reads only the first CHAINED per fragment and chains the others itself later).
template <class P, int N, class = void>
struct chained_planes {
    static constexpr int value = N;
};
template <class P, int N>
struct chained_planes<P, N, decltype((void)P::CHAINED)> {
    static constexpr int value = P::CHAINED < N ? P::CHAINED : N;
};

// The next work item of a persistent raster workgroup: the first unclaimed item of queue
// home, home + 1, ... (mod RASTER_QUEUES); queues found empty are not asked again.  `items`
// items are dealt into the queues as contiguous runs.  NO_ITEM once every queue is empty.
constexpr uint32_t NO_ITEM = 0xFFFFFFFFu;

// Called by a whole wave (lane 0 claims; the values stay wave-uniform, in scalar registers).
__device__ inline uint32_t take_item(uint32_t *queue, uint32_t items, uint32_t home,
                                     uint32_t &tried, int lane)
{
    const uint32_t q = items / RASTER_QUEUES, r = items % RASTER_QUEUES;
    while (tried < RASTER_QUEUES) {
        const uint32_t x = (home + tried) % RASTER_QUEUES;
        const uint32_t first = x * q + (x < r ? x : r), len = q + (x < r ? 1u : 0u);
        uint32_t i = 0;
        if (lane == 0)
            i = atomicAdd(&queue[x * RASTER_QUEUE_STRIDE], 1u);
        i = __builtin_amdgcn_readlane(i, 0);
        if (i < len)
            return first + i;
        ++tried;
    }
    return NO_ITEM;
}

// The raster kernel.  A workgroup takes either raster_waves bins of 8x8 pixels (one wave per
// bin: the bin is one block and its list carries each primitive's block class), or one bin of
// 16x16 or 32x32 pixels whose 4 or 16 blocks its waves share (wave w takes blocks w, w + W,
// ...): each wave reads the bin's list and keeps, in a per-wave ring in LDS, the primitives
// its block classifies as drawn (Rasterizer.h:270-304 with block_may_cover, or the Span
// rows), in list order.
// MODE 0: 8x8 bins; 1: 16x16 bins whose pair values carry the block classes; 2: larger bins
// whose blocks the kernel classifies.
// SPANS = false: every primitive of the draw is a Block primitive (raster mode Block), so the
// kernel carries no Span code and its registers.
template <template <class> class ProgT, int NZ, int NW, int NA, int NP, int GSX, int MODE,
          bool QUEUE, bool SPANS>
__global__ void __launch_bounds__(512) k_raster(RasterArgs a0)
{
    constexpr bool FILT = MODE != 0;
    using L = Layout<NZ, NW, NA, NP>;
    using Prog = ProgT<L>;
    constexpr int N = L::N;
    constexpr int NS = N > 0 ? N : 1;
    constexpr int NC = chained_planes<Prog, N>::value;  // planes phase A chains
    // primitives per group: the program's default, or the tuned value
    constexpr int GS = GSX > 0 ? GSX : (N <= 8 ? 8 : (N <= 16 ? 4 : (N <= 32 ? 2 : 1)));
    static_assert(GS >= 1 && GS <= 8, "a group is at most 8 primitives (one per row slot)");
    constexpr int MAXW = 8;  // waves per workgroup: blockDim.x / 64, at most 8
    constexpr uint32_t RING = 128;  // per-wave ring of a larger bin's drawn primitives
    constexpr int RW = FILT ? MAXW : 1;
    __shared__ float s_rows[MAXW][GS * 64];
    __shared__ unsigned long long s_cov[MAXW][GS];
    __shared__ uint32_t s_ring[RW][RING];

    if (a0.fold && blockIdx.x == 0 && threadIdx.x == 0)
        fold_status(const_cast<unsigned long long *>(a0.status));
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wpb = (int)(blockDim.x >> 6);
    // 8x8 blocks per bin side: 1 for 8x8 bins (FILT = false), else 2 or 4
    const int S = FILT ? 1 << (a0.bin_shift - 3) : 1;
    if (a0.status[ST_OVERFLOW] | a0.status[ST_ERROR])
        return;
    // phase B: lane = pixel (pi, pj); phase A: lane = (slot, row)
    const int pi = lane & 7, pj = lane >> 3;
    const int slot = lane >> 3, row = lane & 7;
    float *rows = s_rows[wv];
    unsigned char *covb = reinterpret_cast<unsigned char *>(s_cov[wv]);
    uint32_t *ring = s_ring[FILT ? wv : 0];
    const unsigned long long below = (1ull << lane) - 1ull;

    // The work items are the workgroups of the launch: wpb 8x8 bins (one per wave), or one
    // larger bin whose blocks the waves share.  With QUEUE, persistent workgroups take items
    // from RASTER_QUEUES queues, each a contiguous run of items (queues QPX x .. QPX x + QPX - 1
    // a band of bin rows for XCD x; workgroup b runs on XCD b % 8), their own queue first, then
    // the next ones (stealing), and leave once all queues are past their ends.  A workgroup
    // keeps its waves on one item at a time, so they share its pair lists and primitive
    // records in the CU's cache as the waves of a launched workgroup do.  Without QUEUE, each
    // workgroup draws the item of its launch index.  An item's bits do not depend on the
    // workgroup that draws it.
    constexpr uint32_t QPX = RASTER_QUEUES / XCDS;  // queues per XCD
    static_assert(QPX >= 1 && QPX * XCDS == RASTER_QUEUES, "whole queues per XCD");
    __shared__ uint32_t s_item;
    const uint32_t items = S == 1 ? ((uint32_t)a0.tiles + (uint32_t)wpb - 1) / (uint32_t)wpb
                                  : (uint32_t)a0.tiles;
    const uint32_t home = (blockIdx.x % XCDS) * QPX + (blockIdx.x / XCDS) % QPX;
    // wave 0 takes the next item while the waves draw the current one
    uint32_t tried = 0, next = 0;  // (wave 0) queues found empty, the item taken
    if (QUEUE && wv == 0)
        next = take_item(a0.queue, items, home, tried, lane);
    bool taken = false;
    for (;;) {
        uint32_t item;
        if constexpr (!QUEUE) {
            if (taken)
                break;
            taken = true;
            item = a0.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
        } else {
            if (threadIdx.x == 0)
                s_item = next;
            __syncthreads();
            item = __builtin_amdgcn_readfirstlane(s_item);
            if (item == NO_ITEM)
                break;
            __syncthreads();  // every wave has read the item before wave 0 writes the next
            if (wv == 0)
                next = take_item(a0.queue, items, home, tried, lane);
        }
        const int tile =
            __builtin_amdgcn_readfirstlane(S == 1 ? (int)item * wpb + wv : (int)item);
        if (tile >= a0.tiles)
            continue;
        const uint32_t beg = a0.ranges[2 * tile], end = a0.ranges[2 * tile + 1];
        if (beg >= end)
            continue;
        RasterArgs a = a0;
        int lt = tile;  // bin index within its target
        if (a0.nscenes) {
            const int lvl = a0.bin_shift - 3;
            const SceneDev &Sc =
                a0.scenes[scene_of_tile(a0.scenes, a0.nscenes, (uint32_t)tile, lvl)];
            a.width = Sc.width;
            a.height = Sc.height;
            a.pitch = Sc.pitch;
            a.tiles_x = Sc.bins_x[lvl];
            a.sc_minx = Sc.sc_minx;
            a.sc_maxx = Sc.sc_maxx;
            a.t0 = Sc.t0;
            a.t1 = Sc.t1;
            a.t2 = Sc.t2;
            lt = tile - (int)Sc.bin0[lvl];
        }
        const int tx = lt % a.tiles_x, ty = lt / a.tiles_x;
        for (int bi = S == 1 ? 0 : wv; bi < S * S; bi += S == 1 ? 1 : wpb) {
            const int bx = (tx << a.bin_shift) + (bi % S) * TILE;
            const int by = (ty << a.bin_shift) + (bi / S) * TILE;
            if (bx >= a.width || by >= a.height)
                continue;
            const int px = bx + pi, py = by + pj;
            const bool inside = px < a.width && py < a.height;
            const size_t o = (size_t)py * (size_t)a.pitch + (size_t)px;
            const int ry = by + row;
            const float ryf = (float)ry;
            const float xf = (float)bx + 0.5f, yf = (float)by + 0.5f;

            typename Prog::State st;
            if (inside)
                Prog::load(st, a, o);
            bool drew = false;  // a larger bin's block may keep no primitive: nothing to store

            uint32_t pos = beg;            // next list entry to read
            uint32_t head = 0, tail = 0;   // ring of the block's drawn primitives (S > 1)
            for (;;) {
                int cnt;
                if (!FILT) {
                    if (pos >= end)
                        break;
                    cnt = (int)((end - pos) < (uint32_t)GS ? (end - pos) : (uint32_t)GS);
                } else {
                    while (tail - head < (uint32_t)GS && pos < end) {
                        const uint32_t i = pos + (uint32_t)lane;
                        uint32_t pv = 0;
                        bool keep = false;
                        if (i < end && MODE == 1) {
                            pv = a.vals[i];
                            const uint32_t cls = (pv >> (PAIR_CLS4_SHIFT + 2 * bi)) & 3u;
                            keep = cls != 0;
                            pv = (pv & PAIR_CLS4_PRIM_MASK) | cls << 30;
                        } else if (i < end) {
                            pv = a.vals[i] & PAIR_PRIM_MASK;
                            const float *rec = a.prims + (size_t)pv * (size_t)a.stride;
                            const uint32_t flags = __float_as_uint(rec[REC_FLAGS]);
                            if (!SPANS || (flags & FLAG_BLOCK)) {
                                // a block of the primitive's block grid (the scissor clamps the
                                // grid, Rasterizer.h:240-250, not the triangle), then its class
                                const uint32_t *ru = reinterpret_cast<const uint32_t *>(rec);
                                const int kx = bx >> 3, ky = by >> 3;
                                int cls = 0;
                                if (kx >= (int)ru[REC_BLOCKS] && kx <= (int)ru[REC_BLOCKS + 2] &&
                                    ky >= (int)ru[REC_BLOCKS + 1] &&
                                    ky <= (int)ru[REC_BLOCKS + 3]) {
                                    cls = classify_block(rec + REC_EDGE, flags, bx, by);
                                    if (cls == 1 && !block_may_cover(rec + REC_EDGE, bx, by))
                                        cls = 0;
                                }
                                keep = cls != 0;
                                pv |= (uint32_t)cls << 30;
                            } else {
                                keep = span_block_any(rec, flags, bx, by, a.sc_minx, a.sc_maxx);
                            }
                        }
                        const unsigned long long m = __ballot(keep);
                        if (keep)
                            ring[(tail + (uint32_t)__popcll(m & below)) & (RING - 1)] = pv;
                        tail += (uint32_t)__popcll(m);
                        pos += 64;
                    }
                    lds_sync();
                    if (tail == head)
                        break;
                    cnt = (int)((tail - head) < (uint32_t)GS ? (tail - head) : (uint32_t)GS);
                }
                drew = true;
                const bool act = slot < cnt;
                uint32_t pv = 0u;
                if (act)
                    pv = !FILT ? a.vals[pos + slot] : ring[(head + (uint32_t)slot) & (RING - 1)];
                if (!FILT)
                    pos += (uint32_t)cnt;
                else
                    head += (uint32_t)cnt;
                const float *__restrict__ rec =
                    a.prims + (size_t)(pv & PAIR_PRIM_MASK) * (size_t)a.stride;
                const int cls = (int)(pv >> 30);
                const uint32_t flags = act ? __float_as_uint(rec[REC_FLAGS]) : 0u;
                const uint32_t key = act ? __float_as_uint(rec[Prog::KEY_WORD]) : 0u;
                const bool blk = !SPANS || (flags & FLAG_BLOCK) != 0;

                // Span: this row's part and [xl, xr) (Rasterizer.h:370-380, 397-407)
                int xl = 0, xr = 0;
                bool rowv = false;
                if (act && !blk)
                    rowv = span_row_extent(rec, flags, ry, a.sc_minx, a.sc_maxx, xl, xr);
                const float xs = (float)xl + 0.5f, ys = ryf + 0.5f;
                const int pre = bx - xl;  // Span: chain steps from xl to the block's first column

                // coverage of this lane's row (8 bits, bit i = column i)
                uint32_t rowmask = 0;
                if (act) {
                    if (blk) {
                        if (cls == 2) {
                            rowmask = 0xFFu;
                        } else {
                            // the three edges' values at this row's start, then stepped along
                            // the row together: a pixel is in when all three tests pass
                            float v[3], ea[3];
                            bool tie[3];
#pragma unroll
                            for (int e = 0; e < 3; ++e) {
                                ea[e] = rec[REC_EDGE + 3 * e];
                                const float eb = rec[REC_EDGE + 3 * e + 1],
                                            ec = rec[REC_EDGE + 3 * e + 2];
                                tie[e] = (flags >> (1 + e)) & 1u;
                                float r = ea[e] * xf + eb * yf + ec;
                                v[e] = r;
#pragma unroll
                                for (int k = 1; k < 8; ++k) {
                                    r = r + eb;
                                    v[e] = k <= row ? r : v[e];
                                }
                            }
#pragma unroll
                            for (int i = 0; i < 8; ++i) {
                                const bool in = edge_pass(v[0], tie[0]) &&
                                                edge_pass(v[1], tie[1]) &&
                                                edge_pass(v[2], tie[2]);
                                rowmask |= in ? (1u << i) : 0u;
#pragma unroll
                                for (int e = 0; e < 3; ++e)
                                    v[e] = v[e] + ea[e];
                            }
                        }
                    } else if (rowv) {
#pragma unroll
                        for (int i = 0; i < 8; ++i)
                            rowmask |= (xl <= bx + i && bx + i < xr) ? (1u << i) : 0u;
                    }
                    covb[slot * 8 + row] = (unsigned char)rowmask;
                }

                // chained values: phase A row walk, then transpose to pixel-per-lane
                float val[GS][NS];
#pragma unroll
                for (int p = 0; p < NC; ++p) {
                    float out[8];
                    if (act) {
                        const float pa = rec[REC_PARAM + 3 * p], pb = rec[REC_PARAM + 3 * p + 1],
                                    pc = rec[REC_PARAM + 3 * p + 2];
                        if (blk) {
                            // row start: the same chain for every row, kept at this lane's row
                            float r = pa * xf + pb * yf + pc, r0 = r;
#pragma unroll
                            for (int k = 1; k < 8; ++k) {
                                r = r + pb;
                                r0 = k <= row ? r : r0;
                            }
                            out[0] = r0;
#pragma unroll
                            for (int i = 1; i < 8; ++i)
                                out[i] = out[i - 1] + pa;
                        } else {
                            // init at xl, then the pre additions up to the block's first column
                            float v = chain_add(pa * xs + pb * ys + pc, pa, pre);
#pragma unroll
                            for (int i = 0; i < 8; ++i) {
                                out[i] = v;
                                if (bx + i >= xl)
                                    v = v + pa;
                            }
                        }
#pragma unroll
                        for (int i = 0; i < 8; ++i)
                            rows[slot * 64 + row * 8 + i] = out[i];
                    }
                    lds_sync();
#pragma unroll
                    for (int s2 = 0; s2 < GS; ++s2)
                        val[s2][p] = rows[s2 * 64 + pj * 8 + pi];
                    lds_sync();
                }
                if (NC == 0)
                    lds_sync();

                // phase B: shade the group's primitives in draw order
#pragma unroll
                for (int s2 = 0; s2 < GS; ++s2) {
                    if (s2 < cnt) {
                        const unsigned long long m = s_cov[wv][s2];
                        const uint32_t k2 = __builtin_amdgcn_readlane(key, s2 * 8);
                        const uint32_t prim2 =
                            Prog::PLANES ? __builtin_amdgcn_readlane(pv, s2 * 8) & PAIR_PRIM_MASK
                                         : 0u;
                        if (((m >> lane) & 1ull) && inside)
                            Prog::shade(st, a, ChainPixel<L, SPANS>{val[s2], prim2}, k2, px,
                                        py);
                    }
                }
                lds_sync();
            }
            if (inside && drew)
                Prog::template store<ChainPixel<L, SPANS>>(st, a, o, px, py);
        }
    }
}

// Workgroups of `threads` threads of `kernel` that the GPU holds at once (its CUs times the
// kernel's occupancy per CU), kept in `known` per device and workgroup size (each kernel
// keeps its own); 0 when the runtime cannot tell.
inline uint32_t resident_groups(const void *kernel, uint32_t threads,
                                std::atomic<uint32_t> (&known)[16][9])
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16 || threads / 64 > 8)
        return 0;
    std::atomic<uint32_t> &slot = known[dev][threads / 64];
    uint32_t v = slot.load(std::memory_order_relaxed);
    if (v)
        return v;
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, (int)threads, 0) !=
            hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    v = (uint32_t)(cus > 0 ? cus : 0) * (uint32_t)(per > 0 ? per : 0);
    slot.store(v, std::memory_order_relaxed);
    return v;
}

// One raster launch of `blocks` workgroups of `w` waves; with a queue, as many persistent
// workgroups as the GPU holds at once (at most `blocks`).
template <template <class> class ProgT, int NZ, int NW, int NA, int NP, int GSX, int MODE,
          bool SPANS>
inline hipError_t launch_mode(const RasterArgs &a, uint32_t blocks, uint32_t w, hipStream_t s)
{
    if (a.queue) {
        static std::atomic<uint32_t> known[16][9];  // [device][threads / 64]
        const uint32_t res = resident_groups(
            reinterpret_cast<const void *>(
                &k_raster<ProgT, NZ, NW, NA, NP, GSX, MODE, true, SPANS>),
            64 * w, known);
        if (res) {
            hipLaunchKernelGGL((k_raster<ProgT, NZ, NW, NA, NP, GSX, MODE, true, SPANS>),
                               dim3(blocks < res ? blocks : res), dim3(64 * w), 0, s, a);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_raster<ProgT, NZ, NW, NA, NP, GSX, MODE, false, SPANS>),
                       dim3(blocks), dim3(64 * w), 0, s, a);
    return hipGetLastError();
}

// 8x8 bins: `waves` bins per workgroup, one wave each; larger bins: one bin per workgroup
// of a.waves_per_bin waves.
template <template <class> class ProgT, int NZ, int NW, int NA, int NP, int GSX = 0>
inline hipError_t launch(const RasterArgs &a, hipStream_t s, int waves)
{
    uint32_t w = (uint32_t)waves, blocks = (uint32_t)((a.tiles + w - 1) / w);
    if (a.bin_shift > 3) {
        w = (uint32_t)a.waves_per_bin;
        blocks = (uint32_t)a.tiles;
    }
    if (!blocks)
        return hipSuccess;
    const int mode = a.bin_shift > 3 ? (a.cls4 ? 1 : 2) : 0;
    if (a.spans) {
        if (mode == 1)
            return launch_mode<ProgT, NZ, NW, NA, NP, GSX, 1, true>(a, blocks, w, s);
        if (mode == 2)
            return launch_mode<ProgT, NZ, NW, NA, NP, GSX, 2, true>(a, blocks, w, s);
        return launch_mode<ProgT, NZ, NW, NA, NP, GSX, 0, true>(a, blocks, w, s);
    }
    if (mode == 1)
        return launch_mode<ProgT, NZ, NW, NA, NP, GSX, 1, false>(a, blocks, w, s);
    if (mode == 2)
        return launch_mode<ProgT, NZ, NW, NA, NP, GSX, 2, false>(a, blocks, w, s);
    return launch_mode<ProgT, NZ, NW, NA, NP, GSX, 0, false>(a, blocks, w, s);
}

}  // namespace swr
