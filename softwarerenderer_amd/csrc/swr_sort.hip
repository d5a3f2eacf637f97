// swr_sort.hip — exclusive scans and the stable (tile, primitive) radix sort.
//
// Binning needs, per 8x8 tile, the primitives touching it in draw order.  The emit pass
// writes (tile, primitive) pairs grouped by primitive in draw order; a stable LSD radix
// sort on the tile id (8 bits per pass) regroups them by tile while keeping draw order
// inside each tile.  Ranks inside a block are made stable with wave ballots: the lanes of
// a wave holding the same digit are found with eight ballots, and a lane's rank is the
// number of such lanes below it.
#include "swr_common.h"
#include "swr_internal.h"

namespace swr {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ inline uint32_t sadd(uint32_t a, uint32_t b) { return a + b; }
__device__ inline uint2 sadd(uint2 a, uint2 b) { return make_uint2(a.x + b.x, a.y + b.y); }
template <typename T> __device__ inline T szero();
template <> __device__ inline uint32_t szero<uint32_t>() { return 0u; }
template <> __device__ inline uint2 szero<uint2>() { return make_uint2(0u, 0u); }

// Wave-wide inclusive scan (64 lanes, shuffles).
__device__ inline uint32_t wave_incl(uint32_t v)
{
    const int lane = (int)(threadIdx.x & 63u);
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(v, off, 64);
        if (lane >= off)
            v += o;
    }
    return v;
}
__device__ inline uint2 wave_incl(uint2 v) { return make_uint2(wave_incl(v.x), wave_incl(v.y)); }
__device__ inline uint32_t ssub(uint32_t a, uint32_t b) { return a - b; }
__device__ inline uint2 ssub(uint2 a, uint2 b) { return make_uint2(a.x - b.x, a.y - b.y); }

// Block-wide exclusive scan of one value per thread; returns the block total in *total.
template <typename T>
__device__ T block_exclusive(T v, T *total)
{
    constexpr int NW = SCAN_THREADS / 64;
    __shared__ T s_w[NW];
    const int lane = (int)(threadIdx.x & 63u), w = (int)(threadIdx.x >> 6);
    const T inc = wave_incl(v);
    if (lane == 63)
        s_w[w] = inc;
    __syncthreads();
    T pre = szero<T>(), tot = szero<T>();
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const T x = s_w[k];
        if (k < w)
            pre = sadd(pre, x);
        tot = sadd(tot, x);
    }
    __syncthreads();  // s_w is reused by the next call
    *total = tot;
    return sadd(pre, ssub(inc, v));
}

// SCAN_ITEMS consecutive elements from `base`: 16-byte loads when the run is whole and
// aligned, element by element (zero past n) otherwise.
template <typename T>
__device__ inline void load_items(const T *in, uint64_t base, uint64_t n, T (&v)[SCAN_ITEMS])
{
    const T *p = in + base;
    if (base + SCAN_ITEMS <= n && ((uintptr_t)p & 15u) == 0) {
        constexpr int Q = SCAN_ITEMS * (int)sizeof(T) / 16;
        uint4 q[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k)
            q[k] = reinterpret_cast<const uint4 *>(p)[k];
        __builtin_memcpy(v, q, sizeof v);
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i)
            v[i] = base + i < n ? p[i] : szero<T>();
    }
}

template <typename T>
__device__ inline void store_items(T *out, uint64_t base, uint64_t n, const T (&v)[SCAN_ITEMS])
{
    T *p = out + base;
    if (base + SCAN_ITEMS <= n && ((uintptr_t)p & 15u) == 0) {
        constexpr int Q = SCAN_ITEMS * (int)sizeof(T) / 16;
        uint4 q[Q];
        __builtin_memcpy(q, v, sizeof q);
#pragma unroll
        for (int k = 0; k < Q; ++k)
            reinterpret_cast<uint4 *>(p)[k] = q[k];
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i)
            if (base + i < n)
                p[i] = v[i];
    }
}

// Sum of scan tile b (SCAN_TILE elements) into sums[b].
template <typename T>
__device__ void reduce_tile(const T *in, uint64_t n, T *sums, uint32_t b)
{
    const uint64_t base = (uint64_t)b * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    load_items<T>(in, base, n, v);
    T acc = szero<T>();
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        acc = sadd(acc, v[i]);
    T total;
    block_exclusive<T>(acc, &total);
    if (threadIdx.x == 0)
        sums[b] = total;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const T *in, uint64_t n, T *sums)
{
    reduce_tile<T>(in, n, sums, blockIdx.x);
}

// One block scans the block sums in place (exclusive), SCAN_TILE sums per step.
template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_sums(T *sums, uint32_t nblocks)
{
    T carry = szero<T>();
    for (uint64_t c0 = 0; c0 < nblocks; c0 += SCAN_TILE) {
        const uint64_t base = c0 + (uint64_t)threadIdx.x * SCAN_ITEMS;
        T v[SCAN_ITEMS];
        load_items<T>(sums, base, nblocks, v);
        T acc = szero<T>();
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i)
            acc = sadd(acc, v[i]);
        T total;
        T run = sadd(block_exclusive<T>(acc, &total), carry);
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i) {
            const T x = v[i];
            v[i] = run;
            run = sadd(run, x);
        }
        store_items<T>(sums, base, nblocks, v);
        carry = sadd(carry, total);
    }
}

// The total of the scan, published by the thread that owns the last element: a u32 scan
// writes it to status[word]; the (prims, pairs) scan writes both totals and flags arena
// overflow against the capacities.
__device__ inline void publish_totals(uint32_t t, unsigned long long *status, int word, uint64_t,
                                      uint64_t)
{
    status[word] = t;
}
__device__ inline void publish_totals(uint2 t, unsigned long long *status, int, uint64_t cap_x,
                                      uint64_t cap_y)
{
    status[ST_PRIMS] = t.x;
    status[ST_PAIRS] = t.y;
    unsigned long long of = 0;
    if (t.x > cap_x) of |= 1ull;
    if (t.y > cap_y) of |= 2ull;
    status[ST_OVERFLOW] = of;
}

// Sum of sums[0 .. b) over the block (the carry into block b when no pass has scanned the
// block sums): b / SCAN_THREADS loads per thread, for scans of few blocks.
template <typename T>
__device__ T block_carry(const T *sums, uint32_t b)
{
    T acc = szero<T>();
    for (uint32_t k = threadIdx.x; k < b; k += SCAN_THREADS)
        acc = sadd(acc, sums[k]);
    T total;
    block_exclusive<T>(acc, &total);
    __syncthreads();  // block_exclusive's LDS is used again by the caller
    return total;
}

// Writes the exclusive scan.  CARRY: the block adds up the block sums before it (the
// two-kernel scan); else sums[] holds their exclusive scan.  With `status` set the thread
// that owns the last element publishes the total (publish_totals).  With `list` (the
// (appended fans, live) scan of the geometry pass): element i with a live flag writes i at
// position (its exclusive live count), compacting the live triangles in order, and the
// live total goes to status[ST_LIVE].
template <typename T, bool CARRY>
__device__ void down_tile(const T *in, uint64_t n, const T *sums, T *out,
                          unsigned long long *status, int word, uint64_t cap_x, uint64_t cap_y,
                          uint32_t b, uint32_t *list = nullptr)
{
    const T carry = CARRY ? block_carry<T>(sums, b) : sums[b];
    const uint64_t base = (uint64_t)b * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    T vals[SCAN_ITEMS];
    load_items<T>(in, base, n, vals);
    T acc = szero<T>();
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        acc = sadd(acc, vals[i]);
    T total;
    T run = sadd(block_exclusive<T>(acc, &total), carry);
    T o[SCAN_ITEMS];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        o[i] = run;
        if constexpr (sizeof(T) == 8) {
            if (list) {
                if (base + i < n && vals[i].y)
                    list[run.y] = (uint32_t)(base + i);
                if (base + i == n - 1)
                    status[ST_LIVE] = run.y + vals[i].y;
            }
        }
        if (status && !list && base + i == n - 1)
            publish_totals(sadd(run, vals[i]), status, word, cap_x, cap_y);
        run = sadd(run, vals[i]);
    }
    store_items<T>(out, base, n, o);
}

template <typename T, bool CARRY>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_down(const T *in, uint64_t n,
                                                          const T *sums, T *out,
                                                          unsigned long long *status, int word,
                                                          uint64_t cap_x, uint64_t cap_y,
                                                          uint32_t *list)
{
    down_tile<T, CARRY>(in, n, sums, out, status, word, cap_x, cap_y, blockIdx.x, list);
}

// The per-triangle (appended fans, live) counts and the per-slot (primitives, pairs) counts
// scanned together: blocks [0, nb_a) take the first, the rest the second.
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce2(const uint2 *a, uint64_t na,
                                                             uint2 *sa, uint32_t nba,
                                                             const uint2 *g, uint64_t ng,
                                                             uint2 *sg)
{
    if (blockIdx.x < nba)
        reduce_tile<uint2>(a, na, sa, blockIdx.x);
    else
        reduce_tile<uint2>(g, ng, sg, blockIdx.x - nba);
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_down2(
    const uint2 *a, uint64_t na, const uint2 *sa, uint2 *oa, uint32_t nba, uint32_t *list,
    const uint2 *g, uint64_t ng, const uint2 *sg, uint2 *og, unsigned long long *status,
    uint64_t cap_x, uint64_t cap_y)
{
    if (blockIdx.x < nba)
        down_tile<uint2, true>(a, na, sa, oa, status, 0, 0, 0, blockIdx.x, list);
    else
        down_tile<uint2, true>(g, ng, sg, og, status, 0, cap_x, cap_y, blockIdx.x - nba);
}

// Scans of up to this many blocks skip the pass over the block sums: each block of the
// down pass adds up the sums before it (at most SCAN_CARRY_BLOCKS / SCAN_THREADS loads per
// thread, one round of independent L2 loads).
constexpr uint32_t SCAN_CARRY_BLOCKS = 8192;

template <typename T>
hipError_t scan_impl(const T *in, T *out, uint64_t n, T *tmp, unsigned long long *status,
                     int word, uint64_t cap_x, uint64_t cap_y, hipStream_t s,
                     uint32_t *list = nullptr)
{
    if (!n)
        return hipSuccess;
    uint32_t nb = (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n, tmp);
    if (nb <= SCAN_CARRY_BLOCKS) {
        hipLaunchKernelGGL((k_scan_down<T, true>), dim3(nb), dim3(SCAN_THREADS), 0, s, in, n,
                           tmp, out, status, word, cap_x, cap_y, list);
    } else {
        hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(SCAN_THREADS), 0, s, tmp, nb);
        hipLaunchKernelGGL((k_scan_down<T, false>), dim3(nb), dim3(SCAN_THREADS), 0, s, in, n,
                           tmp, out, status, word, cap_x, cap_y, list);
    }
    return hipGetLastError();
}

uint64_t scan_tmp_elems(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

hipError_t scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *tmp, hipStream_t s)
{
    return scan_impl<uint32_t>(in, out, n, tmp, nullptr, 0, 0, 0, s);
}

hipError_t scan_pairs(const uint2 *in, uint2 *out, uint64_t n, uint2 *tmp,
                      unsigned long long *status, uint64_t cap_prims, uint64_t cap_pairs,
                      hipStream_t s)
{
    return scan_impl<uint2>(in, out, n, tmp, status, 0, cap_prims, cap_pairs, s);
}

hipError_t scan_fans_and_pairs(const uint2 *app, uint2 *app_ex, uint64_t na, uint32_t *live,
                               const uint2 *grp, uint2 *grp_ex, uint64_t ng, void *tmp,
                               unsigned long long *status, uint64_t cap_prims,
                               uint64_t cap_pairs, hipStream_t s)
{
    const uint32_t nba = (uint32_t)((na + SCAN_TILE - 1) / SCAN_TILE);
    const uint32_t nbg = (uint32_t)((ng + SCAN_TILE - 1) / SCAN_TILE);
    // block sums: those of the first scan, then those of the second
    uint2 *sa = static_cast<uint2 *>(tmp);
    uint2 *sg = sa + scan_tmp_elems(na);
    if (!na || !ng || nba > SCAN_CARRY_BLOCKS || nbg > SCAN_CARRY_BLOCKS) {
        hipError_t e = scan_impl<uint2>(app, app_ex, na, sa, status, 0, 0, 0, s, live);
        if (e != hipSuccess)
            return e;
        return scan_pairs(grp, grp_ex, ng, sg, status, cap_prims, cap_pairs, s);
    }
    hipLaunchKernelGGL(k_scan_reduce2, dim3(nba + nbg), dim3(SCAN_THREADS), 0, s, app, na, sa,
                       nba, grp, ng, sg);
    hipLaunchKernelGGL(k_scan_down2, dim3(nba + nbg), dim3(SCAN_THREADS), 0, s, app, na, sa,
                       app_ex, nba, live, grp, ng, sg, grp_ex, status, cap_prims, cap_pairs);
    return hipGetLastError();
}

// ---- radix sort of (tile, primitive) pairs ---------------------------------------------
constexpr int RADIX_THREADS = 256;

// Pass 0 reads the emitted pairs (ST_PAIRS slots: compacted per expansion chunk, or one run
// with rejected pairs keyed >= tiles, which it drops); later passes and the ranges read the
// kept ones (ST_KEPT).
__device__ inline uint64_t pair_count(const unsigned long long *status, int word)
{
    return (status[ST_OVERFLOW] | status[ST_ERROR]) ? 0 : status[word];
}

__device__ inline void chunk_of(const unsigned long long *status, int word, uint32_t G, uint32_t b,
                                uint64_t &beg, uint64_t &end)
{
    uint64_t m = pair_count(status, word);
    uint64_t ch = (m + G - 1) / G;
    beg = (uint64_t)b * ch;
    end = beg + ch;
    if (beg > m) beg = m;
    if (end > m) end = m;
}

// The expansion's output: chunk c of `size` slots holds its kept[c] pairs at its start
// (the tiles a primitive cannot shade are already dropped).  Radix block b of G owns the
// chunks [k0, k1) and reads their pairs in order as one run, RADIX_THREADS chunks at a
// time: s_off holds the group's exclusive prefix of kept counts, a run position maps to its
// chunk by a binary search.
struct Chunks {
    const uint32_t *kept;  // nullptr: the pairs are one contiguous run (chunk_of)
    uint32_t size;
};

__device__ inline void chunks_of(const unsigned long long *status, uint32_t size, uint32_t G,
                                 uint32_t b, uint64_t &k0, uint64_t &k1)
{
    const uint64_t m = pair_count(status, ST_PAIRS);
    const uint64_t n = (m + size - 1) / size, per = (n + G - 1) / G;
    k0 = (uint64_t)b * per;
    k1 = k0 + per;
    if (k0 > n) k0 = n;
    if (k1 > n) k1 = n;
}

// Loads the kept counts of chunks [g0, g0 + RADIX_THREADS) ∩ [., k1) and their exclusive
// prefix into s_off[0 .. RADIX_THREADS]; returns the group's chunk count.
__device__ inline uint32_t load_chunk_group(const Chunks &ch, uint64_t g0, uint64_t k1,
                                            uint32_t *s_off)
{
    const uint32_t nk = (uint32_t)(k1 - g0 < RADIX_THREADS ? k1 - g0 : RADIX_THREADS);
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    __shared__ uint32_t s_w[RADIX_THREADS / 64];
    const uint32_t c = tid < nk ? ch.kept[g0 + tid] : 0u;
    uint32_t inc = c;
#pragma unroll
    for (int sh = 1; sh < 64; sh <<= 1) {
        const uint32_t o = __shfl_up(inc, sh, 64);
        inc += lane >= (uint32_t)sh ? o : 0u;
    }
    if (lane == 63)
        s_w[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < RADIX_THREADS / 64; ++k) {
        pre += k < (int)w ? s_w[k] : 0u;
        tot += s_w[k];
    }
    s_off[tid] = pre + inc - c;
    if (tid == 0)
        s_off[RADIX_THREADS] = tot;
    __syncthreads();
    return nk;
}

// The slot of run position v of the group (v < s_off[nk]).
__device__ inline uint64_t chunk_slot(const Chunks &ch, uint64_t g0, uint32_t nk,
                                      const uint32_t *s_off, uint32_t v)
{
    uint32_t lo = 0, hi = nk;  // last chunk with s_off <= v
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_off[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return (g0 + lo) * (uint64_t)ch.size + (v - s_off[lo]);
}

__global__ void __launch_bounds__(RADIX_THREADS) k_radix_hist(const uint32_t *keys,
                                                            const unsigned long long *status,
                                                            int word, int shift, uint32_t tiles,
                                                            uint32_t *hist, Chunks ch)
{
    __shared__ uint32_t h[256];
    __shared__ uint32_t s_off[RADIX_THREADS + 1];
    const uint32_t G = gridDim.x, b = blockIdx.x;
    h[threadIdx.x] = 0;
    __syncthreads();
    if (ch.kept) {
        uint64_t k0, k1;
        chunks_of(status, ch.size, G, b, k0, k1);
        for (uint64_t g0 = k0; g0 < k1; g0 += RADIX_THREADS) {
            const uint32_t nk = load_chunk_group(ch, g0, k1, s_off);
            const uint32_t tot = s_off[RADIX_THREADS];
            for (uint32_t v = threadIdx.x; v < tot; v += RADIX_THREADS) {
                const uint32_t k = keys[chunk_slot(ch, g0, nk, s_off, v)];
                if (k < tiles)
                    atomicAdd(&h[(k >> shift) & 255u], 1u);
            }
            __syncthreads();  // s_off is reloaded by the next group
        }
    } else {
        uint64_t beg, end;
        chunk_of(status, word, G, b, beg, end);
        for (uint64_t i = beg + threadIdx.x; i < end; i += RADIX_THREADS) {
            uint32_t k = keys[i];
            if (k < tiles)
                atomicAdd(&h[(k >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * G + b] = h[threadIdx.x];
}

// Stable scatter of one pass: RADIX_ITEMS chunks of RADIX_THREADS pairs per round, chunk j
// element tid at base + j * RADIX_THREADS + tid.  A pair's position is the digit's running
// offset plus the count of its digit in earlier chunks, in earlier waves of its chunk, and
// in lower lanes of its wave (eight ballots find the lanes with the same digit).
constexpr int RADIX_ITEMS = 4;

// One round of RADIX_ITEMS * RADIX_THREADS pairs: run positions [base, end), position i at
// slot slot_of(i); run[] carries each digit's next output position between rounds.
template <class SlotOf>
__device__ inline void scatter_round(const uint32_t *keys, const uint32_t *vals, uint32_t *okeys,
                                     uint32_t *ovals, int shift, uint32_t tiles, uint64_t cap,
                                     uint64_t base, uint64_t end, SlotOf slot_of, uint32_t *run,
                                     uint32_t (*wcnt)[RADIX_THREADS / 64][256],
                                     uint32_t (*off)[RADIX_THREADS / 64][256])
{
    constexpr int NW = RADIX_THREADS / 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    {
        uint32_t k[RADIX_ITEMS], v[RADIX_ITEMS], d[RADIX_ITEMS], rank[RADIX_ITEMS];
        bool valid[RADIX_ITEMS];
#pragma unroll
        for (int j = 0; j < RADIX_ITEMS; ++j) {
            const uint64_t i = base + (uint64_t)j * RADIX_THREADS + tid;
            const uint64_t slot = i < end ? slot_of(i) : 0;
            k[j] = i < end ? keys[slot] : 0u;
            valid[j] = i < end && k[j] < tiles;
            v[j] = valid[j] ? vals[slot] : 0u;
            d[j] = (k[j] >> shift) & 255u;
        }
#pragma unroll
        for (int j = 0; j < RADIX_ITEMS; ++j) {
            unsigned long long m = __ballot(valid[j]);
#pragma unroll
            for (int bit = 0; bit < 8; ++bit) {
                const unsigned long long bb = __ballot((d[j] >> bit) & 1u);
                m &= ((d[j] >> bit) & 1u) ? bb : ~bb;
            }
            rank[j] = (uint32_t)__popcll(m & below);
            if (valid[j] && rank[j] == 0)
                wcnt[j][w][d[j]] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        {
            uint32_t acc = run[tid];
#pragma unroll
            for (int j = 0; j < RADIX_ITEMS; ++j)
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    off[j][q][tid] = acc;
                    acc += wcnt[j][q][tid];
                    wcnt[j][q][tid] = 0;
                }
            run[tid] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RADIX_ITEMS; ++j) {
            if (!valid[j])
                continue;
            const uint32_t pos = off[j][w][d[j]] + rank[j];
            if (pos < cap) {
                okeys[pos] = k[j];
                ovals[pos] = v[j];
            }
        }
    }
}

__global__ void __launch_bounds__(RADIX_THREADS) k_radix_scatter(
    const uint32_t *keys, const uint32_t *vals, uint32_t *okeys, uint32_t *ovals,
    const unsigned long long *status, int word, int shift, uint32_t tiles, const uint32_t *hist_ex,
    Chunks ch)
{
    constexpr int NW = RADIX_THREADS / 64;
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RADIX_ITEMS][NW][256];  // per (chunk, wave, digit) counts
    __shared__ uint32_t off[RADIX_ITEMS][NW][256];   // their exclusive prefix from run
    const uint32_t G = gridDim.x, b = blockIdx.x;
    const int tid = threadIdx.x;
    run[tid] = hist_ex[(size_t)tid * G + b];
#pragma unroll
    for (int j = 0; j < RADIX_ITEMS; ++j)
#pragma unroll
        for (int q = 0; q < NW; ++q)
            wcnt[j][q][tid] = 0;
    __syncthreads();
    const uint64_t cap = pair_count(status, ST_PAIRS);
    if (ch.kept) {
        __shared__ uint32_t s_off[RADIX_THREADS + 1];
        uint64_t k0, k1;
        chunks_of(status, ch.size, G, b, k0, k1);
        for (uint64_t g0 = k0; g0 < k1; g0 += RADIX_THREADS) {
            const uint32_t nk = load_chunk_group(ch, g0, k1, s_off);
            const uint32_t tot = s_off[RADIX_THREADS];
            for (uint64_t base = 0; base < tot; base += (uint64_t)RADIX_THREADS * RADIX_ITEMS)
                scatter_round(keys, vals, okeys, ovals, shift, tiles, cap, base, tot,
                              [&](uint64_t i) { return chunk_slot(ch, g0, nk, s_off, (uint32_t)i); },
                              run, wcnt, off);
            __syncthreads();  // s_off is reloaded by the next group
        }
    } else {
        uint64_t beg, end;
        chunk_of(status, word, G, b, beg, end);
        for (uint64_t base = beg; base < end; base += (uint64_t)RADIX_THREADS * RADIX_ITEMS)
            scatter_round(keys, vals, okeys, ovals, shift, tiles, cap, base, end,
                          [](uint64_t i) { return i; }, run, wcnt, off);
    }
}

// Per-tile [begin, end) into the sorted kept pairs; tiles without pairs keep the zeros the
// begin kernel wrote.
__global__ void __launch_bounds__(256) k_ranges(const uint32_t *keys,
                                              const unsigned long long *status, uint32_t *ranges)
{
    uint64_t m = pair_count(status, ST_KEPT);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t k = keys[i];
        if (i == 0 || keys[i - 1] != k)
            ranges[2 * (size_t)k] = (uint32_t)i;
        if (i == m - 1 || keys[i + 1] != k)
            ranges[2 * (size_t)k + 1] = (uint32_t)(i + 1);
    }
}

int radix_passes(uint32_t tiles)
{
    int bits = 0;
    while (bits < 32 && (1ull << bits) < (unsigned long long)tiles)
        ++bits;
    return bits == 0 ? 1 : (bits + 7) / 8;
}

// Rejected pairs carry a key >= tiles; every key below `tiles` sorts within the passes.
uint32_t pair_sentinel(uint32_t) { return 0xFFFFFFFFu; }

hipError_t sort_pairs(SortBuffers &sb, const unsigned long long *status, uint32_t tiles,
                      hipStream_t s)
{
    int passes = radix_passes(tiles);
    uint32_t *ka = sb.keys_a, *va = sb.vals_a, *kb = sb.keys_b, *vb = sb.vals_b;
    const uint32_t nh = 256 * sb.blocks;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        const int word = p == 0 ? ST_PAIRS : ST_KEPT;
        // pass 0 reads the expansion's chunks (when it compacted them), later passes one run
        Chunks ch{p == 0 ? sb.chunk_kept : nullptr, sb.chunk};
        hipLaunchKernelGGL(k_radix_hist, dim3(sb.blocks), dim3(RADIX_THREADS), 0, s, ka, status,
                           word, shift, tiles, sb.hist, ch);
        // pass 0's histogram total is the number of kept pairs (ST_KEPT)
        hipError_t e = scan_impl<uint32_t>(sb.hist, sb.hist_ex, nh, sb.scan_tmp,
                                           p == 0 ? (unsigned long long *)status : nullptr,
                                           ST_KEPT, 0, 0, s);
        if (e != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(sb.blocks), dim3(RADIX_THREADS), 0, s, ka, va, kb,
                           vb, status, word, shift, tiles, sb.hist_ex, ch);
        uint32_t *t = ka;
        ka = kb;
        kb = t;
        t = va;
        va = vb;
        vb = t;
    }
    sb.sorted_keys = ka;
    sb.sorted_vals = va;
    return hipGetLastError();
}

hipError_t launch_ranges(const SortBuffers &sb, const unsigned long long *status,
                         uint32_t tiles, hipStream_t s)
{
    (void)tiles;
    hipLaunchKernelGGL(k_ranges, dim3(1024), dim3(256), 0, s, sb.sorted_keys, status, sb.ranges);
    return hipGetLastError();
}

}  // namespace swr
