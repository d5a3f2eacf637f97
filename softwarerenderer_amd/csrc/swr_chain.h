// swr_chain.h — the result of a chain of fp32 additions of one constant, without running it.
//
// The reference steps every interpolant along a span by repeated addition
// (PixelData::stepX, PixelData.h:104-118; PixelShaderBase::drawSpan, PixelShaderBase.h:96-112):
// the value at pixel x of a span starting at xl is init followed by (x - xl) additions of the
// plane's a, each rounded to fp32.  Replaying that chain costs one addition per pixel of the
// span before x.  chain_add gives the same bits in a number of steps that grows with the
// number of binades the chain crosses (logarithmic in its length), not with its length.
//
// Why a run of additions inside one binade is one multiplication.  Let v lie in the binade
// [2^e, 2^(e+1)) (or its negative), so v = m*u with u = 2^(e-23) and m an integer, and let
// a = q*u (q real, exact in double).  While the exact sum v + a stays in the binade, fp32
// spacing there is u, so round-to-nearest-even gives fl(v + a) = RNE(m + q)*u.  If q is not
// a half-integer, RNE(m + q) = m + RNE(q) for every m: each addition adds d = RNE(q)*u, and
// k additions add k*d exactly.  If q is a half-integer (a tie), RNE(m + q) = m + RNE(q) when
// m is even, and the result is then even again: after one addition inside the binade (which
// leaves m even whatever it was), every later one adds d.  The closed form is used only while
// every intermediate result stays at least u inside the binade (so each exact sum is within
// u/2 of it and cannot round across an edge); each edge crossing is one explicit addition.
// Non-finite values and magnitudes near the subnormal range take explicit additions too.
#pragma once

#include <stdint.h>

namespace swr {

// 1 / x approximately (the hardware reciprocal on the GPU, a rounded division on the host);
// chain_add corrects what the estimate gets wrong.
__host__ __device__ inline double approx_rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcp(x);
#else
    return 1.0 / x;
#endif
}

// fp32 value after n sequential additions of a: `for (k = 0; k < n; ++k) v = v + a;`
// (round to nearest even, subnormals kept, no contraction).
__host__ __device__ inline float chain_add(float v, float a, int n)
{
    if (n <= 16) {  // short chains: the additions themselves are cheaper
        for (int k = 0; k < n; ++k)
            v = v + a;
        return v;
    }
    int binade = 0x7fffffff;  // exponent of the binade the evenness below refers to
    bool even = false;        // v / u is known to be even (a tie's additions are regular)
    while (n > 0) {
        const float av = v < 0.0f ? -v : v;
        const float aa = a < 0.0f ? -a : a;
        if (n <= 2 || !(av >= 0x1p-100f) || !(av < 0x1p+125f) || !(aa < 0x1p+125f)) {
            v = v + a;
            --n;
            continue;
        }
        const int e = (int)(__builtin_bit_cast(uint32_t, av) >> 23) - 127;
        if (e != binade) {
            binade = e;
            even = false;
        }
        const double lo = __builtin_bit_cast(double, (uint64_t)(e + 1023) << 52);  // 2^e
        const double hi = 2.0 * lo;
        const double u = __builtin_bit_cast(double, (uint64_t)(e - 23 + 1023) << 52);
        const double u_inv = __builtin_bit_cast(double, (uint64_t)(23 - e + 1023) << 52);
        const double q = (double)a * u_inv;  // exact: a power-of-two scaling (a / u)
        const double fq = __builtin_floor(q);
        const bool tie = q - fq == 0.5;
        if (tie && !even) {  // one addition inside the binade makes v / u even
            v = v + a;
            --n;
            const float nv = v < 0.0f ? -v : v;
            even = nv >= 0x1p-100f && (int)(__builtin_bit_cast(uint32_t, nv) >> 23) - 127 == e;
            continue;
        }
        const bool fq_even = __builtin_floor(fq * 0.5) == fq * 0.5;
        const double r = tie ? (fq_even ? fq : fq + 1.0)
                             : __builtin_floor(q + 0.5);  // RNE(q); q - fq != 0.5 here
        const double d = (v < 0.0f ? -r : r) * u;  // change of |v| per addition
        const double m0 = (double)av;
        if (d == 0.0) {
            // |a| <= u/2: every addition rounds back to v, unless v sits on the binade's
            // lower edge and a points to zero (the spacing below it is u/2)
            const bool down = (v < 0.0f) != (a < 0.0f) && a != 0.0f;
            if (!(down && m0 == lo))
                return v;
            v = v + a;
            --n;
            continue;
        }
        // the largest k with every |v| + j*d (1 <= j <= k) inside [lo + u, hi - u]
        double room = d > 0.0 ? (hi - u) - m0 : m0 - (lo + u);
        const double step = d > 0.0 ? d : -d;
        // floor(room / step) is at most 2^23 (room < 2^e, step >= u): estimated from a
        // reciprocal, then made exact by the two loops
        double k = room < 0.0 ? 0.0 : __builtin_floor(room * approx_rcp(step));
        while (k > 0.0 && k * step > room)
            k -= 1.0;
        while ((k + 1.0) * step <= room)
            k += 1.0;
        if (k > (double)n)
            k = (double)n;
        const double nm = m0 + k * d;  // exact: a multiple of u inside the binade
        v = (float)(v < 0.0f ? -nm : nm);
        n -= (int)k;
        if (n > 0) {  // the addition that reaches or crosses the binade's edge
            v = v + a;
            --n;
        }
    }
    return v;
}

}  // namespace swr
