// Device helpers of the syncmer scan shared by the read kernels
// (kernels.hip) and the reference seeding of the index build
// (ref_seed.hip). Positions are int in a read and 64-bit in a genome, so
// the helpers take either.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gis_core.h"

namespace gis {

// any non-ACGT base in [a, b] (b - a < 64)
template <typename I>
__device__ __forceinline__ bool any_n(const uint64_t *nm, I a, I b) {
    const I wa = a >> 6, wb = b >> 6;
    if (wa == wb) {
        const uint64_t mask = (~0ULL >> (63 - (int)(b - a))) << (int)(a & 63);
        return (nm[wa] & mask) != 0;
    }
    const uint64_t m1 = nm[wa] >> (int)(a & 63);
    const uint64_t m2 = nm[wb] & (~0ULL >> (63 - (int)(b & 63)));
    return (m1 | m2) != 0;
}

// bit m of v to bit 2m
__device__ __forceinline__ uint64_t spread32(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFULL;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFULL;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0FULL;
    x = (x | (x << 2)) & 0x3333333333333333ULL;
    x = (x | (x << 1)) & 0x5555555555555555ULL;
    return x;
}

// the 2-bit groups of x in reverse order
__device__ __forceinline__ uint64_t rev_pairs(uint64_t x) {
    x = __builtin_bitreverse64(x);
    return ((x >> 1) & 0x5555555555555555ULL) | ((x & 0x5555555555555555ULL) << 1);
}

// Canonical code of the n-mer (n <= 32) starting at base j: the smaller of
// the forward code (first base most significant, the rolling xs[0] of
// step2/randstrobes.cu:25-27) and the reverse-complement code (first base's
// complement least significant, xs[1]). In the LSB-first packing the window
// is already the reverse complement up to complementing every base.
template <typename I>
__device__ __forceinline__ uint64_t canonical_code(const uint64_t *pk, I j, int n, uint64_t mask) {
    const I q = j >> 5;
    const int sh = (int)(j & 31) * 2;
    uint64_t g = pk[q] >> sh;
    if (sh) g |= pk[q + 1] << (64 - sh);
    g &= mask;
    const uint64_t r = g ^ mask;
    const uint64_t f = rev_pairs(g) >> (64 - 2 * n);
    return f < r ? f : r;
}

// canonical_code for n <= 16: the same value from 32-bit operations. The
// n-mer's 2n bits are the low bits of the 32-bit window at bit 2j of the
// packed bases (v_alignbit over the two words it spans).
__device__ __forceinline__ uint32_t canonical_code32(const uint64_t *pk, int j, int n, uint32_t mask) {
    const uint32_t *w = reinterpret_cast<const uint32_t *>(pk);
    const int b = 2 * j;
    const int wi = b >> 5;
    const uint32_t lo = w[wi], hi = w[wi + 1];
    const uint32_t g = __builtin_amdgcn_alignbit(hi, lo, (uint32_t)(b & 31)) & mask;
    const uint32_t r = g ^ mask;
    uint32_t x = __builtin_bitreverse32(g);
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    const uint32_t f = x >> (32 - 2 * n);
    return f < r ? f : r;
}

// anchor_state (gis_core.h) for a window of W s-mers known at compile time:
// the loads issue together and the minimum, its first index and the count
// of equal values come from straight-line code. The minimum's first index
// and the number of values equal to it are what the running scan of
// anchor_state finds (values before the first minimum are larger).
template <int W, typename I>
__device__ __forceinline__ uint8_t anchor_state_w(const uint64_t *h, I p, bool seg_first) {
    uint64_t v[W];
#pragma unroll
    for (int o = 0; o < W; o++) v[o] = h[p + o];
    uint64_t minv = v[0];
    int first = 0;
#pragma unroll
    for (int o = 1; o < W; o++) {
        const bool lt = v[o] < minv;
        minv = lt ? v[o] : minv;
        first = lt ? o : first;
    }
    int eq = 0;
#pragma unroll
    for (int o = 0; o < W; o++) eq += v[o] == minv ? 1 : 0;
    if (seg_first) return minv == UINT64_MAX ? ST_NONE : (uint8_t)first;
    if (eq == 1 && minv != UINT64_MAX) return (uint8_t)first;
    return ST_UNKNOWN;
}

}  // namespace gis
