// Wave64 helpers shared by the gfx950 kernels: intra-wave ordering of LDS
// and scratch memory, lane masks, and wave-wide reductions and scans.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gis {

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ordering of scratch memory (LDS or global) written by other lanes of the
// same wave; a workgroup's waves share one CU's L1
__device__ __forceinline__ void wave_sync_mem() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// A wave-uniform value moved to scalar registers. The compiler cannot see
// that a wave's index (threadIdx.x / 64), and what is loaded at it, is the
// same on every lane, and otherwise keeps such values in VGPRs, which cost
// occupancy.
__device__ __forceinline__ uint32_t uni32(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    return ((uint64_t)uni32((uint32_t)(x >> 32)) << 32) | uni32((uint32_t)x);
}
// A wave-uniform read through the constant address space, which becomes a
// scalar load (a read; the kernel writes nothing through it). The scalar
// cache is not kept coherent with vector stores, so the element read must not
// be written while the kernel runs: counts, offsets and id lists that earlier
// kernels or the host wrote.
template <typename T>
__device__ __forceinline__ T sload(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}
// this wave's index in the grid, for kernels with wpb waves per block
__device__ __forceinline__ uint32_t wave_index(uint32_t wpb) {
    return uni32(blockIdx.x * wpb + (threadIdx.x >> 6));
}

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t y = __shfl_xor(x, d, 64);
        x = x > y ? x : y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
    uint32_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x - v;
}

// Inclusive prefix sum over the wave from DPP row shifts (1, 2, 4, 8 within
// each row of 16 lanes) and the row broadcasts of lanes 15 and 31 (GFX9
// DPP), without the LDS round trips of __shfl. Every lane must be active.
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// wave total of x (the last lane's inclusive prefix)
__device__ __forceinline__ uint32_t wave_sum_dpp(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan_dpp(x), 63);
}

// wave maximum of x by the same DPP steps (lanes a step does not reach read
// 0, which no unsigned value is below). Every lane must be active.
__device__ __forceinline__ uint32_t wave_max_dpp(uint32_t x) {
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false));
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// lanes for which p holds
__device__ __forceinline__ uint32_t wave_count(bool p) { return (uint32_t)__popcll(__ballot(p)); }

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

}  // namespace gis
