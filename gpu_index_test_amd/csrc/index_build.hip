// Device half of strobemer index construction (StrobemerIndex::populate,
// step2/index.cu:143-244): sort the RefRandstrobe records by (hash,
// position, packed), fill the bucket starts with the reference's rule for
// where a bucket begins, and histogram the runs of equal hashes.
//
// The sort is two stable LSD radix sorts of the library (rocPRIM): first by
// the secondary key position << 32 | packed, then by hash. The bucket starts
// and the histogram are plain grid-stride kernels over the sorted table.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "index_build.h"

namespace gis {

namespace {

__global__ void k_split(const uint4 *__restrict__ rs, uint64_t n, uint64_t *__restrict__ sec,
                        uint64_t *__restrict__ hash) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 v = rs[i];
        hash[i] = (uint64_t)v.x | ((uint64_t)v.y << 32);
        sec[i] = ((uint64_t)v.z << 32) | v.w;
    }
}

__global__ void k_pack(uint4 *__restrict__ rs, uint64_t n, const uint64_t *__restrict__ hash,
                       const uint64_t *__restrict__ sec) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t h = hash[i], p = sec[i];
        rs[i] = make_uint4((uint32_t)h, (uint32_t)(h >> 32), (uint32_t)(p >> 32), (uint32_t)p);
    }
}

__global__ void k_fill(uint64_t *__restrict__ a, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

__device__ __forceinline__ uint64_t hash_at(const uint4 *rs, uint64_t i) {
    const uint2 v = reinterpret_cast<const uint2 *>(rs)[2 * i];
    return (uint64_t)v.x | ((uint64_t)v.y << 32);
}

// A bucket start is written only where the hash changes: at change point i
// the buckets after the previous change point's bucket (all of them from 0
// when the previous run is the table's first) up to bucket(hash[i]) start at
// i, so the first hash's run is never inside a bucket (step2/index.cu:187-212).
__global__ void k_starts(const uint4 *__restrict__ rs, uint64_t n, int bits, uint64_t *__restrict__ starts) {
    const uint64_t h0 = n ? hash_at(rs, 0) : 0;
    for (uint64_t i = 1 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t h = hash_at(rs, i), hp = hash_at(rs, i - 1);
        if (h == hp) continue;
        const uint64_t top = h >> (64 - bits);
        const uint64_t lo = hp == h0 ? 0 : (hp >> (64 - bits)) + 1;
        for (uint64_t b = lo; b <= top; b++) starts[b] = i;
    }
}

// runs of equal hashes: hist[c] for c = 1..100, hist[101] for longer runs
__global__ void k_runs(const uint4 *__restrict__ rs, uint64_t n, unsigned long long *__restrict__ hist) {
    __shared__ unsigned long long h[RUN_BINS];
    for (int t = threadIdx.x; t < RUN_BINS; t += blockDim.x) h[t] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = hash_at(rs, i);
        if (i > 0 && hash_at(rs, i - 1) == v) continue;
        uint64_t j = i + 1;
        while (j < n && j - i <= 100 && hash_at(rs, j) == v) j++;
        const uint64_t c = j - i;
        atomicAdd(&h[c > 100 ? 101 : c], 1ull);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < RUN_BINS; t += blockDim.x)
        if (h[t]) atomicAdd(&hist[t], h[t]);
}

unsigned grid_of(uint64_t n) {
    const uint64_t g = (n + 255) / 256;
    return (unsigned)(g < 8192 ? (g ? g : 1) : 8192);
}

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw BuildError{e, what};
}

}  // namespace

void device_sort_records(uint4 *rs, uint64_t n, hipStream_t st) {
    if (n < 2) return;
    uint64_t *buf = nullptr;
    check(hipMalloc(&buf, 4 * n * sizeof(uint64_t)), "hipMalloc of the sort buffers");
    void *tmp = nullptr;
    try {
        uint64_t *sec0 = buf, *sec1 = buf + n, *h0 = buf + 2 * n, *h1 = buf + 3 * n;
        hipLaunchKernelGGL(k_split, dim3(grid_of(n)), dim3(256), 0, st, rs, n, sec0, h0);
        rocprim::double_buffer<uint64_t> keys(sec0, sec1), vals(h0, h1);
        size_t tb1 = 0, tb2 = 0;
        check(rocprim::radix_sort_pairs(nullptr, tb1, keys, vals, (size_t)n, 0, 64, st), "radix sort size");
        rocprim::double_buffer<uint64_t> keys2(h0, h1), vals2(sec0, sec1);
        check(rocprim::radix_sort_pairs(nullptr, tb2, keys2, vals2, (size_t)n, 0, 64, st), "radix sort size");
        const size_t tb = tb1 > tb2 ? tb1 : tb2;
        check(hipMalloc(&tmp, tb ? tb : 1), "hipMalloc of the sort workspace");
        // stable by (position, packed), then stable by hash
        check(rocprim::radix_sort_pairs(tmp, tb1, keys, vals, (size_t)n, 0, 64, st), "radix sort by position");
        rocprim::double_buffer<uint64_t> k2(vals.current(), vals.alternate()), v2(keys.current(), keys.alternate());
        size_t tb3 = tb;
        check(rocprim::radix_sort_pairs(tmp, tb3, k2, v2, (size_t)n, 0, 64, st), "radix sort by hash");
        hipLaunchKernelGGL(k_pack, dim3(grid_of(n)), dim3(256), 0, st, rs, n, k2.current(), v2.current());
        check(hipGetLastError(), "index sort kernels");
        check(hipStreamSynchronize(st), "index sort");
    } catch (...) {
        (void)hipStreamSynchronize(st);
        if (tmp) (void)hipFree(tmp);
        (void)hipFree(buf);
        throw;
    }
    check(hipFree(tmp), "hipFree");
    check(hipFree(buf), "hipFree");
}

namespace {

// a table the lookups can trust: sorted by (hash, position), and every
// record inside [starts[b], starts[b + 1]) carries b in its top hash bits
__global__ void k_validate(const uint4 *__restrict__ rs, uint64_t n, const uint64_t *__restrict__ starts, int bits,
                           unsigned *__restrict__ bad) {
    const uint64_t nb = 1ull << bits;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n || i < nb;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (i >= 1 && i < n) {
            const uint4 a = rs[i - 1], c = rs[i];
            const uint64_t ha = (uint64_t)a.x | ((uint64_t)a.y << 32), hc = (uint64_t)c.x | ((uint64_t)c.y << 32);
            if (ha > hc || (ha == hc && a.z > c.z)) atomicOr(bad, 1u);
        }
        if (i < nb) {
            const uint64_t s = starts[i], e = starts[i + 1];
            if (s < e && e <= n &&
                ((hash_at(rs, s) >> (64 - bits)) != i || (hash_at(rs, e - 1) >> (64 - bits)) != i))
                atomicOr(bad, 2u);
        }
    }
}

}  // namespace

const char *device_validate_index(const uint4 *rs, uint64_t n, const uint64_t *starts, int bits, hipStream_t st) {
    unsigned *bad = nullptr;
    check(hipMalloc(&bad, sizeof(unsigned)), "hipMalloc");
    unsigned h = 0;
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(unsigned), st);
    if (e == hipSuccess) {
        const uint64_t items = std::max<uint64_t>(n, 1ull << bits);
        hipLaunchKernelGGL(k_validate, dim3(grid_of(items)), dim3(256), 0, st, rs, n, starts, bits, bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, bad, sizeof(unsigned), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(bad);
    check(e, "index validation");
    if (h & 1u) return "randstrobes not sorted by (hash, position)";
    if (h & 2u) return "bucket starts do not match the records' hashes";
    return nullptr;
}

void device_finish_index(const uint4 *rs, uint64_t n, int bits, uint64_t *starts, uint64_t *hist_out,
                         hipStream_t st) {
    const uint64_t nb = (1ull << bits) + 1;
    unsigned long long *hist = nullptr;
    check(hipMalloc(&hist, RUN_BINS * sizeof(unsigned long long)), "hipMalloc of the run histogram");
    hipLaunchKernelGGL(k_fill, dim3(grid_of(nb)), dim3(256), 0, st, starts, nb, n);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, reinterpret_cast<uint64_t *>(hist),
                       (uint64_t)RUN_BINS, (uint64_t)0);
    if (n > 1) hipLaunchKernelGGL(k_starts, dim3(grid_of(n)), dim3(256), 0, st, rs, n, bits, starts);
    if (n > 0) hipLaunchKernelGGL(k_runs, dim3(grid_of(n)), dim3(256), 0, st, rs, n, hist);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hist_out, hist, RUN_BINS * sizeof(uint64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(hist);
    check(e, "index starts and run histogram");
}

}  // namespace gis
