// psr_sort.hip -- stable LSD radix sort of (key, value) pairs and an
// exclusive scan, for the octree build (keys by octree path, fields) and the
// device ray emission (rays by source).
//
// Radix sort: 4-bit digits.  A pass over bits [shift, shift + 4) runs three
// kernels over tiles of SORT_TILE consecutive elements:
//   k_digit_hist   the tile's count of each digit -> cnt[digit][tile]
//   k_scan_*       exclusive scan of cnt in (digit, tile) order: where each
//                  tile's elements of each digit start in the output
//   k_digit_scatter  every thread counts the digits of its SORT_ITEMS
//                  consecutive elements in registers; a block scan over
//                  (digit, thread) gives each thread its place within the
//                  tile's digit run, so the scatter keeps the input order
//                  among equal digits (stable)
// Passes alternate between the two buffers; the result is left in the
// output buffers (one extra copy when the pass count is odd).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "psr_internal.h"

namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;
constexpr int RADIX = 16;
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

template <typename K>
__device__ __forceinline__ int digit_of(K k, int shift) {
    return (int)((k >> shift) & (K)(RADIX - 1));
}

// exclusive scan of one value per thread over a block of SORT_THREADS or
// SCAN_THREADS threads; *total = the block's sum
template <int NT, typename T>
__device__ __forceinline__ T block_exclusive(T v, T *total) {
    __shared__ T wsum[NT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    T before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; k++) {
        const T s = wsum[k];
        if (k < w) before += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return before + x - v;
}

template <typename K>
__global__ __launch_bounds__(SORT_THREADS) void k_digit_hist(const K *__restrict__ keys, int64_t n, int shift,
                                                             uint32_t *__restrict__ cnt, int64_t ntile) {
    __shared__ uint32_t h[RADIX];
    if (threadIdx.x < RADIX) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t tile = blockIdx.x;
    const int64_t base = tile * SORT_TILE;
    uint32_t c[RADIX];
#pragma unroll
    for (int d = 0; d < RADIX; d++) c[d] = 0;
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; it++) {
        const int64_t i = base + it * SORT_THREADS + threadIdx.x;
        if (i < n) {
            const int d = digit_of(keys[i], shift);
#pragma unroll
            for (int q = 0; q < RADIX; q++) c[q] += (q == d);
        }
    }
#pragma unroll
    for (int d = 0; d < RADIX; d++)
        if (c[d]) atomicAdd(&h[d], c[d]);
    __syncthreads();
    if (threadIdx.x < RADIX) cnt[(int64_t)threadIdx.x * ntile + tile] = h[threadIdx.x];
}

template <typename K, typename V>
__global__ __launch_bounds__(SORT_THREADS) void k_digit_scatter(const K *__restrict__ kin, const V *__restrict__ vin,
                                                                K *__restrict__ kout, V *__restrict__ vout,
                                                                int64_t n, int shift,
                                                                const uint32_t *__restrict__ off, int64_t ntile) {
    __shared__ uint32_t tstart[RADIX];
    const int64_t tile = blockIdx.x;
    const int64_t base = tile * SORT_TILE + (int64_t)threadIdx.x * SORT_ITEMS;  // this thread's run
    K k[SORT_ITEMS];
    V v[SORT_ITEMS];
    uint32_t c[RADIX];
#pragma unroll
    for (int d = 0; d < RADIX; d++) c[d] = 0;
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; it++) {
        const int64_t i = base + it;
        if (i < n) {
            k[it] = kin[i];
            v[it] = vin[i];
            const int d = digit_of(k[it], shift);
#pragma unroll
            for (int q = 0; q < RADIX; q++) c[q] += (q == d);
        }
    }
    // place of this thread's first element of each digit among the tile's
    // elements of that digit: exclusive scan over threads, digit by digit
    uint32_t pos[RADIX];
#pragma unroll
    for (int d = 0; d < RADIX; d++) {
        uint32_t tot;
        pos[d] = block_exclusive<SORT_THREADS, uint32_t>(c[d], &tot);
        (void)tot;
    }
    if (threadIdx.x < RADIX) tstart[threadIdx.x] = off[(int64_t)threadIdx.x * ntile + tile];
    __syncthreads();
#pragma unroll
    for (int d = 0; d < RADIX; d++) pos[d] += tstart[d];
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; it++) {
        const int64_t i = base + it;
        if (i < n) {
            const int d = digit_of(k[it], shift);
            uint32_t p = 0;
#pragma unroll
            for (int q = 0; q < RADIX; q++)
                if (q == d) p = pos[q]++;
            kout[p] = k[it];
            vout[p] = v[it];
        }
    }
}

// exclusive scan, three kernels: tile sums, a scan of the tile sums (one
// block, looping), the tiles' scans with their offsets
template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const T *__restrict__ in, int64_t n,
                                                             T *__restrict__ sums) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    T s = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; it++)
        if (base + it < n) s += in[base + it];
    T tot;
    (void)block_exclusive<SCAN_THREADS, T>(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(T *__restrict__ sums, int64_t m) {
    __shared__ T carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < m; b0 += SCAN_THREADS) {
        const int64_t i = b0 + threadIdx.x;
        const T v = i < m ? sums[i] : (T)0;
        T tot;
        const T ex = block_exclusive<SCAN_THREADS, T>(v, &tot);
        const T c = carry;
        if (i < m) sums[i] = c + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot;
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const T *__restrict__ in, int64_t n,
                                                             const T *__restrict__ sums, T *__restrict__ out) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; it++) {
        v[it] = base + it < n ? in[base + it] : (T)0;
        s += v[it];
    }
    T tot;
    T run = block_exclusive<SCAN_THREADS, T>(s, &tot) + sums[blockIdx.x];
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; it++) {
        if (base + it < n) out[base + it] = run;
        run += v[it];
    }
}

template <typename K, typename V>
__global__ void k_copy_pairs(const K *__restrict__ ki, const V *__restrict__ vi, K *__restrict__ ko,
                             V *__restrict__ vo, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        ko[i] = ki[i];
        vo[i] = vi[i];
    }
}

inline unsigned grid_of(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

template <typename T>
size_t scan_scratch(int64_t n) {
    return sizeof(T) * (size_t)std::max<int64_t>(1, (n + SCAN_TILE - 1) / SCAN_TILE);
}

template <typename T>
hipError_t scan_run(const T *in, T *out, int64_t n, void *scratch, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t m = (n + SCAN_TILE - 1) / SCAN_TILE;
    T *sums = (T *)scratch;
    k_scan_tiles<T><<<(unsigned)m, SCAN_THREADS, 0, st>>>(in, n, sums);
    k_scan_sums<T><<<1, SCAN_THREADS, 0, st>>>(sums, m);
    k_scan_apply<T><<<(unsigned)m, SCAN_THREADS, 0, st>>>(in, n, sums, out);
    return hipGetLastError();
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <typename K, typename V>
size_t sort_scratch(int64_t n) {
    const int64_t ntile = std::max<int64_t>(1, (n + SORT_TILE - 1) / SORT_TILE);
    const int64_t nc = RADIX * ntile;
    // counts, their scan, the scan's tile sums, two ping-pong buffers
    return 2 * align256(sizeof(uint32_t) * (size_t)nc) + align256(scan_scratch<uint32_t>(nc)) +
           2 * (align256(sizeof(K) * (size_t)n) + align256(sizeof(V) * (size_t)n));
}

// keys / values sorted by bits [begin, end) of the key into kout / vout;
// the inputs are left as they are
template <typename K, typename V>
hipError_t sort_run(const K *kin, const V *vin, K *kout, V *vout, int64_t n, int begin, int end, void *scratch,
                    hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t ntile = (n + SORT_TILE - 1) / SORT_TILE;
    const int64_t nc = RADIX * ntile;
    char *w = (char *)scratch;
    uint32_t *cnt = (uint32_t *)w;
    w += align256(sizeof(uint32_t) * (size_t)nc);
    uint32_t *off = (uint32_t *)w;
    w += align256(sizeof(uint32_t) * (size_t)nc);
    void *sscr = w;
    w += align256(scan_scratch<uint32_t>(nc));
    K *kb[2];
    V *vb[2];
    for (int q = 0; q < 2; q++) {
        kb[q] = (K *)w;
        w += align256(sizeof(K) * (size_t)n);
        vb[q] = (V *)w;
        w += align256(sizeof(V) * (size_t)n);
    }
    const int passes = (end - begin + 3) / 4;
    if (passes <= 0) {
        k_copy_pairs<K, V><<<grid_of(n, 256 * 8), 256, 0, st>>>(kin, vin, kout, vout, n);
        return hipGetLastError();
    }
    const K *ks = kin;
    const V *vs = vin;
    for (int p = 0; p < passes; p++) {
        const int shift = begin + 4 * p;
        K *kd = p == passes - 1 ? kout : kb[p & 1];
        V *vd = p == passes - 1 ? vout : vb[p & 1];
        k_digit_hist<K><<<(unsigned)ntile, SORT_THREADS, 0, st>>>(ks, n, shift, cnt, ntile);
        hipError_t e = scan_run<uint32_t>(cnt, off, nc, sscr, st);
        if (e != hipSuccess) return e;
        k_digit_scatter<K, V><<<(unsigned)ntile, SORT_THREADS, 0, st>>>(ks, vs, kd, vd, n, shift, off, ntile);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        ks = kd;
        vs = vd;
    }
    return hipSuccess;
}

}  // namespace

size_t psr_sort_scratch_u64(int64_t n) { return sort_scratch<unsigned long long, uint32_t>(n); }
size_t psr_sort_scratch_u16(int64_t n) { return sort_scratch<uint16_t, uint32_t>(n); }
size_t psr_sort_scratch_u32(int64_t n) { return sort_scratch<uint32_t, uint32_t>(n); }
size_t psr_scan_scratch_i32(int64_t n) { return scan_scratch<int32_t>(n); }

hipError_t psr_sort_pairs_u64(const unsigned long long *kin, const uint32_t *vin, unsigned long long *kout,
                              uint32_t *vout, int64_t n, int begin, int end, void *scratch, hipStream_t st) {
    return sort_run<unsigned long long, uint32_t>(kin, vin, kout, vout, n, begin, end, scratch, st);
}
hipError_t psr_sort_pairs_u16(const uint16_t *kin, const uint32_t *vin, uint16_t *kout, uint32_t *vout, int64_t n,
                              int begin, int end, void *scratch, hipStream_t st) {
    return sort_run<uint16_t, uint32_t>(kin, vin, kout, vout, n, begin, end, scratch, st);
}
hipError_t psr_sort_pairs_u32(const uint32_t *kin, const uint32_t *vin, uint32_t *kout, uint32_t *vout, int64_t n,
                              int begin, int end, void *scratch, hipStream_t st) {
    return sort_run<uint32_t, uint32_t>(kin, vin, kout, vout, n, begin, end, scratch, st);
}
hipError_t psr_exclusive_sum_i32(const int32_t *in, int32_t *out, int64_t n, void *scratch, hipStream_t st) {
    return scan_run<int32_t>(in, out, n, scratch, st);
}
