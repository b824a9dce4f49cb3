/*
 * st_sort.hip - stable LSD radix sort of (32-bit key, 32-bit value) pairs on
 * the device (gfx950, wave64), used for the Morton-code sorts of the scene
 * and mesh LBVH builds.
 *
 * Each pass sorts by one 8-bit digit:
 *   k_rs_count    per tile of RS_TILE keys, the count of each digit (LDS
 *                 counters), written digit-major: hist[digit * tiles + tile];
 *   k_scan_*      exclusive prefix sum of hist (a tile's share of a digit
 *                 starts after every earlier digit and the digit's share of
 *                 every earlier tile);
 *   k_rs_scatter  each tile's keys in index order, 256 at a time: a key's
 *                 rank among equal digits is the count in earlier rounds of
 *                 the tile (LDS), in earlier waves of the round (LDS) and in
 *                 earlier lanes of its wave (eight ballots give the lanes
 *                 with the same digit); the pair goes to hist + rank.
 * Equal keys keep their input order in every pass, so the sort is stable.
 */
#include <hip/hip_runtime.h>

#include "st_sort.h"

#define RS_THREADS 256u              /* = digit values per pass */
#define RS_WAVES (RS_THREADS / 64u)
#define RS_ITEMS 16u
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define SC_THREADS 1024u
#define SC_BLOCK (2u * SC_THREADS)   /* elements per scan block */

__global__ void __launch_bounds__(RS_THREADS) k_rs_count(const uint32_t *__restrict__ keys, uint32_t n,
                                                         uint32_t shift, uint32_t tiles, uint32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[RS_THREADS];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
    for (uint32_t r = 0; r < RS_ITEMS; r++) {
        const uint32_t i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

/* inclusive prefix sum over the lanes of a wave */
__device__ static inline uint32_t wave_scan(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t off = 1; off < 64u; off <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

/* exclusive prefix sum of a[b * SC_BLOCK, ...) in place; the block's total
   to sums[b] (when sums) */
__global__ void __launch_bounds__(SC_THREADS) k_scan_blocks(uint32_t *a, uint32_t len, uint32_t *sums) {
    __shared__ uint32_t wsum[SC_THREADS / 64u];
    const uint32_t t = threadIdx.x, i = blockIdx.x * SC_BLOCK + 2u * t;
    const uint32_t x0 = i < len ? a[i] : 0u, x1 = i + 1u < len ? a[i + 1u] : 0u;
    const uint32_t incl = wave_scan(x0 + x1);
    if ((t & 63u) == 63u) wsum[t >> 6] = incl;
    __syncthreads();
    if (t < 64u) {
        const uint32_t v = t < SC_THREADS / 64u ? wsum[t] : 0u;
        const uint32_t s = wave_scan(v);
        if (t < SC_THREADS / 64u) wsum[t] = s - v;   /* exclusive wave offsets */
    }
    __syncthreads();
    const uint32_t excl = wsum[t >> 6] + incl - (x0 + x1);
    if (i < len) a[i] = excl;
    if (i + 1u < len) a[i + 1u] = excl + x0;
    if (sums && t == SC_THREADS - 1u) sums[blockIdx.x] = excl + x0 + x1;
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_add(uint32_t *a, uint32_t len, const uint32_t *sums) {
    const uint32_t add = sums[blockIdx.x];
    const uint32_t i = blockIdx.x * SC_BLOCK + threadIdx.x;
    if (i < len) a[i] += add;
    if (i + SC_THREADS < len) a[i + SC_THREADS] += add;
}

__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint32_t *__restrict__ kin,
                                                           const uint32_t *__restrict__ vin,
                                                           uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                           uint32_t n, uint32_t shift, uint32_t tiles,
                                                           const uint32_t *__restrict__ hist) {
    __shared__ uint32_t run[RS_THREADS];            /* next output position of each digit */
    __shared__ uint32_t wc[RS_WAVES][RS_THREADS];   /* this round: keys of each digit per wave */
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    run[t] = hist[(size_t)t * tiles + blockIdx.x];
    for (uint32_t w = 0; w < RS_WAVES; w++) wc[w][t] = 0;
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t base = blockIdx.x * RS_TILE;
    for (uint32_t r = 0; r < RS_ITEMS; r++) {
        const uint32_t i = base + r * RS_THREADS + t;
        const bool valid = i < n;
        uint32_t k = 0, v = 0, d = 0;
        if (valid) {
            k = kin[i];
            v = vin[i];
            d = (k >> shift) & 0xFFu;
        }
        /* lanes of this wave with the same digit */
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8u; b++) {
            const uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t below = (uint32_t)__builtin_popcountll(peers & lt);
        if (valid && below == 0) wc[wv][d] = (uint32_t)__builtin_popcountll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = run[d] + below;
            for (uint32_t w = 0; w < wv; w++) off += wc[w][d];
            kout[off] = k;
            vout[off] = v;
        }
        __syncthreads();
        uint32_t s = 0;
        for (uint32_t w = 0; w < RS_WAVES; w++) {
            s += wc[w][t];
            wc[w][t] = 0;
        }
        run[t] += s;
        __syncthreads();
    }
}

static size_t r256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t st_radix_sort_scratch_bytes(uint32_t n) {
    const size_t tiles = (n + RS_TILE - 1) / RS_TILE;
    const size_t len = tiles * RS_THREADS;
    const size_t b1 = (len + SC_BLOCK - 1) / SC_BLOCK;
    return r256(4 * len) + r256(4 * b1) + 256;
}

int st_scan_exclusive(hipStream_t stream, uint32_t *a, uint32_t len, uint32_t *sums) {
    const uint32_t b1 = (len + SC_BLOCK - 1) / SC_BLOCK;
    if (len == 0) return 0;
    if (b1 > SC_BLOCK) return -1;
    hipLaunchKernelGGL(k_scan_blocks, dim3(b1), dim3(SC_THREADS), 0, stream, a, len, b1 > 1 ? sums : nullptr);
    if (b1 > 1) {
        hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(SC_THREADS), 0, stream, sums, b1, nullptr);
        hipLaunchKernelGGL(k_scan_add, dim3(b1), dim3(SC_THREADS), 0, stream, a, len, (const uint32_t *)sums);
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int st_radix_sort_pairs(hipStream_t stream, uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp,
                        uint32_t n, uint32_t key_bits, void *scratch) {
    if (n <= 1) return 0;
    if (key_bits == 0 || key_bits > 32) return -1;
    const uint32_t tiles = (n + RS_TILE - 1) / RS_TILE;
    const uint32_t len = tiles * RS_THREADS;
    const uint32_t b1 = (len + SC_BLOCK - 1) / SC_BLOCK;
    if (b1 > SC_BLOCK) return -1;   /* the block sums are scanned by one block */
    uint32_t *hist = (uint32_t *)scratch;
    uint32_t *sums = (uint32_t *)((char *)scratch + r256(4ull * len));
    uint32_t *kin = keys, *vin = vals, *kout = keys_tmp, *vout = vals_tmp;
    for (uint32_t shift = 0; shift < key_bits; shift += 8u) {
        hipLaunchKernelGGL(k_rs_count, dim3(tiles), dim3(RS_THREADS), 0, stream, (const uint32_t *)kin, n, shift,
                           tiles, hist);
        hipLaunchKernelGGL(k_scan_blocks, dim3(b1), dim3(SC_THREADS), 0, stream, hist, len, b1 > 1 ? sums : nullptr);
        if (b1 > 1) {
            hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(SC_THREADS), 0, stream, sums, b1, nullptr);
            hipLaunchKernelGGL(k_scan_add, dim3(b1), dim3(SC_THREADS), 0, stream, hist, len, (const uint32_t *)sums);
        }
        hipLaunchKernelGGL(k_rs_scatter, dim3(tiles), dim3(RS_THREADS), 0, stream, (const uint32_t *)kin,
                           (const uint32_t *)vin, kout, vout, n, shift, tiles, (const uint32_t *)hist);
        uint32_t *tk = kin, *tv = vin;
        kin = kout;
        vin = vout;
        kout = tk;
        vout = tv;
    }
    if (kin != keys) {
        if (hipMemcpyAsync(keys, kin, 4ull * n, hipMemcpyDeviceToDevice, stream) != hipSuccess) return -2;
        if (hipMemcpyAsync(vals, vin, 4ull * n, hipMemcpyDeviceToDevice, stream) != hipSuccess) return -2;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
