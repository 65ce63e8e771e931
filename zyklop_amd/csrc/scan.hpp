// Device-wide exclusive prefix sum over u32, built from a wave64 shuffle scan
// and an LDS pass over wave totals (no library scan).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace zk {

constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_PER_THREAD = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_PER_THREAD;

// exclusive scan of v across the block; returns the block total in *total
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (wid == 0) {
        uint32_t w = lane < nw ? lds[lane] : 0u;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t y = __shfl_up(w, d, 64);
            if (lane >= d) w += y;
        }
        if (lane < nw) lds[lane] = w;
    }
    __syncthreads();
    uint32_t prefix = wid ? lds[wid - 1] : 0u;
    *total = lds[nw - 1];
    __syncthreads();
    return prefix + x - v;
}

__global__ void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                             uint32_t* __restrict__ tile_sums, size_t n) {
    __shared__ uint32_t lds[16];
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(s, lds, &total);
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// single block: exclusive scan of the tile sums (n_tiles <= SCAN_TILE)
__global__ void k_scan_tile_sums(uint32_t* __restrict__ sums, int n_tiles) {
    __shared__ uint32_t lds[16];
    int base = threadIdx.x * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        v[k] = (base + k < n_tiles) ? sums[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(s, lds, &total);
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        if (base + k < n_tiles) sums[base + k] = ex;
        ex += v[k];
    }
}

__global__ void k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ tile_sums, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tile_sums[i / SCAN_TILE];
}

// out[0..n) = exclusive scan of in[0..n); tmp needs >= cdiv(n, SCAN_TILE) u32.
inline void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t* tmp, size_t n, hipStream_t st) {
    unsigned tiles = cdiv(n, SCAN_TILE);
    ZK_REQUIRE(tiles <= (unsigned)SCAN_TILE, ZK_E_ARG, "scan too large");
    k_scan_tiles<<<tiles, SCAN_THREADS, 0, st>>>(in, out, tmp, n);
    k_scan_tile_sums<<<1, SCAN_THREADS, 0, st>>>(tmp, (int)tiles);
    k_scan_add<<<cdiv(n, 256), 256, 0, st>>>(out, tmp, n);
    ZK_HIP(hipGetLastError());
}

}  // namespace zk
