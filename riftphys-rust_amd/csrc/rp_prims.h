// Device-wide primitives used by the step: exclusive scan of u32 counts and a stable LSD radix
// sort of (u32 key, u32 value) pairs.  Both are deterministic: no floating-point accumulation,
// and the sort's intra-block ranks come from 64-lane ballots (stable within a wavefront) plus
// per-wavefront digit counts in LDS, never from atomic arrival order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RP_SCAN_BLOCK 1024  // elements per scan block (256 threads x 4)

// Block-local exclusive scan of 4 consecutive u32 per thread (256 threads).
__device__ __forceinline__ uint32_t rp_block_exscan256(uint32_t v, uint32_t *lds /*[256]*/, uint32_t *total) {
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t x = t >= off ? lds[t - off] : 0u;
        __syncthreads();
        lds[t] += x;
        __syncthreads();
    }
    uint32_t incl = lds[t];
    if (total) *total = lds[255];
    __syncthreads();
    return incl - v;
}

// Pass 1: per-block sums.
static __global__ void __launch_bounds__(256) k_scan_reduce(const uint32_t *in, uint32_t n, uint32_t *block_sums) {
    __shared__ uint32_t lds[256];
    uint32_t base = blockIdx.x * RP_SCAN_BLOCK + threadIdx.x * 4;
    uint32_t s = 0;
    for (int k = 0; k < 4; k++) s += (base + k < n) ? in[base + k] : 0u;
    uint32_t tot;
    rp_block_exscan256(s, lds, &tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// Pass 2: one block scans the block sums (any count), writes the grand total.
static __global__ void __launch_bounds__(256) k_scan_sums(uint32_t *sums, uint32_t nblocks, uint32_t *total_out) {
    __shared__ uint32_t lds[256];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nblocks; base += 1024) {
        uint32_t idx = base + threadIdx.x * 4;
        uint32_t v[4], s = 0;
        for (int k = 0; k < 4; k++) { v[k] = (idx + k < nblocks) ? sums[idx + k] : 0u; s += v[k]; }
        uint32_t tot;
        uint32_t ex = rp_block_exscan256(s, lds, &tot);
        uint32_t run = carry + ex;
        for (int k = 0; k < 4; k++) { if (idx + k < nblocks) sums[idx + k] = run; run += v[k]; }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// Pass 3: exclusive scan within each block plus the block's offset.  out may alias in.
static __global__ void __launch_bounds__(256) k_scan_apply(const uint32_t *in, uint32_t n, const uint32_t *block_sums, uint32_t *out) {
    __shared__ uint32_t lds[256];
    uint32_t base = blockIdx.x * RP_SCAN_BLOCK + threadIdx.x * 4;
    uint32_t v[4], s = 0;
    for (int k = 0; k < 4; k++) { v[k] = (base + k < n) ? in[base + k] : 0u; s += v[k]; }
    uint32_t ex = rp_block_exscan256(s, lds, nullptr);
    uint32_t run = block_sums[blockIdx.x] + ex;
    for (int k = 0; k < 4; k++) { if (base + k < n) out[base + k] = run; run += v[k]; }
}

// ---------------------------------------------------------------- radix sort
#define RP_RS_TILE 1024  // items per block per pass (256 threads x 4 rounds)

// n = min(*n_ptr, cap): the device count can exceed the buffers' capacity on an overflowing tick
// (which is then rolled back), so the sort never reads or writes past cap.
static __global__ void __launch_bounds__(256) k_rs_hist(const uint32_t *keys, const uint32_t *n_ptr, uint32_t cap, uint32_t shift,
                                                 uint32_t nblocks_alloc, uint32_t *hist) {
    __shared__ uint32_t h[256];
    const uint32_t n = min(*n_ptr, cap);
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * RP_RS_TILE;
    for (int r = 0; r < 4; r++) {
        uint32_t i = base + r * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks_alloc + blockIdx.x] = h[threadIdx.x];
}

static __global__ void __launch_bounds__(256) k_rs_scatter(const uint32_t *keys, const uint32_t *vals, const uint32_t *n_ptr, uint32_t cap,
                                                    uint32_t shift, uint32_t nblocks_alloc, const uint32_t *hist_off, uint32_t *keys_out,
                                                    uint32_t *vals_out) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[4][256];
    const uint32_t n = min(*n_ptr, cap);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    run[t] = hist_off[t * nblocks_alloc + blockIdx.x];
    uint32_t base = blockIdx.x * RP_RS_TILE;
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int r = 0; r < 4; r++) {
        for (int w = 0; w < 4; w++) wcnt[w][t] = 0;
        __syncthreads();
        uint32_t i = base + r * 256 + t;
        bool valid = i < n;
        uint32_t key = valid ? keys[i] : 0u;
        uint32_t dig = (key >> shift) & 255u;
        uint64_t active = __ballot(valid);
        uint64_t peers = active;
        for (int b = 0; b < 8; b++) {
            uint64_t m = __ballot((dig >> b) & 1u);
            peers &= ((dig >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[wave][dig] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[dig] + rank;
            for (int w = 0; w < wave; w++) pos += wcnt[w][dig];
            keys_out[pos] = key;
            vals_out[pos] = vals[i];
        }
        __syncthreads();
        run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        __syncthreads();
    }
}
