// Device helpers shared by the HIP translation units: literal helpers, the
// open-addressing strash table on 64-bit keys, and the frontier controller used
// by the level-synchronous sweeps.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define LID(l) ((l) >> 1)
#define LC(l) ((l) & 1u)
#define INVALID32 0xFFFFFFFFu
#define KEY_EMPTY 0xFFFFFFFFFFFFFFFFull

constexpr int kBlock = 256;

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ uint64_t and_key(uint32_t l0, uint32_t l1) {
    return (static_cast<uint64_t>(l0) << 32) | l1;
}

// Insert `key` with value `val`, keeping the minimum value over all inserts of
// the same key (values must start at ~0).  Returns the slot, or -1 if the table
// is full (probing is bounded by the capacity; *err is set).
__device__ __forceinline__ int64_t ht_insert_min(unsigned long long *keys, unsigned long long *vals,
                                                 uint64_t mask, uint64_t key, uint64_t val, uint32_t *err) {
    uint64_t s = mix64(key) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        unsigned long long prev = atomicCAS(&keys[s], KEY_EMPTY, (unsigned long long)key);
        if (prev == KEY_EMPTY || prev == key) {
            atomicMin(&vals[s], (unsigned long long)val);
            return (int64_t)s;
        }
        s = (s + 1) & mask;
    }
    atomicOr(err, 1u);
    return -1;
}

// Frontier controller of a level-synchronous sweep.  order[] holds the
// frontiers back to back: F_{k-1} = [prevStart, +prevSize), F_k = [curStart,
// +curSize), F_{k+1} is appended after F_k.  The last block of each round
// advances the window, so the host can launch rounds without reading back.
struct SweepCtl {
    uint32_t prevStart, prevSize, curStart, curSize;
    uint32_t nextSize, blocksDone, depth, err;
    uint32_t processed, pad[7];
};

__device__ __forceinline__ void sweep_round_end(SweepCtl *ctl, uint32_t k) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        uint32_t done = atomicAdd(&ctl->blocksDone, 1u);
        if (done == gridDim.x - 1) {
            __threadfence();
            uint32_t cur = atomicAdd(&ctl->curSize, 0u);
            uint32_t nxt = atomicAdd(&ctl->nextSize, 0u);
            uint32_t cs = ctl->curStart;
            if (cur > 0) ctl->depth = k;
            ctl->processed += cur;
            ctl->prevStart = cs;
            ctl->prevSize = cur;
            ctl->curStart = cs + cur;
            ctl->curSize = nxt;
            ctl->nextSize = 0;
            __threadfence();
            atomicExch(&ctl->blocksDone, 0u);
        }
    }
}

// Adds every thread's x to *dst with one global atomic per workgroup; every
// thread of the workgroup must call it.
__device__ __forceinline__ void block_add_u64(unsigned long long *dst, uint32_t x) {
    __shared__ unsigned long long s_sum;
    if (threadIdx.x == 0) s_sum = 0;
    __syncthreads();
    if (x) atomicAdd(&s_sum, (unsigned long long)x);
    __syncthreads();
    if (threadIdx.x == 0 && s_sum) atomicAdd(dst, s_sum);
}

inline int grid_for(int64_t n, int cap = 4096) {
    int64_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<int>(g);
}
