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

// Appending to a shared buffer with one atomic per wavefront instead of one
// per item: the lanes that want a slot (any subset of the active lanes) take
// consecutive slots after the leader's reservation, in lane order.
__device__ __forceinline__ uint32_t wave_append(uint32_t *count, bool want) {
    const uint64_t m = __ballot(want);
    const uint32_t lane = threadIdx.x & 63u;
    const int leader = m ? __ffsll((unsigned long long)m) - 1 : 0;
    uint32_t base = 0;
    if (m && lane == (uint32_t)leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// n entries for every lane with one atomic per wavefront (n may be 0); all 64
// lanes must be active
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *count, uint32_t n) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t x = n;   // inclusive prefix sum over the lanes
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, d);
        if (lane >= (uint32_t)d) x += y;
    }
    const uint32_t total = (uint32_t)__shfl((int)x, 63);
    uint32_t base = 0;
    if (lane == 63u && total) base = atomicAdd(count, total);
    base = (uint32_t)__shfl((int)base, 63);
    return base + x - n;
}

// Per-tag reductions (a count or a maximum per member of a batch) over a
// range of items.  One global atomic per item on a per-member word would
// serialize on that word (a member's items all hit it), so each workgroup
// takes a contiguous chunk of the range, reduces tags below TAG_BINS in LDS
// and adds its non-empty bins to the global words; larger tags go to the
// global words directly.  Launch with tag_grid(items) workgroups of kBlock.
constexpr int TAG_BINS = 8192;
inline int tag_grid(int64_t items) {
    int64_t g = (items + 4 * kBlock - 1) / (4 * kBlock);
    return (int)(g < 1 ? 1 : g > 512 ? 512 : g);
}
struct TagBins {
    uint32_t *bins;   // TAG_BINS words of LDS
    __device__ void init() {
        for (int i = threadIdx.x; i < TAG_BINS; i += blockDim.x) bins[i] = 0;
        __syncthreads();
    }
    __device__ void add(uint32_t *global, uint32_t tag, uint32_t x) {
        if (tag < (uint32_t)TAG_BINS) atomicAdd(&bins[tag], x);
        else atomicAdd(&global[tag], x);
    }
    __device__ void max(uint32_t *global, uint32_t tag, uint32_t x) {
        if (tag < (uint32_t)TAG_BINS) atomicMax(&bins[tag], x);
        else atomicMax(&global[tag], x);
    }
    // nTags: the global words that exist (bins past it stay empty)
    __device__ void flush_add(uint32_t *global, uint32_t nTags) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nTags && i < (uint32_t)TAG_BINS; i += blockDim.x)
            if (bins[i]) atomicAdd(&global[i], bins[i]);
    }
    __device__ void flush_max(uint32_t *global, uint32_t nTags) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nTags && i < (uint32_t)TAG_BINS; i += blockDim.x)
            if (bins[i]) atomicMax(&global[i], bins[i]);
    }
};
// this workgroup's chunk [begin, end) of [first, last)
__device__ __forceinline__ void tag_chunk(int64_t first, int64_t last, int64_t &begin, int64_t &end) {
    const int64_t total = last - first, chunk = (total + gridDim.x - 1) / gridDim.x;
    begin = first + (int64_t)blockIdx.x * chunk;
    end = begin + chunk < last ? begin + chunk : last;
}

inline int grid_for(int64_t n, int cap = 4096) {
    int64_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<int>(g);
}
