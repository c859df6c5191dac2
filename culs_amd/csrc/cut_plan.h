// Task lists of the planned cut schedule (GLS_RW_PLAN_CUTS): the kernels
// that list each launch's cones once per pass.  Kept apart from the cut
// kernels so that a host build can run them serially with bounds checks.
#pragma once

constexpr int PLAN_MAX_LEVELS = 6;
constexpr int PLAN_MAX_CONE = (1 << PLAN_MAX_LEVELS) - 1;   // a position-p root's cone has < 2^(p+1) nodes
constexpr uint32_t PLAN_NONE = 0xFFFFFFFFu;
constexpr uint32_t PLAN_GLOBAL = 0xFFu;

// position of v in its launch (lv: the first ids of the K levels, then the end)
__device__ __forceinline__ int plan_pos(const int32_t *__restrict__ lv, int K, uint32_t v) {
    int p = 0;
    while (p + 1 < K && v >= (uint32_t)lv[p + 1]) ++p;
    return p;
}

// the cone of root r in launch j (r first), breadth first; returns its size
__device__ int plan_cone(uint32_t r, int32_t j, const int32_t *__restrict__ nodeLaunch,
                         const uint32_t *__restrict__ f0, const uint32_t *__restrict__ f1, uint32_t *list) {
    int n = 1;
    list[0] = r;
    for (int i = 0; i < n; i++) {
        const uint32_t x = list[i];
        for (int k = 0; k < 2; k++) {
            const uint32_t f = LID(k == 0 ? f0[x] : f1[x]);
            if (nodeLaunch[f] != j) continue;
            bool seen = false;
            for (int q = 0; q < n; q++) seen |= list[q] == f;
            if (!seen && n < PLAN_MAX_CONE) list[n++] = f;
        }
    }
    return n;
}

// launch index of every node of a planned launch (one workgroup per launch)
__global__ void k_plan_fill(int32_t K, const int32_t *__restrict__ lv, int32_t *nodeLaunch) {
    const int32_t *l = lv + (size_t)blockIdx.x * (K + 1);
    for (int v = l[0] + (int)threadIdx.x; v < l[K]; v += blockDim.x) nodeLaunch[v] = (int32_t)blockIdx.x;
}

// roots of position p: nodes there that no higher root's cone holds; each
// bids (atomic min of its id) for the cone nodes it would store
__global__ void k_plan_roots(int32_t n, int32_t nPIs, int32_t K, int32_t p, const uint32_t *__restrict__ f0,
                             const uint32_t *__restrict__ f1, const int32_t *__restrict__ lv,
                             const int32_t *__restrict__ nodeLaunch, uint32_t *own, uint8_t *coneSize) {
    uint32_t list[PLAN_MAX_CONE];
    for (int v = nPIs + 1 + blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const int32_t j = nodeLaunch[v];
        if (j < 0 || own[v] != PLAN_NONE) continue;
        const int32_t *l = lv + (size_t)j * (K + 1);
        if (plan_pos(l, K, (uint32_t)v) != p) continue;
        const int c = plan_cone((uint32_t)v, j, nodeLaunch, f0, f1, list);
        coneSize[v] = (uint8_t)c;
        for (int q = 1; q < c; q++) atomicMin(&own[list[q]], (uint32_t)v);
    }
}

// flag[v]: 1 for a packed position-0 node, 1 << 32 for a root above position 0
__global__ void k_plan_flags(int32_t n, int32_t K, const int32_t *__restrict__ lv,
                             const int32_t *__restrict__ nodeLaunch, const uint32_t *__restrict__ own,
                             unsigned long long *flag) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += gridDim.x * blockDim.x) {
        unsigned long long f = 0;
        const int32_t j = v < n ? nodeLaunch[v] : -1;
        if (j >= 0 && own[v] == PLAN_NONE) {
            const int32_t *l = lv + (size_t)j * (K + 1);
            f = (K > 1 && v >= l[1]) ? (1ull << 32) : 1ull;
        }
        flag[v] = f;
    }
}

// tasks per launch: ceil(packed / 4) + roots
__global__ void k_plan_counts(int32_t nL, int32_t K, const int32_t *__restrict__ lv,
                              const unsigned long long *__restrict__ S, uint32_t *cnt) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > nL) return;
    if (j == nL) { cnt[j] = 0; return; }
    const int32_t *l = lv + (size_t)j * (K + 1);
    const unsigned long long c = S[l[K]] - S[l[0]];
    cnt[j] = ((uint32_t)c + 3u) / 4u + (uint32_t)(c >> 32);
}

// the task of node v (a root or a packed node), or PLAN_NONE; q = its rank
// among the launch's packed nodes
__device__ __forceinline__ uint32_t plan_task(int v, int32_t K, const int32_t *__restrict__ lv,
                                              const int32_t *__restrict__ nodeLaunch, const uint32_t *__restrict__ own,
                                              const unsigned long long *__restrict__ S,
                                              const uint32_t *__restrict__ toff, uint32_t &q, uint32_t &c1,
                                              int32_t &j) {
    j = nodeLaunch[v];
    if (j < 0 || own[v] != PLAN_NONE) return PLAN_NONE;
    const int32_t *l = lv + (size_t)j * (K + 1);
    const unsigned long long Ss = S[l[0]], c = S[l[K]] - Ss, r = S[v] - Ss;
    c1 = (uint32_t)c;
    if (K > 1 && v >= l[1]) {
        q = 0;
        return toff[j] + (c1 + 3u) / 4u + (uint32_t)(r >> 32);
    }
    q = (uint32_t)r;
    return toff[j] + q / 4u;
}

// entries per task
__global__ void k_plan_ecount(int32_t n, int32_t nPIs, int32_t K, const int32_t *__restrict__ lv,
                              const int32_t *__restrict__ nodeLaunch, const uint32_t *__restrict__ own,
                              const unsigned long long *__restrict__ S, const uint32_t *__restrict__ toff,
                              const uint8_t *__restrict__ coneSize, uint32_t *ecnt) {
    for (int v = nPIs + 1 + blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        uint32_t q = 0, c1 = 0;
        int32_t j = 0;
        const uint32_t t = plan_task(v, K, lv, nodeLaunch, own, S, toff, q, c1, j);
        if (t == PLAN_NONE) continue;
        const int32_t *l = lv + (size_t)j * (K + 1);
        if (K > 1 && v >= l[1]) ecnt[t] = coneSize[v];
        else if (q % 4u == 0) ecnt[t] = min(4u, c1 - q);
    }
}

__global__ void k_plan_scatter(int32_t n, int32_t nPIs, int32_t K, const uint32_t *__restrict__ f0,
                               const uint32_t *__restrict__ f1, const int32_t *__restrict__ lv,
                               const int32_t *__restrict__ nodeLaunch, const uint32_t *__restrict__ own,
                               const unsigned long long *__restrict__ S, const uint32_t *__restrict__ toff,
                               const uint32_t *__restrict__ eoff, uint4 *heads, uint4 *ents) {
    uint32_t list[PLAN_MAX_CONE];
    uint8_t posOf[PLAN_MAX_CONE];
    for (int v = nPIs + 1 + blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        uint32_t q = 0, c1 = 0;
        int32_t j = 0;
        const uint32_t t = plan_task(v, K, lv, nodeLaunch, own, S, toff, q, c1, j);
        if (t == PLAN_NONE) continue;
        const int32_t *l = lv + (size_t)j * (K + 1);
        const uint32_t e0 = eoff[t];
        if (!(K > 1 && v >= l[1])) {            // packed position-0 node
            const uint32_t k = q % 4u;
            ents[e0 + k] = make_uint4((uint32_t)v | 1u << 31, f0[v], f1[v], PLAN_GLOBAL | PLAN_GLOBAL << 8);
            if (k == 0) {
                const uint32_t nb = min(4u, c1 - q);
                heads[t] = make_uint4(e0, nb, nb, 0u);
            }
            continue;
        }
        // a root: its cone ordered by position, then id
        const int c = plan_cone((uint32_t)v, j, nodeLaunch, f0, f1, list);
        for (int i = 0; i < c; i++) posOf[i] = (uint8_t)plan_pos(l, K, list[i]);
        for (int i = 1; i < c; i++) {
            const uint32_t x = list[i];
            const uint8_t px = posOf[i];
            int k = i - 1;
            while (k >= 0 && (posOf[k] > px || (posOf[k] == px && list[k] > x))) {
                list[k + 1] = list[k];
                posOf[k + 1] = posOf[k];
                --k;
            }
            list[k + 1] = x;
            posOf[k + 1] = px;
        }
        uint32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < c; i++) {
            const uint32_t x = list[i];
            const uint32_t l0 = f0[x], l1 = f1[x];
            uint32_t src0 = PLAN_GLOBAL, src1 = PLAN_GLOBAL;
            for (int k = 0; k < i; k++) {
                if (list[k] == LID(l0)) src0 = (uint32_t)k;
                if (list[k] == LID(l1)) src1 = (uint32_t)k;
            }
            const bool root = x == (uint32_t)v;
            const uint32_t stored = (root || own[x] == (uint32_t)v) ? 1u : 0u;
            const uint32_t kept = root ? 0u : 1u;
            ents[e0 + i] = make_uint4(x | stored << 31 | kept << 30, l0, l1, src0 | src1 << 8);
            cnt[posOf[i]]++;
        }
        const uint32_t top = posOf[c - 1];
        heads[t] = make_uint4(e0, (uint32_t)c | top << 16, cnt[0] | cnt[1] << 8 | cnt[2] << 16 | cnt[3] << 24,
                              cnt[4] | cnt[5] << 8 | cnt[6] << 16 | cnt[7] << 24);
    }
}

