// Balance (`b`) on the device.
//
// Semantics of balancePerformV2 (src/algorithms/balance.cu:912-1382) with
// sortDecId = 1, made exact:
//   1. covers: from every root AND node a DFS (fanin0 first) collects the
//      cover's inputs: literals that are complemented, not AND nodes, have
//      more than one fanout (POs count), or sit at the stack / size limits
//      (128 / 1024); a repeated input is dropped and an input seen in both
//      polarities makes the cover constant FALSE (balance.cu:666-750).
//      Roots are found level by level from the POs; a root's level is the
//      last (deepest) level it appears at (balance.cu:958-1037, 753-834).
//   2. reconstruction, deepest level first, covers of a level in id order,
//      in batches of 32768 (balance.cu:1129-1229): a cover's inputs are
//      sorted by (level desc, literal desc); then, step by step over the whole
//      batch, the last input may be swapped with an earlier one of the same
//      level that already forms an AND with the last (sharedNodeDrivenPermute,
//      :187-247); the last two inputs are popped and their AND is looked up
//      (trivial rules, then the table, :113-185); new ANDs take id
//      counter + cover index, the smallest id wins among equal keys, the
//      counter moves past the largest id (:1188-1210); the result goes back
//      into the cover in level order unless already present (:429-467).
//   3. created nodes are renumbered in id order (:1237-1262); POs follow
//      their cover's result (:882-910).
// Literal values decide ties in the input order; there the reference's own
// values are used (lit 0 = TRUE, common.h:285-286), not the AIGER ones.
//
// Mapping to MI355X: cover discovery is two frontier sweeps (reachability,
// then a topological sweep of the cover graph; each root is touched twice,
// several rounds per host check).  Reconstruction of
// narrow levels runs in ONE persistent workgroup across levels and steps, so
// a step costs a few workgroup barriers instead of kernel launches and a host
// round trip; the hash tables are global and all of their traffic is atomics,
// which the L2 serves.  Wide levels (more than balance_narrow_max covers, a
// gls_tuning setting) run each phase of a step as a grid-wide kernel.
#include "gls_internal.h"
#include "dev_common.h"
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <algorithm>
#include <chrono>
#include <vector>

namespace {

constexpr int kB = 256;
constexpr int BAL_STACK = 128;        // DFS_COVER_STACK_SIZE, balance.cu:18
constexpr int BAL_MAX_COVER = 1024;   // MAX_LOCAL_COVER_SIZE, balance.cu:17
constexpr int BAL_BATCH = 32768;      // iterLen, balance.cu:1129
constexpr int BAL_WG = 1024;          // threads of the persistent reconstruction workgroup
constexpr int COVER_THREADS = 16384;  // threads of a cover-discovery launch (each owns COVER_SCRATCH slots),
constexpr int COVER_THREADS_MAX = 131072;   // scaled up to this with the graph (a level's frontier in one round)
// the size limit is checked before an expansion, so a cover can end up to a
// few inputs above BAL_MAX_COVER: room for that
constexpr int COVER_SCRATCH = BAL_MAX_COVER + 64;
constexpr int COVER_LEVELS_PER_CHECK = 32;  // cover-discovery levels launched per host check
constexpr uint32_t UNSET = 0xFFFFFFFFu;
constexpr unsigned long long EMPTY64 = ~0ull;

inline int nblocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kB - 1) / kB, 1 << 20)); }

__device__ __forceinline__ uint32_t culs_val(uint32_t l) { return l < 2u ? 1u - l : l; }

// ---------------------------------------------------------------------------
// hash tables (open addressing, linear probing; keys never removed)
// ---------------------------------------------------------------------------
struct Tables {
    unsigned long long *rk;   // reconstruct table: key (lit0 << 32 | lit1)
    uint32_t *rv;             //   value: the node's id (smallest on a tie)
    uint64_t rmask;
    uint32_t *lk, *lv;        // level table: id -> level
    uint64_t lmask;
};

__device__ __forceinline__ uint32_t r_find(const Tables &t, unsigned long long key) {
    for (uint64_t s = mix64(key) & t.rmask, probes = 0; probes <= t.rmask; s = (s + 1) & t.rmask, probes++) {
        const unsigned long long k = __hip_atomic_load(&t.rk[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == key) return __hip_atomic_load(&t.rv[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == EMPTY64) return UNSET;
    }
    return UNSET;
}

// insert key with value val, keeping the smallest value; returns false if full
__device__ __forceinline__ bool r_insert_min(const Tables &t, unsigned long long key, uint32_t val) {
    for (uint64_t s = mix64(key) & t.rmask, probes = 0; probes <= t.rmask; s = (s + 1) & t.rmask, probes++) {
        unsigned long long prev = EMPTY64;
        __hip_atomic_compare_exchange_strong(&t.rk[s], &prev, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        if (prev == EMPTY64 || prev == key) {
            __hip_atomic_fetch_min(&t.rv[s], val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ bool l_insert(const Tables &t, uint32_t id, uint32_t lev) {
    for (uint64_t s = mix64(id) & t.lmask, probes = 0; probes <= t.lmask; s = (s + 1) & t.lmask, probes++) {
        uint32_t prev = UNSET;
        __hip_atomic_compare_exchange_strong(&t.lk[s], &prev, id, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        if (prev == UNSET || prev == id) {
            __hip_atomic_fetch_min(&t.lv[s], lev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ uint32_t l_find(const Tables &t, uint32_t id) {
    for (uint64_t s = mix64(id) & t.lmask, probes = 0; probes <= t.lmask; s = (s + 1) & t.lmask, probes++) {
        const uint32_t k = __hip_atomic_load(&t.lk[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == id) return __hip_atomic_load(&t.lv[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == UNSET) return UNSET;
    }
    return UNSET;
}

// trivial AND rules (hash_table.h:435-449) in AIGER literals; UNSET if none
__device__ __forceinline__ uint32_t and_trivial(uint32_t a, uint32_t b) {
    if (a == b) return a;
    if (a == (b ^ 1u)) return 0u;
    if (a == 1u) return b;
    if (b == 1u) return a;
    if (a == 0u || b == 0u) return 0u;
    return UNSET;
}

// retrieveAndNode (balance.cu:87-111): trivial result, existing node, or UNSET
__device__ __forceinline__ uint32_t and_lookup(const Tables &t, uint32_t a, uint32_t b) {
    const uint32_t tr = and_trivial(a, b);
    if (tr != UNSET) return tr;
    if (a > b) { const uint32_t x = a; a = b; b = x; }
    const uint32_t v = r_find(t, (unsigned long long)a << 32 | b);
    return v == UNSET ? UNSET : 2u * v;
}

__device__ __forceinline__ uint32_t lit_level(const Tables &t, uint32_t lit, int32_t nPIs) {
    const uint32_t id = lit >> 1;
    if ((int32_t)id <= nPIs) return 0u;
    const uint32_t l = l_find(t, id);
    return l == UNSET ? 0u : l;
}

// ---------------------------------------------------------------------------
// phase 1: covers
// ---------------------------------------------------------------------------
__global__ void k_fanouts(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0, const uint32_t *__restrict__ f1,
                          int32_t nPOs, const uint32_t *__restrict__ pos, uint32_t *nFo) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += stride) {
        atomicAdd(&nFo[LID(f0[v])], 1u);
        atomicAdd(&nFo[LID(f1[v])], 1u);
    }
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nPOs; i += stride)
        atomicAdd(&nFo[LID(pos[i])], 1u);
}

// first frontier: the AND nodes driving POs, each once
__global__ void k_po_frontier(int32_t nPOs, int32_t nPIs, const uint32_t *__restrict__ pos, uint32_t *stamp,
                              uint32_t *front, uint32_t *frontLen) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nPOs; i += stride) {
        const uint32_t id = LID(pos[i]);
        const bool add = (int32_t)id > nPIs && atomicExch(&stamp[id], 1u) != 1u;
        const uint32_t q = wave_append(frontLen, add);
        if (add) front[q] = id;
    }
}

struct CoverArgs {
    const uint32_t *f0, *f1, *nFo;
    int32_t nPIs;
    uint32_t *coverOff, *coverLen;   // coverLen UNSET: not computed
    uint32_t *pool, *poolNext;
    uint64_t poolCap;
    uint32_t *scratch;               // COVER_SCRATCH entries per thread
    uint32_t *lastLevel, *stamp;
    uint32_t *indeg;                 // parent covers of a root not yet taken by the topological sweep
    uint32_t *err;                   // 1: cover pool full
};

// localCoverTravToTable (balance.cu:666-750) of root v into in[]; returns len
// (0 = constant FALSE)
__device__ uint32_t cover_dfs(const CoverArgs &a, uint32_t v, uint32_t *in) {
    uint32_t stack[BAL_STACK];
    int top = -1;
    uint32_t len = 0;
    const uint32_t rootLit = 2u * v;
    stack[++top] = rootLit;
    while (top != -1) {
        const uint32_t cur = stack[top--];
        bool rep = false;
        for (uint32_t i = 0; i < len; i++) {
            const uint32_t x = in[i];
            if (x == cur) { rep = true; break; }
            if (x == (cur ^ 1u)) return 0u;
        }
        if (rep) continue;
        const uint32_t id = cur >> 1;
        if (cur != rootLit && ((cur & 1u) || (int32_t)id <= a.nPIs || a.nFo[id] > 1u || top + 2 >= BAL_STACK ||
                               (int)len + 2 * (top + 1) + 1 >= BAL_MAX_COVER)) {
            in[len++] = cur;
            continue;
        }
        stack[++top] = a.f1[id];
        stack[++top] = a.f0[id];
    }
    return len;
}

// Root levels.  The reference sweeps levels from the POs, putting on level
// L + 1 every AND input of a cover on level L (again on every level it is
// reached on), and keeps for each root the last level it appeared on
// (balance.cu:958-1037): a root's level is the length of the longest cover
// path from a PO to it, and the number of levels is one more than the
// longest.  Here each root is visited twice instead of once per path length:
//   reach: a sweep from the POs in which each root is put on a frontier the
//          first time it is seen; its cover is computed then, and each AND
//          input of the cover counts one more parent cover;
//   kahn:  a synchronous topological sweep of the cover graph: round 0 holds
//          the roots without a parent cover (PO drivers), and a root goes on
//          the next round's frontier when its last parent cover is taken.
// A root taken in round L has its longest path from a PO equal to L.
// flen[r] = size of round r's frontier, kept on the device so that several
// rounds are launched per host check; round r's frontier is in
// fronts[r & 1].
__global__ void k_cover_reach(CoverArgs a, uint32_t *front0, uint32_t *front1, uint32_t *flen, uint32_t round) {
    const uint32_t frontLen = flen[round];
    const uint32_t *front = (round & 1u) ? front1 : front0;
    uint32_t *next = (round & 1u) ? front0 : front1;
    uint32_t *nextLen = &flen[round + 1];
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    uint32_t *in = a.scratch + (size_t)tid * COVER_SCRATCH;
    const int64_t lane = threadIdx.x & 63;
    // the wavefront runs the same number of iterations (its first lane's
    // index decides), so the pool reservation sees all 64 lanes
    for (int64_t k = tid; k - lane < (int64_t)frontLen; k += stride) {
        const bool has = k < (int64_t)frontLen;
        const uint32_t v = has ? front[k] : 0u;
        const uint32_t len = has ? cover_dfs(a, v, in) : 0u;
        const uint32_t off = wave_reserve(a.poolNext, len);
        if (!has) continue;
        if ((uint64_t)off + len > a.poolCap) {
            atomicOr(a.err, 1u);
            a.coverLen[v] = 0;
            continue;
        }
        for (uint32_t i = 0; i < len; i++) {
            const uint32_t lit = in[i], id = lit >> 1;
            a.pool[off + i] = lit;
            bool add = false;
            if ((int32_t)id > a.nPIs) {
                atomicAdd(&a.indeg[id], 1u);
                add = atomicExch(&a.stamp[id], 1u) != 1u;
            }
            const uint32_t q = wave_append(nextLen, add);
            if (add) next[q] = id;
        }
        a.coverOff[v] = off;
        a.coverLen[v] = len;
    }
}

// round 0 of the topological sweep: the PO drivers that no cover has as input
__global__ void k_kahn_first(int32_t nPOs, int32_t nPIs, const uint32_t *__restrict__ pos,
                             const uint32_t *__restrict__ indeg, uint32_t *stamp, uint32_t *front,
                             uint32_t *frontLen) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nPOs; i += stride) {
        const uint32_t id = LID(pos[i]);
        const bool add = (int32_t)id > nPIs && indeg[id] == 0u && atomicExch(&stamp[id], 2u) != 2u;
        const uint32_t q = wave_append(frontLen, add);
        if (add) front[q] = id;
    }
}

__global__ void k_cover_kahn(CoverArgs a, uint32_t *front0, uint32_t *front1, uint32_t *flen, uint32_t round) {
    const uint32_t frontLen = flen[round];
    const uint32_t *front = (round & 1u) ? front1 : front0;
    uint32_t *next = (round & 1u) ? front0 : front1;
    uint32_t *nextLen = &flen[round + 1];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < frontLen; k += stride) {
        const uint32_t v = front[k];
        a.lastLevel[v] = round;
        const uint32_t len = a.coverLen[v], off = a.coverOff[v];
        for (uint32_t i = 0; i < len; i++) {
            const uint32_t id = a.pool[off + i] >> 1;
            const bool add = (int32_t)id > a.nPIs && atomicSub(&a.indeg[id], 1u) == 1u;
            const uint32_t q = wave_append(nextLen, add);
            if (add) next[q] = id;
        }
    }
}

__global__ void k_root_flags(int32_t n, const uint32_t *__restrict__ lastLevel, uint8_t *flag) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += stride)
        flag[v] = lastLevel[v] != UNSET;
}

__global__ void k_gather_level(int32_t m, const uint32_t *__restrict__ ids, const uint32_t *__restrict__ lastLevel,
                               uint32_t *keys) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += stride) keys[i] = lastLevel[ids[i]];
}

// ---------------------------------------------------------------------------
// phase 2: reconstruction
// ---------------------------------------------------------------------------
struct RecArgs {
    Tables t;
    int32_t nPIs;
    const uint32_t *roots;            // roots by (level, id)
    const uint32_t *coverOff, *coverLen, *pool;
    uint32_t *newLit;                 // per root id
    uint32_t *arr, *alv;              // batch workspace: maxLen entries per cover
    uint32_t *alen;
    uint32_t maxLen;
    uint8_t *mask;
    unsigned long long *skey;
    uint32_t *sval, *slev;
    uint32_t *counter;                // [0] next tentative id, [1] step max, [2] any, [3] created
    uint32_t *err;                    // 2: table full
    const uint32_t *tagIn;            // optional: per-node tag of the input (batched balance)
    Tables tagTab;                    // created id -> tag of the cover that made it
    uint32_t shuffle;                 // 1: `b -s` input order (sortDecId = 0)
};

// thrust::shuffle(thrust::seq, ..., default_random_engine(seed)) as an index
// map, the tie-break of `b -s` (balance.cu:301, 404-411).  The algorithm is
// thrust's (rocThrust thrust/system/detail/generic/shuffle.inl, ROCm 7.2): a
// 24-round Feistel bijection over the next power of two (at least 16) with
// round keys from minstd_rand seeded by `seed` (0 taken as 1); position k of
// the shuffled sequence holds the input whose index is the k-th value below
// len that the bijection yields for 0, 1, 2, ...
struct ThrustShuffle {
    uint32_t key[24];
    uint32_t lb, rb, lmask, rmask, span;
    __device__ ThrustShuffle(uint32_t len, uint32_t seed) {
        uint32_t x = seed % 2147483647u;
        if (x == 0) x = 1;
        uint32_t bits = 4;
        if (len > 16) {
            bits = 0;
            for (uint32_t m = len - 1; m; m >>= 1) bits++;
        }
        lb = bits / 2;
        rb = bits - lb;
        lmask = (1u << lb) - 1u;
        rmask = (1u << rb) - 1u;
        span = 1u << bits;
        for (int i = 0; i < 24; i++) {
            x = (uint32_t)(((uint64_t)x * 48271u) % 2147483647u);
            key[i] = x;
        }
    }
    __device__ uint32_t operator()(uint32_t j) const {
        uint32_t s0 = j >> rb, s1 = j & rmask;
        for (int r = 0; r < 24; r++) {
            const uint64_t prod = 0xD2B74407B1CE6E93ull * (uint64_t)s0;
            const uint32_t hi = (uint32_t)(prod >> 32);
            const uint32_t lo = ((uint32_t)prod << (rb - lb)) | (s1 >> lb);
            s0 = ((hi ^ key[r]) ^ s1) & lmask;
            s1 = lo & rmask;
        }
        return (s0 << rb) | s1;
    }
};

__device__ __forceinline__ void rec_tag(const RecArgs &r, uint32_t c, uint32_t root) {
    if (r.tagIn && r.mask[c] == 1 && !l_insert(r.tagTab, r.sval[c], r.tagIn[root])) atomicOr(r.err, 2u);
}

// prepareReconstructArrays (balance.cu:272-427), one cover
__device__ void rec_gather(const RecArgs &r, uint32_t c, uint32_t root) {
    uint32_t *A = r.arr + (size_t)c * r.maxLen, *LV = r.alv + (size_t)c * r.maxLen;
    const uint32_t len = r.coverLen[root], off = r.coverOff[root];
    r.alen[c] = len;
    auto put = [&](uint32_t k, uint32_t src) {
        const uint32_t lit = r.pool[off + src], id = lit >> 1;
        const uint32_t nl = (int32_t)id <= r.nPIs ? lit : (r.newLit[id] ^ (lit & 1u));
        A[k] = nl;
        LV[k] = lit_level(r.t, nl, r.nPIs);
    };
    if (r.shuffle) {
        // `b -s`: the inputs in the order of the shuffle seeded with the
        // cover's index in its batch
        if (len > 0) {
            const ThrustShuffle sh(len, c);
            uint32_t k = 0;
            for (uint32_t j = 0; j < sh.span && k < len; j++) {
                const uint32_t src = sh(j);
                if (src < len) put(k++, src);
            }
        }
    } else {
        for (uint32_t i = 0; i < len; i++) put(i, i);
    }
    // level descending, then (`b`) the reference's literal value descending;
    // insertion sort keeps equal keys in order, as thrust's sequential sort does
    for (uint32_t i = 1; i < len; i++) {
        const uint32_t x = A[i], xl = LV[i];
        int j = (int)i - 1;
        while (j >= 0 && (LV[j] < xl || (!r.shuffle && LV[j] == xl && culs_val(A[j]) < culs_val(x)))) {
            A[j + 1] = A[j];
            LV[j + 1] = LV[j];
            j--;
        }
        A[j + 1] = x;
        LV[j + 1] = xl;
    }
}

// sharedNodeDrivenPermute then prepareDataToInsert, one cover; returns mask
__device__ uint32_t rec_prepare(const RecArgs &r, uint32_t c, uint32_t counter) {
    uint32_t *A = r.arr + (size_t)c * r.maxLen, *LV = r.alv + (size_t)c * r.maxLen;
    const uint32_t L = r.alen[c];
    if (L <= 1) { r.mask[c] = 0; return 0; }
    int left;
    if (L < 3) {
        left = 0;
    } else {
        left = (int)L - 2;
        const uint32_t border = LV[left];
        for (left--; left >= 0; left--)
            if (LV[left] != border) break;
        left++;
    }
    const int right = (int)L - 2;
    if (left < right) {
        const uint32_t l1 = A[right + 1], l2 = A[right];
        for (int i = right; i >= left; i--) {
            const uint32_t l3 = A[i];
            if (and_lookup(r.t, l1, l3) != UNSET) {
                if (l3 != l2) {
                    A[i] = l2;
                    A[right] = l3;
                    const uint32_t tl = LV[i];
                    LV[i] = LV[right];
                    LV[right] = tl;
                }
                break;
            }
        }
    }
    uint32_t l1 = A[L - 1], l2 = A[L - 2];
    const uint32_t v1 = LV[L - 1], v2 = LV[L - 2];
    r.alen[c] = L - 2;
    const uint32_t found = and_lookup(r.t, l1, l2);
    if (found != UNSET) {
        r.mask[c] = 2;
        r.sval[c] = found;
        r.slev[c] = lit_level(r.t, found, r.nPIs);
        return 2;
    }
    if (l1 > l2) { const uint32_t x = l1; l1 = l2; l2 = x; }
    r.mask[c] = 1;
    r.skey[c] = (unsigned long long)l1 << 32 | l2;
    r.sval[c] = counter + c;
    r.slev[c] = (v1 > v2 ? v1 : v2) + 1u;
    return 1;
}

// addBackLocalArrays (balance.cu:429-467), one cover
__device__ void rec_add_back(const RecArgs &r, uint32_t c) {
    const uint8_t m = r.mask[c];
    if (m == 0) return;
    uint32_t *A = r.arr + (size_t)c * r.maxLen, *LV = r.alv + (size_t)c * r.maxLen;
    const uint32_t lit = m == 2 ? r.sval[c] : 2u * r.sval[c];
    const uint32_t L = r.alen[c];
    for (uint32_t i = 0; i < L; i++)
        if (A[i] == lit) return;
    A[L] = lit;
    LV[L] = r.slev[c];
    r.alen[c] = L + 1;
    for (uint32_t i = L; i > 0; i--) {
        if (LV[i] <= LV[i - 1]) break;
        const uint32_t x = A[i];
        A[i] = A[i - 1];
        A[i - 1] = x;
        const uint32_t y = LV[i];
        LV[i] = LV[i - 1];
        LV[i - 1] = y;
    }
}

__device__ __forceinline__ void rec_record(const RecArgs &r, uint32_t c, uint32_t root) {
    r.newLit[root] = r.alen[c] == 0 ? 0u : r.arr[(size_t)c * r.maxLen];
}

// The persistent workgroup: levels [lFirst .. lLast] (descending), every one
// at most balance_narrow_max covers wide; starts[l] = first index of level l
// in roots.
__global__ void __launch_bounds__(BAL_WG) k_rec_narrow(RecArgs r, const uint32_t *__restrict__ starts, int32_t lFirst,
                                                       int32_t lLast) {
    __shared__ uint32_t s_any, s_max, s_counter;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_counter = __hip_atomic_load(&r.counter[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int32_t l = lFirst; l >= lLast; l--) {
        const uint32_t s = starts[l], nb = starts[l + 1] - s;
        for (uint32_t c = tid; c < nb; c += BAL_WG) rec_gather(r, c, r.roots[s + c]);
        __syncthreads();
        for (;;) {
            if (tid == 0) { s_any = 0; s_max = 0; }
            __syncthreads();
            const uint32_t counter = s_counter;
            bool any = false;
            for (uint32_t c = tid; c < nb; c += BAL_WG) any |= rec_prepare(r, c, counter) != 0;
            if (any) atomicOr(&s_any, 1u);
            __syncthreads();
            if (!s_any) break;
            for (uint32_t c = tid; c < nb; c += BAL_WG)
                if (r.mask[c] == 1 && !r_insert_min(r.t, r.skey[c], r.sval[c])) atomicOr(r.err, 2u);
            __syncthreads();
            uint32_t mx = 0;
            for (uint32_t c = tid; c < nb; c += BAL_WG) {
                if (r.mask[c] != 1) continue;
                const uint32_t got = r_find(r.t, r.skey[c]);
                if (got == r.sval[c]) atomicAdd(&r.counter[3], 1u);
                r.sval[c] = got;
                mx = got > mx ? got : mx;
            }
            if (mx) atomicMax(&s_max, mx);
            __syncthreads();
            if (tid == 0 && s_max) s_counter = s_max + 1u;
            for (uint32_t c = tid; c < nb; c += BAL_WG) {
                if (r.mask[c] == 1 && !l_insert(r.t, r.sval[c], r.slev[c])) atomicOr(r.err, 2u);
                rec_tag(r, c, r.roots[s + c]);
                rec_add_back(r, c);
            }
            __syncthreads();
        }
        for (uint32_t c = tid; c < nb; c += BAL_WG) rec_record(r, c, r.roots[s + c]);
        __syncthreads();
    }
    if (tid == 0) __hip_atomic_store(&r.counter[0], s_counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// wide levels: one kernel per phase
__global__ void k_rec_gather(RecArgs r, uint32_t s, uint32_t nb) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nb; c += stride)
        rec_gather(r, (uint32_t)c, r.roots[s + c]);
}

__global__ void k_rec_prepare(RecArgs r, uint32_t nb) {
    const uint32_t counter = r.counter[0];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    bool any = false;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nb; c += stride)
        any |= rec_prepare(r, (uint32_t)c, counter) != 0;
    if (any) atomicOr(&r.counter[2], 1u);
}

__global__ void k_rec_insert(RecArgs r, uint32_t nb) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nb; c += stride)
        if (r.mask[c] == 1 && !r_insert_min(r.t, r.skey[c], r.sval[c])) atomicOr(r.err, 2u);
}

__global__ void k_rec_retrieve(RecArgs r, uint32_t nb) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    uint32_t mx = 0, made = 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nb; c += stride) {
        if (r.mask[c] != 1) continue;
        const uint32_t got = r_find(r.t, r.skey[c]);
        made += got == r.sval[c] ? 1u : 0u;
        r.sval[c] = got;
        mx = got > mx ? got : mx;
    }
    // one atomic per workgroup for each counter
    __shared__ uint32_t s_made, s_mx;
    if (threadIdx.x == 0) { s_made = 0; s_mx = 0; }
    __syncthreads();
    if (made) atomicAdd(&s_made, made);
    if (mx) atomicMax(&s_mx, mx);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_made) atomicAdd(&r.counter[3], s_made);
        if (s_mx) atomicMax(&r.counter[1], s_mx);
    }
}

__global__ void k_rec_add_back(RecArgs r, uint32_t s, uint32_t nb) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nb; c += stride) {
        if (r.mask[c] == 1 && !l_insert(r.t, r.sval[c], r.slev[c])) atomicOr(r.err, 2u);
        rec_tag(r, (uint32_t)c, r.roots[s + c]);
        rec_add_back(r, (uint32_t)c);
    }
}

// counter[0] = step max + 1 when a node was made; clears the step fields
__global__ void k_rec_step_end(uint32_t *counter) {
    if (counter[1]) counter[0] = counter[1] + 1u;
    counter[1] = 0;
    counter[2] = 0;
}

__global__ void k_rec_record(RecArgs r, uint32_t s, uint32_t nb) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nb; c += stride)
        rec_record(r, (uint32_t)c, r.roots[s + c]);
}

// ---------------------------------------------------------------------------
// phase 3: renumbering
// ---------------------------------------------------------------------------
__global__ void k_created(uint64_t cap, const unsigned long long *__restrict__ rk, const uint32_t *__restrict__ rv,
                          unsigned long long *keys, uint32_t *ids, uint32_t *count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < (int64_t)cap; s += stride) {
        const bool full = rk[s] != EMPTY64;
        const uint32_t k = wave_append(count, full);
        if (!full) continue;
        keys[k] = rk[s];
        ids[k] = rv[s];
    }
}

// ids sorted ascending: created node q gets output id nPIs + 1 + q
__global__ void k_out_map(int32_t m, int32_t nPIs, const uint32_t *__restrict__ sortedIds, Tables t, uint32_t *err) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < m; q += stride)
        if (!l_insert(t, sortedIds[q], (uint32_t)(nPIs + 1 + q))) atomicOr(err, 2u);
}

__device__ __forceinline__ uint32_t out_lit(const Tables &outMap, uint32_t lit, int32_t nPIs) {
    const uint32_t id = lit >> 1;
    if ((int32_t)id <= nPIs) return lit;
    return 2u * l_find(outMap, id) | (lit & 1u);
}

__global__ void k_out_emit(int32_t m, int32_t nPIs, const unsigned long long *__restrict__ sortedKeys,
                           const uint32_t *__restrict__ sortedIds, Tables outMap, uint32_t *of0, uint32_t *of1,
                           uint32_t *olevel, const uint32_t *__restrict__ tagIn, Tables tagTab, uint32_t *tagOut) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < m; q += stride) {
        const unsigned long long k = sortedKeys[q];
        const int64_t o = nPIs + 1 + q;
        of0[o] = out_lit(outMap, (uint32_t)(k >> 32), nPIs);
        of1[o] = out_lit(outMap, (uint32_t)k, nPIs);
        olevel[o] = 0;
        if (tagOut) tagOut[o] = l_find(tagTab, sortedIds[q]);
    }
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v <= nPIs; v += stride) {
        of0[v] = 0;
        of1[v] = 0;
        olevel[v] = 0;
        if (tagOut) tagOut[v] = tagIn[v];
    }
}

__global__ void k_out_pos(int32_t nPOs, int32_t nPIs, const uint32_t *__restrict__ pos,
                          const uint32_t *__restrict__ newLit, Tables outMap, uint32_t *opos) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nPOs; i += stride) {
        const uint32_t l = pos[i], id = l >> 1;
        if ((int32_t)id <= nPIs) { opos[i] = l; continue; }
        const uint32_t nl = newLit[id];
        opos[i] = out_lit(outMap, nl, nPIs) ^ (l & 1u);
    }
}

// (level << 32 | tag) of every root in level order, for the batch-split check
__global__ void k_level_tag_keys(uint32_t nRoots, const uint32_t *__restrict__ levels,
                                 const uint32_t *__restrict__ roots, const uint32_t *__restrict__ tag,
                                 unsigned long long *keys) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nRoots; i += stride)
        keys[i] = (unsigned long long)levels[i] << 32 | tag[roots[i]];
}

// In sorted (level, tag) keys, a key equal to the one BAL_BATCH places
// before it belongs to a tag with more than BAL_BATCH covers on that level:
// its tag goes to the list (at most cap entries; *count counts them all).
__global__ void k_split_tags(uint32_t nRoots, const unsigned long long *__restrict__ sortedKeys, uint32_t *list,
                             uint32_t cap, uint32_t *count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)BAL_BATCH + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nRoots; i += stride) {
        const unsigned long long k = sortedKeys[i];
        // the first place of each offending run past the limit reports it
        if (k == sortedKeys[i - BAL_BATCH] && (i == BAL_BATCH || sortedKeys[i - BAL_BATCH - 1] != k)) {
            const uint32_t q = atomicAdd(count, 1u);
            if (q < cap) list[q] = (uint32_t)k;
        }
    }
}

// the largest set cover length (at least 1)
__global__ void k_max_cover_len(int32_t n, const uint32_t *__restrict__ coverLen, uint32_t *out) {
    __shared__ uint32_t s_m;
    if (threadIdx.x == 0) s_m = 1;
    __syncthreads();
    uint32_t m = 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += stride) {
        const uint32_t l = coverLen[v];
        if (l != UNSET && l > m) m = l;
    }
    atomicMax(&s_m, m);
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, s_m);
}

// steps of every reconstruction batch of the wide levels: the largest
// (cover length - 1) of its roots; batchOff[l] = first batch of level l
// (UNSET for a level that the persistent workgroup takes)
__global__ void k_batch_steps(uint32_t nRoots, const uint32_t *__restrict__ levels,
                              const uint32_t *__restrict__ roots, const uint32_t *__restrict__ coverLen,
                              const uint32_t *__restrict__ starts, const uint32_t *__restrict__ batchOff,
                              uint32_t batchLimit, uint32_t *steps) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nRoots; i += stride) {
        const uint32_t l = levels[i], bo = batchOff[l];
        if (bo == UNSET) continue;
        const uint32_t len = coverLen[roots[i]];
        const uint32_t st = len > 1 ? len - 1 : 0u;
        // most roots of a batch see a maximum at least as large already: the
        // atomic is taken only by those that raise it
        uint32_t *dst = &steps[bo + ((uint32_t)i - starts[l]) / batchLimit];
        if (st > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(dst, st);
    }
}

struct Tmp {
    gls_ctx *ctx;
    std::vector<void *> ptrs;
    bool ok = true;
    template <class T> T *get(size_t n) {
        T *p = pool_get<T>(ctx, std::max<size_t>(n, 1));
        if (p) ptrs.push_back(p);
        else ok = false;
        return p;
    }
    ~Tmp() {
        for (void *p : ptrs) pool_put(ctx, p);
    }
};

uint32_t bits_of(uint32_t x) {
    uint32_t b = 1;
    while (b < 32 && (x >> b) != 0) ++b;
    return b;
}

} // namespace

int dev_balance(gls_ctx *ctx, const DevGraph &in, DevGraph &out, gls_op_info *info, const uint32_t *tagIn,
                uint32_t **tagOut, std::vector<uint32_t> *splitTags, int sortDecId) {
    const bool tagged = tagIn && tagOut;
    if (sortDecId != 0 && sortDecId != 1) return gls_fail(ctx, GLS_ERR_ARG, "balance: sortDecId must be 0 or 1");
    // the shuffle is seeded with a cover's index in its batch, which in a
    // batched arena is not its index when balanced alone
    if (sortDecId == 0 && tagged) return gls_fail(ctx, GLS_ERR_ARG, "balance: `b -s` is not batched");
    const int32_t n = in.nObjs, nPIs = in.nPIs, nPOs = in.nPOs;
    const uint32_t narrowMax = (uint32_t)tuning_for(ctx, n).balance_narrow_max;
    hipStream_t st = ctx->stream;
    Tmp t{ctx};
    const auto t0 = std::chrono::steady_clock::now();
    // ---- phase 1: covers and root levels
    uint32_t *nFo = t.get<uint32_t>(n), *coverOff = t.get<uint32_t>(n), *coverLen = t.get<uint32_t>(n);
    uint32_t *lastLevel = t.get<uint32_t>(n), *stamp = t.get<uint32_t>(n), *indeg = t.get<uint32_t>(n);
    uint32_t *frontA = t.get<uint32_t>(n), *frontB = t.get<uint32_t>(n);
    uint32_t *ctr = t.get<uint32_t>(8);   // [1] pool next, [2] err, [4..7] reconstruction counter
    uint32_t *flen = t.get<uint32_t>((size_t)n + 2 + COVER_LEVELS_PER_CHECK);   // frontier size per level
    const uint64_t poolCap = 2ull * (uint64_t)n + 64;   // cover inputs are distinct fanin edges
    uint32_t *pool = t.get<uint32_t>(poolCap);
    // one cover-discovery thread per 8 objects within the bounds: a wide
    // level's frontier is then walked in one or two rounds per thread
    const int coverThreads = (int)std::min<int64_t>(
        COVER_THREADS_MAX, std::max<int64_t>(COVER_THREADS, ((int64_t)n / 8 + kB - 1) / kB * kB));
    uint32_t *scratch = t.get<uint32_t>((size_t)coverThreads * COVER_SCRATCH);
    if (!t.ok) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: allocation (n=%d)", n);
    GLS_HIP(ctx, hipMemsetAsync(nFo, 0, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(coverLen, 0xFF, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(lastLevel, 0xFF, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(stamp, 0, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(indeg, 0, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(ctr, 0, sizeof(uint32_t) * 8, st));
    GLS_HIP(ctx, hipMemsetAsync(flen, 0, sizeof(uint32_t) * ((size_t)n + 2 + COVER_LEVELS_PER_CHECK), st));
    k_fanouts<<<nblocks(n), kB, 0, st>>>(n, nPIs, in.f0, in.f1, nPOs, in.pos, nFo);
    if (nPOs > 0) k_po_frontier<<<nblocks(nPOs), kB, 0, st>>>(nPOs, nPIs, in.pos, stamp, frontA, &flen[0]);
    GLS_HIP(ctx, hipGetLastError());
    CoverArgs ca{in.f0, in.f1, nFo, nPIs, coverOff, coverLen, pool, &ctr[1], poolCap, scratch, lastLevel, stamp,
                 indeg, &ctr[2]};
    // rounds are launched COVER_LEVELS_PER_CHECK at a time; a launch past the
    // last round finds an empty frontier and does nothing.  Returns the
    // number of rounds with a non-empty frontier.
    auto sweep = [&](bool kahn, uint32_t &rounds) -> int {
        for (uint32_t base = 0;; base += COVER_LEVELS_PER_CHECK) {
            if ((uint64_t)base > (uint64_t)n) return gls_fail(ctx, GLS_ERR_INTERNAL, "balance: cover levels do not end");
            for (uint32_t r = base; r < base + COVER_LEVELS_PER_CHECK; r++) {
                if (kahn) k_cover_kahn<<<coverThreads / kB, kB, 0, st>>>(ca, frontA, frontB, flen, r);
                else k_cover_reach<<<coverThreads / kB, kB, 0, st>>>(ca, frontA, frontB, flen, r);
            }
            GLS_HIP(ctx, hipGetLastError());
            uint32_t h[COVER_LEVELS_PER_CHECK + 1];
            GLS_TRY(stream_copy(ctx, h, &flen[base], sizeof(h), hipMemcpyDeviceToHost));
            for (int i = 0; i <= COVER_LEVELS_PER_CHECK; i++)
                if (h[i] == 0) { rounds = base + (uint32_t)i; return GLS_OK; }
        }
    };
    uint32_t reachRounds = 0, levels = 0;
    GLS_TRY(sweep(false, reachRounds));
    // every reached root now has its cover and its count of parent covers
    GLS_HIP(ctx, hipMemsetAsync(flen, 0, sizeof(uint32_t) * ((size_t)reachRounds + 2), st));
    if (nPOs > 0) k_kahn_first<<<nblocks(nPOs), kB, 0, st>>>(nPOs, nPIs, in.pos, indeg, stamp, frontA, &flen[0]);
    GLS_HIP(ctx, hipGetLastError());
    GLS_TRY(sweep(true, levels));
    uint32_t herr = 0, poolUsed = 0;
    GLS_TRY(stream_copy(ctx, &herr, &ctr[2], sizeof(uint32_t), hipMemcpyDeviceToHost));
    GLS_TRY(stream_copy(ctx, &poolUsed, &ctr[1], sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (herr) return gls_fail(ctx, GLS_ERR_INTERNAL, "balance: cover pool overflow");

    // roots by (level, id)
    uint8_t *flag = t.get<uint8_t>(n);
    uint32_t *roots = t.get<uint32_t>(n), *rootsSorted = t.get<uint32_t>(n);
    uint32_t *keys = t.get<uint32_t>(n), *keysSorted = t.get<uint32_t>(n), *d_cnt = t.get<uint32_t>(2);
    if (!t.ok) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: allocation");
    k_root_flags<<<nblocks(n), kB, 0, st>>>(n, lastLevel, flag);
    size_t tb = 0;
    rocprim::counting_iterator<uint32_t> ids(0u);
    GLS_HIP(ctx, rocprim::select(nullptr, tb, ids, flag, roots, d_cnt, (size_t)n, st));
    void *tmp = t.get<uint8_t>(tb);
    if (!tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: select storage");
    GLS_HIP(ctx, rocprim::select(tmp, tb, ids, flag, roots, d_cnt, (size_t)n, st));
    uint32_t nRoots = 0;
    GLS_TRY(stream_copy(ctx, &nRoots, d_cnt, sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> starts(levels + 2, 0);
    if (nRoots > 0) {
        k_gather_level<<<nblocks(nRoots), kB, 0, st>>>((int32_t)nRoots, roots, lastLevel, keys);
        size_t tb2 = 0;
        const uint32_t eb = bits_of(levels);
        GLS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb2, keys, keysSorted, roots, rootsSorted, (size_t)nRoots, 0u,
                                               eb, st));
        void *tmp2 = t.get<uint8_t>(tb2);
        if (!tmp2) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: sort storage");
        GLS_HIP(ctx, rocprim::radix_sort_pairs(tmp2, tb2, keys, keysSorted, roots, rootsSorted, (size_t)nRoots, 0u,
                                               eb, st));
        std::vector<uint32_t> hk(nRoots);
        GLS_TRY(stream_copy(ctx, hk.data(), keysSorted, sizeof(uint32_t) * nRoots, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nRoots; i++) starts[hk[i] + 1]++;
        for (uint32_t l = 1; l <= levels + 1; l++) starts[l] += starts[l - 1];
    }
    // Batches: 32768 covers of a level at a time.  With tags (a union of
    // independent AIGs) a level is one batch, which gives every AIG the batches
    // it gets alone as long as none has more than 32768 covers on a level: the
    // (level, tag) keys of the roots are sorted on the device, and a tag with
    // a longer run is reported to the caller (splitTags), which balances it
    // alone.
    if (tagged && nRoots > (uint32_t)BAL_BATCH) {
        unsigned long long *gk = t.get<unsigned long long>(nRoots), *gkS = t.get<unsigned long long>(nRoots);
        uint32_t *lst = t.get<uint32_t>(1024), *lcnt = t.get<uint32_t>(1);
        if (!t.ok) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: split check");
        k_level_tag_keys<<<nblocks(nRoots), kB, 0, st>>>(nRoots, keysSorted, rootsSorted, tagIn, gk);
        size_t tbk = 0;
        GLS_HIP(ctx, rocprim::radix_sort_keys(nullptr, tbk, gk, gkS, (size_t)nRoots, 0u, 64u, st));
        void *tmpk = t.get<uint8_t>(tbk);
        if (!tmpk) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: split check sort");
        GLS_HIP(ctx, rocprim::radix_sort_keys(tmpk, tbk, gk, gkS, (size_t)nRoots, 0u, 64u, st));
        GLS_HIP(ctx, hipMemsetAsync(lcnt, 0, sizeof(uint32_t), st));
        k_split_tags<<<nblocks(nRoots), kB, 0, st>>>(nRoots, gkS, lst, 1024u, lcnt);
        GLS_HIP(ctx, hipGetLastError());
        uint32_t nsplit = 0;
        GLS_TRY(stream_copy(ctx, &nsplit, lcnt, sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (nsplit > 0) {
            if (splitTags) {
                splitTags->assign(std::min<uint32_t>(nsplit, 1024u), 0u);
                GLS_TRY(stream_copy(ctx, splitTags->data(), lst, sizeof(uint32_t) * splitTags->size(),
                                    hipMemcpyDeviceToHost));
                if (nsplit > 1024u) splitTags->clear();   // too many to list: the caller splits them all
            }
            return GLS_ERR_BATCH_SPLIT;
        }
    }
    const uint32_t batchLimit = tagged ? 0xFFFFFFFFu : (uint32_t)BAL_BATCH;
    // the persistent workgroup takes a whole level as one batch, so a level
    // goes there only if it is not wider than a batch
    const uint32_t narrowCap = std::min(narrowMax, batchLimit);
    // cover lengths: the workspace width, and the steps a batch needs (a step
    // takes two inputs of every unfinished cover and gives back at most one),
    // reduced on the device
    uint32_t maxLen = 1;
    {
        uint32_t *d_ml = t.get<uint32_t>(1);
        if (!d_ml) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: cover length");
        GLS_HIP(ctx, hipMemsetAsync(d_ml, 0, sizeof(uint32_t), st));
        k_max_cover_len<<<std::min<int64_t>(nblocks(n), 1024), kB, 0, st>>>(n, coverLen, d_ml);
        GLS_HIP(ctx, hipGetLastError());
        GLS_TRY(stream_copy(ctx, &maxLen, d_ml, sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    // the wide levels' batches and their step counts
    std::vector<uint32_t> batchOff(levels + 1, UNSET), batchSteps;
    uint32_t nBatches = 0;
    for (uint32_t l = 0; l < levels; l++) {
        const uint32_t cnt = starts[l + 1] - starts[l];
        if (cnt <= narrowCap) continue;
        batchOff[l] = nBatches;
        nBatches += (cnt + std::min(cnt, batchLimit) - 1) / std::min(cnt, batchLimit);
    }
    if (nBatches > 0) {
        uint32_t *d_off = t.get<uint32_t>(levels + 1), *d_st = t.get<uint32_t>(levels + 2);
        uint32_t *d_steps = t.get<uint32_t>(nBatches);
        if (!d_off || !d_st || !d_steps) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: batch steps");
        GLS_TRY(stream_copy(ctx, d_off, batchOff.data(), sizeof(uint32_t) * (levels + 1), hipMemcpyHostToDevice));
        GLS_TRY(stream_copy(ctx, d_st, starts.data(), sizeof(uint32_t) * (levels + 2), hipMemcpyHostToDevice));
        GLS_HIP(ctx, hipMemsetAsync(d_steps, 0, sizeof(uint32_t) * nBatches, st));
        k_batch_steps<<<nblocks(nRoots), kB, 0, st>>>(nRoots, keysSorted, rootsSorted, coverLen, d_st, d_off,
                                                      batchLimit, d_steps);
        GLS_HIP(ctx, hipGetLastError());
        batchSteps.resize(nBatches);
        GLS_TRY(stream_copy(ctx, batchSteps.data(), d_steps, sizeof(uint32_t) * nBatches, hipMemcpyDeviceToHost));
    }

    // ---- phase 2: reconstruction
    uint64_t cap = 1024;
    while (cap < 2ull * ((uint64_t)poolUsed + 16)) cap <<= 1;
    unsigned long long *rk = t.get<unsigned long long>(cap);
    uint32_t *rv = t.get<uint32_t>(cap), *lk = t.get<uint32_t>(cap), *lv = t.get<uint32_t>(cap);
    uint32_t *newLit = t.get<uint32_t>(n);
    size_t maxBatch = 1;
    for (uint32_t l = 0; l < levels; l++)
        maxBatch = std::max<size_t>(maxBatch, std::min<uint32_t>(starts[l + 1] - starts[l], batchLimit));
    uint32_t *arr = t.get<uint32_t>(maxBatch * maxLen), *alv = t.get<uint32_t>(maxBatch * maxLen);
    uint32_t *alen = t.get<uint32_t>(maxBatch);
    uint8_t *mask = t.get<uint8_t>(maxBatch);
    unsigned long long *skey = t.get<unsigned long long>(maxBatch);
    uint32_t *sval = t.get<uint32_t>(maxBatch), *slev = t.get<uint32_t>(maxBatch);
    uint32_t *d_starts = t.get<uint32_t>(levels + 2);
    if (!t.ok) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: reconstruction allocation");
    GLS_HIP(ctx, hipMemsetAsync(rk, 0xFF, sizeof(unsigned long long) * cap, st));
    GLS_HIP(ctx, hipMemsetAsync(rv, 0xFF, sizeof(uint32_t) * cap, st));
    GLS_HIP(ctx, hipMemsetAsync(lk, 0xFF, sizeof(uint32_t) * cap, st));
    GLS_HIP(ctx, hipMemsetAsync(lv, 0xFF, sizeof(uint32_t) * cap, st));
    GLS_HIP(ctx, hipMemsetAsync(newLit, 0xFF, sizeof(uint32_t) * n, st));
    GLS_TRY(stream_copy(ctx, d_starts, starts.data(), sizeof(uint32_t) * (levels + 2), hipMemcpyHostToDevice));
    const uint32_t firstId = (uint32_t)nPIs + 1u;
    GLS_TRY(stream_copy(ctx, &ctr[4], &firstId, sizeof(uint32_t), hipMemcpyHostToDevice));
    Tables tabs{rk, rv, cap - 1, lk, lv, cap - 1};
    Tables tagTab{nullptr, nullptr, 0, nullptr, nullptr, 0};
    if (tagged) {
        uint32_t *tk = t.get<uint32_t>(cap), *tv = t.get<uint32_t>(cap);
        if (!tk || !tv) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: tag table");
        GLS_HIP(ctx, hipMemsetAsync(tk, 0xFF, sizeof(uint32_t) * cap, st));
        GLS_HIP(ctx, hipMemsetAsync(tv, 0xFF, sizeof(uint32_t) * cap, st));
        tagTab = Tables{nullptr, nullptr, 0, tk, tv, cap - 1};
    }
    RecArgs ra{tabs, nPIs, rootsSorted, coverOff, coverLen, pool, newLit, arr, alv, alen, maxLen, mask, skey,
               sval, slev, &ctr[4], &ctr[2], tagged ? tagIn : nullptr, tagTab, sortDecId == 0 ? 1u : 0u};
    int32_t l = (int32_t)levels - 1;
    while (l >= 0) {
        // a run of narrow levels goes to the persistent workgroup
        int32_t e = l;
        while (e >= 0 && starts[e + 1] - starts[e] <= narrowCap) e--;
        if (e < l) {
            k_rec_narrow<<<1, BAL_WG, 0, st>>>(ra, d_starts, l, e + 1);
            GLS_HIP(ctx, hipGetLastError());
            l = e;
            continue;
        }
        // one wide level, in batches, one kernel per phase
        uint32_t bi = batchOff[l];
        for (uint32_t b0 = starts[l]; b0 < starts[l + 1]; b0 += std::min<uint32_t>(starts[l + 1] - b0, batchLimit)) {
            const uint32_t nb = std::min<uint32_t>(starts[l + 1] - b0, batchLimit);
            const uint32_t steps = batchSteps[bi++];
            k_rec_gather<<<nblocks(nb), kB, 0, st>>>(ra, b0, nb);
            for (uint32_t stp = 0; stp < steps; stp++) {
                k_rec_prepare<<<nblocks(nb), kB, 0, st>>>(ra, nb);
                k_rec_insert<<<nblocks(nb), kB, 0, st>>>(ra, nb);
                k_rec_retrieve<<<nblocks(nb), kB, 0, st>>>(ra, nb);
                k_rec_add_back<<<nblocks(nb), kB, 0, st>>>(ra, b0, nb);
                k_rec_step_end<<<1, 1, 0, st>>>(&ctr[4]);
                GLS_HIP(ctx, hipGetLastError());
            }
            k_rec_record<<<nblocks(nb), kB, 0, st>>>(ra, b0, nb);
            GLS_HIP(ctx, hipGetLastError());
        }
        l--;
    }
    GLS_TRY(stream_copy(ctx, &herr, &ctr[2], sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (herr) return gls_fail(ctx, GLS_ERR_INTERNAL, "balance: table overflow (%u)", herr);

    // ---- phase 3: created nodes in id order
    uint32_t created = 0;
    GLS_TRY(stream_copy(ctx, &created, &ctr[7], sizeof(uint32_t), hipMemcpyDeviceToHost));
    unsigned long long *ckeys = t.get<unsigned long long>(created), *ckeysS = t.get<unsigned long long>(created);
    uint32_t *cids = t.get<uint32_t>(created), *cidsS = t.get<uint32_t>(created);
    uint32_t *ok = t.get<uint32_t>(cap), *ov = t.get<uint32_t>(cap);
    if (!t.ok) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: output allocation");
    GLS_HIP(ctx, hipMemsetAsync(d_cnt, 0, sizeof(uint32_t), st));
    k_created<<<nblocks((int64_t)cap), kB, 0, st>>>(cap, rk, rv, ckeys, cids, d_cnt);
    uint32_t m = 0;
    GLS_TRY(stream_copy(ctx, &m, d_cnt, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (m != created) return gls_fail(ctx, GLS_ERR_INTERNAL, "balance: %u nodes in the table, %u created", m, created);
    if (m > 0) {
        size_t tb3 = 0;
        GLS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb3, cids, cidsS, ckeys, ckeysS, (size_t)m, 0u, 32u, st));
        void *tmp3 = t.get<uint8_t>(tb3);
        if (!tmp3) return gls_fail(ctx, GLS_ERR_NOMEM, "balance: sort storage");
        GLS_HIP(ctx, rocprim::radix_sort_pairs(tmp3, tb3, cids, cidsS, ckeys, ckeysS, (size_t)m, 0u, 32u, st));
    }
    GLS_HIP(ctx, hipMemsetAsync(ok, 0xFF, sizeof(uint32_t) * cap, st));
    GLS_HIP(ctx, hipMemsetAsync(ov, 0xFF, sizeof(uint32_t) * cap, st));
    Tables outMap{nullptr, nullptr, 0, ok, ov, cap - 1};
    if (m > 0) k_out_map<<<nblocks(m), kB, 0, st>>>((int32_t)m, nPIs, cidsS, outMap, &ctr[2]);
    if ((int64_t)nPIs + 1 + m > 0x7FFFFFF0LL) return gls_fail(ctx, GLS_ERR_ARG, "balance: result too large");
    GLS_TRY(dev_graph_alloc(ctx, out, nPIs + 1 + (int32_t)m, nPIs, nPOs));
    uint32_t *tags = nullptr;
    if (tagged) {
        tags = pool_get<uint32_t>(ctx, (size_t)nPIs + 1 + m);
        if (!tags) { dev_graph_free(ctx, out); return gls_fail(ctx, GLS_ERR_NOMEM, "balance: tags"); }
        *tagOut = tags;
    }
    k_out_emit<<<nblocks(std::max<int64_t>(m, nPIs + 1)), kB, 0, st>>>((int32_t)m, nPIs, ckeysS, cidsS, outMap,
                                                                     out.f0, out.f1, out.level, tagIn, tagTab, tags);
    if (nPOs > 0) k_out_pos<<<nblocks(nPOs), kB, 0, st>>>(nPOs, nPIs, in.pos, newLit, outMap, out.pos);
    GLS_HIP(ctx, hipGetLastError());
    GLS_TRY(stream_copy(ctx, &herr, &ctr[2], sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (herr) { dev_graph_free(ctx, out); return gls_fail(ctx, GLS_ERR_INTERNAL, "balance: output map overflow"); }
    if (info) {
        info->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        info->levels_enumerated = levels;
        info->new_nodes = m;
        info->candidates = nRoots;
    }
    return GLS_OK;
}
