// Structural hashing on the device.
//
// Semantics follow the reference GPU strash (src/aig/strash.cu:190-314):
// dangling nodes (not in the transitive fanin of a PO) are removed, the
// remaining nodes are processed in level waves (a node's wave is its level in
// the dangling-free input), every AND is folded by the trivial rules
// (src/hash_table/hash_table.h:435-449) or merged with a structurally identical
// node, and the survivors are renumbered consecutively in (wave, old id) order.
// Where the reference keeps whichever duplicate won an atomicCAS
// (hash_table.h:478-486), this pass keeps the smallest (wave, old id): the
// table value is (wave << 32 | id) updated by atomicMin, so the result is the
// same on every run.
//
// Mapping to MI355X: the reference rescans every remaining node for readiness
// once per level (O(n * levels), strash.cu:247-277).  Here readiness is
// propagated along a fanout CSR (Kahn frontiers), so each node is touched once;
// one kernel per wave, launched in batches without host round trips (the last
// workgroup of a wave advances the frontier window on the device).
#include "gls_internal.h"
#include "dev_common.h"
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

namespace {

__global__ void k_count_refs(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0,
                             const uint32_t *__restrict__ f1, int32_t nPOs,
                             const uint32_t *__restrict__ pos, uint32_t *ref) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + tid; v < n; v += stride) {
        atomicAdd(&ref[LID(f0[v])], 1u);
        atomicAdd(&ref[LID(f1[v])], 1u);
    }
    for (int64_t i = tid; i < nPOs; i += stride) atomicAdd(&ref[LID(pos[i])], 1u);
}

// flag[v] = 1 for a live AND whose count is 0 (the first frontier of the
// dangling cascade or of the wave sweep); dead may be null
__global__ void k_zero_flags(int32_t n, int32_t nPIs, const uint32_t *__restrict__ cnt,
                             const uint8_t *__restrict__ dead, uint8_t *flag) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < n; v += stride) flag[v] = (v > nPIs && cnt[v] == 0 && !(dead && dead[v])) ? 1u : 0u;
}

// One round of the dangling cascade (strash.cu:175-188 as a frontier).
__global__ void k_dangle_round(uint32_t k, SweepCtl *ctl, int32_t nPIs, const uint32_t *__restrict__ f0,
                               const uint32_t *__restrict__ f1, uint32_t *ref, uint8_t *dead,
                               uint32_t *order) {
    uint32_t cs = ctl->curStart, cn = ctl->curSize;
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (uint32_t i = tid; i < cn; i += stride) {
        uint32_t v = order[cs + i];
        dead[v] = 1;
        uint32_t a = LID(f0[v]), b = LID(f1[v]);
        if (a > (uint32_t)nPIs && atomicSub(&ref[a], 1u) == 1u)
            order[cs + cn + atomicAdd(&ctl->nextSize, 1u)] = a;
        if (b > (uint32_t)nPIs && atomicSub(&ref[b], 1u) == 1u)
            order[cs + cn + atomicAdd(&ctl->nextSize, 1u)] = b;
    }
    sweep_round_end(ctl, k);
}

constexpr uint32_t PEND_ONE = 0x101u;

__global__ void k_kahn_count(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0,
                             const uint32_t *__restrict__ f1, const uint8_t *__restrict__ dead,
                             uint32_t *cnt, uint32_t *pending) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + tid; v < n; v += stride) {
        if (dead[v]) { pending[v] = 0; continue; }
        uint32_t a = LID(f0[v]), b = LID(f1[v]);
        uint32_t p = 0;
        if (a > (uint32_t)nPIs) { atomicAdd(&cnt[a], 1u); ++p; }
        if (b > (uint32_t)nPIs) { atomicAdd(&cnt[b], 1u); ++p; }
        // a node with one non-PI fanin carries PEND_ONE: the in-workgroup sweep
        // releases it without touching the counter, and the counter's low byte
        // is the count
        pending[v] = p == 1 ? PEND_ONE : p;
    }
}

// Fanout CSR: adj holds the fanout, adjY its other fanin (the fanin that is
// not the list's owner), with Y_OFL set when the fanout has more fanouts of
// its own than its node record holds (so its overflow record is read with it).
constexpr uint32_t Y_OFL = 0x80000000u;
constexpr uint32_t Y_ID = 0x7FFFFFFFu;
constexpr uint32_t REC_INLINE = 6;      // fanouts held in a node record

__global__ void k_kahn_fill(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0,
                            const uint32_t *__restrict__ f1, const uint8_t *__restrict__ dead,
                            const uint32_t *__restrict__ cnt, uint32_t *cursor, uint32_t *adj, uint32_t *adjY) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + tid; v < n; v += stride) {
        if (dead[v]) continue;
        uint32_t a = LID(f0[v]), b = LID(f1[v]);
        const uint32_t ofl = cnt[v] > REC_INLINE ? Y_OFL : 0u;
        if (a > (uint32_t)nPIs) {
            const uint32_t e = atomicAdd(&cursor[a], 1u);
            adj[e] = (uint32_t)v;
            adjY[e] = b | ofl;
        }
        if (b > (uint32_t)nPIs) {
            const uint32_t e = atomicAdd(&cursor[b], 1u);
            adj[e] = (uint32_t)v;
            adjY[e] = a | ofl;
        }
    }
}

// Pre-inserts the key of every live non-trivial AND as read from its own
// fanin literals; the slot's value field collects the smallest id with that
// key; slotTmp[v] = its slot (NO_SLOT for a trivial key).  k_canon then gives
// every such node the value location of its key: the entry val[m] of that
// smallest id m, so that the nodes of one wave, whose ids are close, update
// values on a few cache lines instead of one random table line each.  The
// table keeps the key (for probes by nodes whose fanins changed) with the
// value field set to VAL_MARK | m, which sends those probes to val[m] too.
constexpr uint32_t NO_SLOT = 0xFFFFFFFFu;
constexpr uint32_t LOC_HT = 0x80000000u;                 // location is a table slot
constexpr unsigned long long VAL_MARK = 0xFFFFFFFE00000000ull;

__global__ void k_preinsert(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0, const uint32_t *__restrict__ f1,
                            const uint8_t *__restrict__ dead, unsigned long long *ht, uint64_t mask, uint32_t *slotTmp,
                            uint32_t *err, unsigned long long *probes) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    uint32_t np = 0;
    for (int64_t v = nPIs + 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += stride) {
        if (dead[v]) continue;
        uint32_t l0 = f0[v], l1 = f1[v];
        if (l0 > l1) { const uint32_t t = l0; l0 = l1; l1 = t; }
        if (l0 == l1 || l0 == (l1 ^ 1u) || l0 < 2u) { slotTmp[v] = NO_SLOT; continue; }
        const uint64_t key = and_key(l0, l1);
        uint64_t s = mix64(key) & mask;
        uint32_t slot = NO_SLOT;
        for (uint64_t probe = 0; probe <= mask; ++probe) {
            unsigned long long prev = KEY_EMPTY;
            __hip_atomic_compare_exchange_strong(&ht[2 * s], &prev, (unsigned long long)key, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ++np;
            if (prev == KEY_EMPTY || prev == key) { slot = (uint32_t)s; break; }
            s = (s + 1) & mask;
        }
        if (slot == NO_SLOT) { atomicOr(err, 1u); continue; }
        atomicMin(&ht[2 * (size_t)slot + 1], (unsigned long long)v);
        slotTmp[v] = slot;
    }
    block_add_u64(probes, np);
}

__global__ void k_canon(int32_t n, int32_t nPIs, const uint8_t *__restrict__ dead, unsigned long long *ht,
                        const uint32_t *__restrict__ slotTmp, uint32_t *locOf) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += stride) {
        if (dead[v]) continue;
        const uint32_t slot = slotTmp[v];
        if (slot == NO_SLOT) { locOf[v] = NO_SLOT; continue; }
        // the low word is the smallest id before and after the owner marks it
        const uint32_t m = (uint32_t)__hip_atomic_load(&ht[2 * (size_t)slot + 1], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
        locOf[v] = m;
        if (m == (uint32_t)v)
            __hip_atomic_store(&ht[2 * (size_t)slot + 1], VAL_MARK | m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

struct WaveArgs {
    const uint32_t *f0, *f1;
    int32_t nPIs;
    unsigned long long *rl;         // per node: literal over representative ids | level << 32;
                                    // the literal is INVALID32 until the representative is known
    uint32_t *locOf, *wave;         // value location of non-trivial nodes, wave of every live node
    unsigned long long *ht;         // slot s: ht[2s] = key, ht[2s + 1] = value or VAL_MARK | val index
    unsigned long long *val;        // per-node values of pre-inserted keys: min (wave << 32 | id)
    uint64_t mask;
    const uint32_t *adjOff, *adj;
    uint32_t *pending, *order;
    uint32_t *fcnt, *fstart;        // size and start (in order[]) of every wave's frontier
    uint32_t *depth, *err;
    unsigned long long *probes;     // table slot inspections (key probes and value updates)
};

__device__ __forceinline__ unsigned long long pack_rl(uint32_t lit, uint32_t lev) {
    return (static_cast<unsigned long long>(lev) << 32) | lit;
}

// Value of a location: val[loc], or the table slot's value field for LOC_HT.
__device__ __forceinline__ unsigned long long *vptr(const WaveArgs &a, uint32_t loc) {
    return (loc & LOC_HT) ? &a.ht[2 * (size_t)(loc & ~LOC_HT) + 1] : &a.val[loc];
}

// Location of a key found or inserted by a probe at slot s.
template <int SCOPE>
__device__ __forceinline__ uint32_t loc_of_slot(const WaveArgs &a, uint64_t s) {
    const unsigned long long w = __hip_atomic_load(&a.ht[2 * s + 1], __ATOMIC_RELAXED, SCOPE);
    return (w & 0xFFFFFFFF00000000ull) == VAL_MARK ? (uint32_t)w : ((uint32_t)s | LOC_HT);
}

template <int SCOPE>
__device__ __forceinline__ uint32_t rep_of_loc(const WaveArgs &a, uint32_t loc) {
    unsigned long long v = __hip_atomic_load(vptr(a, loc), __ATOMIC_RELAXED, SCOPE);
    return 2u * (uint32_t)(v & 0xFFFFFFFFull);
}

// Fanin literal over representatives, and its level.  A fanin of the previous
// wave may still be unresolved; its table value is final by now.
template <int SCOPE>
__device__ __forceinline__ uint32_t map_fanin(const WaveArgs &a, uint32_t lit, uint32_t &lev) {
    const uint32_t u = LID(lit);
    if (u <= (uint32_t)a.nPIs) { lev = 0; return lit; }
    const unsigned long long x = a.rl[u];
    uint32_t r = (uint32_t)x;
    lev = (uint32_t)(x >> 32);
    if (r == INVALID32) r = rep_of_loc<SCOPE>(a, a.locOf[u]);
    return r ^ LC(lit);
}

// Hashes node v of wave k.  Trivial ANDs (hash_table.h:435-449 rules) are
// resolved at once; the others are inserted with value (k << 32 | v) kept by
// minimum, so the representative of duplicates is the smallest (wave, id).
// Returns true for a non-trivial node (slot and level filled, rl not yet).
template <int SCOPE>
__device__ __forceinline__ bool hash_node_lits(const WaveArgs &a, uint32_t k, uint32_t v, uint32_t raw0,
                                               uint32_t raw1, uint32_t &slotOut, uint32_t &levOut, uint32_t &np) {
    uint32_t la, lb;
    uint32_t l0 = map_fanin<SCOPE>(a, raw0, la), l1 = map_fanin<SCOPE>(a, raw1, lb);
    if (l0 > l1) { uint32_t t = l0; l0 = l1; l1 = t; t = la; la = lb; lb = t; }
    a.wave[v] = k;
    uint32_t r = INVALID32, rlev = 0;
    if (l0 == l1) { r = l0; rlev = la; }
    else if (l0 == (l1 ^ 1u)) r = 0;
    else if (l0 == 1u) { r = l1; rlev = lb; }
    else if (l0 == 0u) r = 0;
    if (r != INVALID32) {
        a.rl[v] = pack_rl(r, rlev);
        return false;
    }
    levOut = 1u + (la > lb ? la : lb);
    const uint64_t key = and_key(l0, l1), val = (static_cast<uint64_t>(k) << 32) | v;
    {
        // fanins that are their own representatives: the pre-inserted key holds
        const uint32_t r0 = raw0 < raw1 ? raw0 : raw1, r1 = raw0 < raw1 ? raw1 : raw0;
        const uint32_t pre = a.locOf[v];
        if (l0 == r0 && l1 == r1 && pre != NO_SLOT) {
            __hip_atomic_fetch_min(vptr(a, pre), (unsigned long long)val, __ATOMIC_RELAXED, SCOPE);
            ++np;
            slotOut = pre;
            return true;
        }
    }
    uint64_t s = mix64(key) & a.mask;
    uint32_t loc = NO_SLOT;
    for (uint64_t probe = 0; probe <= a.mask; ++probe) {
        unsigned long long prev = KEY_EMPTY;
        __hip_atomic_compare_exchange_strong(&a.ht[2 * s], &prev, (unsigned long long)key, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, SCOPE);
        ++np;
        if (prev == KEY_EMPTY || prev == key) {
            loc = loc_of_slot<SCOPE>(a, s);
            __hip_atomic_fetch_min(vptr(a, loc), (unsigned long long)val, __ATOMIC_RELAXED, SCOPE);
            break;
        }
        s = (s + 1) & a.mask;
    }
    if (loc == NO_SLOT) { atomicOr(a.err, 1u); loc = LOC_HT; }
    slotOut = loc;
    a.locOf[v] = loc;
    return true;
}

template <int SCOPE>
__device__ __forceinline__ bool hash_node(const WaveArgs &a, uint32_t k, uint32_t v, uint32_t &slotOut,
                                          uint32_t &levOut, uint32_t &np) {
    return hash_node_lits<SCOPE>(a, k, v, a.f0[v], a.f1[v], slotOut, levOut, np);
}

// Releases the fanouts of v whose last pending fanin is v.  Up to four
// decrements are in flight before any result is inspected.
template <int SCOPE, class Emit>
__device__ __forceinline__ void release_range(const WaveArgs &a, uint32_t e0, uint32_t ee, Emit emit) {
    for (uint32_t e = e0; e < ee; e += 4) {
        uint32_t w[4], r[4];
#pragma unroll
        for (int j = 0; j < 4; j++) w[j] = (e + j < ee) ? a.adj[e + j] : 0u;
#pragma unroll
        for (int j = 0; j < 4; j++)
            r[j] = (e + j < ee) ? __hip_atomic_fetch_add(&a.pending[w[j]], 0xFFFFFFFFu, __ATOMIC_RELAXED, SCOPE) : 0u;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (e + j < ee && (r[j] & 0xFFu) == 1u) emit(w[j]);
    }
}

template <int SCOPE, class Emit>
__device__ __forceinline__ void release(const WaveArgs &a, uint32_t v, Emit emit) {
    release_range<SCOPE>(a, a.adjOff[v], a.adjOff[v + 1], emit);
}

template <int SCOPE>
__device__ __forceinline__ void resolve(const WaveArgs &a, uint32_t v) {
    const unsigned long long x = a.rl[v];
    if ((uint32_t)x == INVALID32) a.rl[v] = pack_rl(rep_of_loc<SCOPE>(a, a.locOf[v]), (uint32_t)(x >> 32));
}

// One wave across the whole device (wide frontiers): items below pn resolve
// the previous wave, the rest hash this wave.
__global__ void k_wave_multi(uint32_t k, WaveArgs a) {
    const uint32_t pn = a.fcnt[k - 1], ps = a.fstart[k - 1], cn = a.fcnt[k], cs = a.fstart[k];
    const uint32_t nb = cs + cn;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    if (tid == 0) {
        a.fstart[k + 1] = nb;
        if (cn > 0) atomicMax(a.depth, k);
    }
    uint32_t np = 0;
    for (uint32_t i = tid; i < pn + cn; i += stride) {
        if (i < pn) {
            resolve<__HIP_MEMORY_SCOPE_AGENT>(a, a.order[ps + i]);
            continue;
        }
        const uint32_t v = a.order[cs + i - pn];
        uint32_t slot = 0, lev = 0;
        if (hash_node<__HIP_MEMORY_SCOPE_AGENT>(a, k, v, slot, lev, np)) a.rl[v] = pack_rl(INVALID32, lev);
        release<__HIP_MEMORY_SCOPE_AGENT>(a, v, [&](uint32_t w) { a.order[nb + atomicAdd(&a.fcnt[k + 1], 1u)] = w; });
    }
    block_add_u64(a.probes, np);
}

// Packed per-node record for the in-workgroup sweep, one 64-byte line:
// fanin literals, value location, fanout count, then the first REC_INLINE
// fanouts with their other fanin (adjY's word, flag included).  A node with
// more fanouts also has an overflow record: fanouts 6..11 and where its
// fanout list starts in adj / adjY (for the rest).
constexpr int REC_PAIRS = 6;
struct alignas(64) NodeRec {
    uint32_t f0, f1, loc, cnt;
    uint32_t wy[2 * REC_PAIRS];     // fanout j: wy[2j], its other fanin: wy[2j + 1]
};
struct alignas(64) NodeOfl {
    uint32_t wy[2 * REC_PAIRS];     // fanouts 6..11
    uint32_t start, pad0, pad1, pad2;
};
static_assert(sizeof(NodeRec) == 64 && sizeof(NodeOfl) == 64 && REC_PAIRS == (int)REC_INLINE,
              "64-byte records, read as four 16-byte loads");

__global__ void k_node_recs(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0,
                            const uint32_t *__restrict__ f1, const uint8_t *__restrict__ dead,
                            const uint32_t *__restrict__ locOf, const uint32_t *__restrict__ adjOff,
                            const uint32_t *__restrict__ adj, const uint32_t *__restrict__ adjY, NodeRec *recs,
                            NodeOfl *ofls) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += stride) {
        if (dead[v]) continue;
        NodeRec r;
        r.f0 = f0[v];
        r.f1 = f1[v];
        r.loc = locOf[v];
        const uint32_t e0 = adjOff[v];
        r.cnt = adjOff[v + 1] - e0;
#pragma unroll
        for (int j = 0; j < REC_PAIRS; j++) {
            r.wy[2 * j] = (uint32_t)j < r.cnt ? adj[e0 + j] : 0u;
            r.wy[2 * j + 1] = (uint32_t)j < r.cnt ? adjY[e0 + j] : 0u;
        }
        recs[v] = r;
        if (r.cnt > REC_INLINE) {
            NodeOfl o;
#pragma unroll
            for (int j = 0; j < REC_PAIRS; j++) {
                const uint32_t jj = REC_INLINE + (uint32_t)j;
                o.wy[2 * j] = jj < r.cnt ? adj[e0 + jj] : 0u;
                o.wy[2 * j + 1] = jj < r.cnt ? adjY[e0 + jj] : 0u;
            }
            o.start = e0;
            o.pad0 = o.pad1 = o.pad2 = 0;
            ofls[v] = o;
        }
    }
}

struct RecFull {
    uint32_t f0, f1, loc, cnt;
    uint4 c[3];                     // pairs 2k, 2k + 1 in c[k]
};
struct OflFull {
    uint4 c[3];                     // pairs 6 + 2k, 7 + 2k in c[k]
    uint32_t start;
};

__device__ __forceinline__ RecFull load_full(const NodeRec *recs, uint32_t v) {
    const uint4 *p = reinterpret_cast<const uint4 *>(recs + v);
    const uint4 h = p[0];
    RecFull r;
    r.f0 = h.x; r.f1 = h.y; r.loc = h.z; r.cnt = h.w;
#pragma unroll
    for (int k = 0; k < 3; k++) r.c[k] = p[1 + k];
    return r;
}

__device__ __forceinline__ OflFull load_ofl(const NodeOfl *ofls, uint32_t v) {
    const uint4 *p = reinterpret_cast<const uint4 *>(ofls + v);
    OflFull o;
#pragma unroll
    for (int k = 0; k < 3; k++) o.c[k] = p[k];
    o.start = p[3].x;
    return o;
}

// Appends a node's fanout edges (fanout, other fanin, source entry q) at
// position eb of an LDS edge list: fanouts 0..5 from its record, 6..11 from
// its overflow record o (read when the count exceeds 6), the rest from adj.
__device__ __forceinline__ void put_edges(uint32_t *gw, uint32_t *gy, uint16_t *gq, uint32_t eb, const RecFull &r,
                                          const OflFull &o, uint16_t q, const uint32_t *__restrict__ adj,
                                          const uint32_t *__restrict__ adjY) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (r.cnt > 2u * k) { gw[eb + 2 * k] = r.c[k].x; gy[eb + 2 * k] = r.c[k].y; gq[eb + 2 * k] = q; }
        if (r.cnt > 2u * k + 1) { gw[eb + 2 * k + 1] = r.c[k].z; gy[eb + 2 * k + 1] = r.c[k].w; gq[eb + 2 * k + 1] = q; }
    }
    if (r.cnt <= REC_INLINE) return;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint32_t j = REC_INLINE + 2u * k;
        if (r.cnt > j) { gw[eb + j] = o.c[k].x; gy[eb + j] = o.c[k].y; gq[eb + j] = q; }
        if (r.cnt > j + 1) { gw[eb + j + 1] = o.c[k].z; gy[eb + j + 1] = o.c[k].w; gq[eb + j + 1] = q; }
    }
    for (uint32_t j = 2 * REC_INLINE; j < r.cnt; j++) {
        gw[eb + j] = adj[o.start + j];
        gy[eb + j] = adjY[o.start + j];
        gq[eb + j] = q;
    }
}

// Consecutive narrow waves inside one workgroup, with every hand-over between
// waves kept in LDS so that a wave costs about two dependent trips to memory:
//   - the frontier entries (node, fanin literals, pre-inserted slot and where
//     each fanin's representative comes from) and the frontier's fanout edges
//     (fanout, its other fanin, source entry) are in LDS, written by the
//     thread that released the node;
//   - phase A: one thread per node hashes it (fanin representatives from LDS,
//     one returning atomic on the table) while, on other threads, one thread
//     per fanout edge decrements the fanout's pending count; a fanout that
//     becomes ready reads its whole NodeRec and the representative of its
//     other fanin in one trip and enters the next frontier (its other fanin is
//     resolved, or is in this wave, whose frontier position its rl then holds);
//   - phase B: each node's representative is the table minimum: a value from
//     an earlier wave, returned by its own atomic, is final; otherwise it is the
//     smallest id among this wave's nodes on the same slot, kept in an LDS table
//     whose keys carry the wave number (stale entries count as empty).
// rl of a node that has entered the frontier and is not hashed yet holds
// INVALID32 | position << 32.  Runs while the frontier fits (at most `thresh`
// <= WL_FM nodes, at most WL_EM edges); *kOut = next wave to run (== k0 when
// even the first frontier's edges do not fit).
constexpr int WG1 = 1024;
constexpr int WL_FM = 1024;     // frontier entries
constexpr int WL_EM = 2048;     // fanout edges of a frontier
constexpr int WL_DT_BITS = 10;  // same-wave duplicate table: 1024 entries (a wave inserts at most WL_FM keys)
constexpr int WL_DT = 1 << WL_DT_BITS;
constexpr unsigned long long SRC_POS = 1ull << 63;
static_assert(WL_FM <= WG1 && WL_EM <= 2 * WG1, "one node and two edges per thread");
static_assert(WL_FM <= 65536, "entry index in 16 bits");
// Fanout counters of a window of ids kept in LDS as one bit each: a node with
// two non-PI fanins is ready on its second release, and its bit toggles on
// each (PEND_ONE nodes are ready on their only release).  The window is
// circular (bit of id w at w mod PB_WIN) and moves in steps of PB_STEP ids;
// ids outside it use the global counters.  Ownership of an arrival moves with
// the window: an id leaving it with its bit set gets one global decrement; an
// id entering it with one global arrival (counter 1) gets its bit set and
// its counter back to 2.  At exit every set bit is written back.
constexpr int PB_WORDS = 3840;
constexpr uint32_t PB_WIN = (uint32_t)PB_WORDS * 32u;   // 122,880 ids
constexpr uint32_t PB_STEP = PB_WIN / 4;
constexpr uint32_t PB_MARGIN = 8192;    // ids kept below the frontier's smallest
static_assert(PB_STEP % 32 == 0, "window steps by whole words");
// A second, fixed bitmap holds the counters of the highest ids [tailStart, n):
// nodes a rewrite adds are numbered after all others, so their counters would
// otherwise never be inside the sliding window.  The sliding window stops at
// tailStart, so no id is in both.
constexpr int PT_WORDS = 4096;
constexpr uint32_t PT_WIN = (uint32_t)PT_WORDS * 32u;   // 131,072 ids
static_assert(WL_DT >= WL_FM, "every node of a wave finds a duplicate-table entry");

// Source of a fanin's representative: SRC_POS | frontier position of this
// wave, or the resolved rl value (literal | level << 32).
__device__ __forceinline__ unsigned long long src_of(unsigned long long r) {
    return (uint32_t)r == INVALID32 ? (SRC_POS | (r >> 32)) : r;
}

// ids [lo, hi) leave the window: their set bits become one global decrement
// (lo, hi multiples of 32; `circ` = the sliding window's circular indexing,
// else word (id - off) / 32)
__device__ __forceinline__ uint32_t pb_word(uint32_t base, bool circ, uint32_t off) {
    return circ ? (base % PB_WIN) / 32 : (base - off) / 32;
}

__device__ __forceinline__ void pb_release_range(uint32_t *s_pb, uint32_t lo, uint32_t hi, uint32_t *pending,
                                                 int32_t n, bool circ = true, uint32_t off = 0) {
    for (uint32_t wd = lo / 32 + threadIdx.x; wd < hi / 32; wd += WG1) {
        const uint32_t base = wd * 32;
        uint32_t bits = s_pb[pb_word(base, circ, off)];
        s_pb[pb_word(base, circ, off)] = 0;
        while (bits) {
            const uint32_t j = __ffs(bits) - 1;
            bits &= bits - 1;
            if (base + j < (uint32_t)n)
                __hip_atomic_fetch_add(&pending[base + j], 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ids [lo, hi) enter the window (their words are clear): an id with one
// global arrival takes it over as its bit
__device__ __forceinline__ void pb_take_range(uint32_t *s_pb, uint32_t lo, uint32_t hi, uint32_t *pending, int32_t n,
                                              bool circ = true, uint32_t off = 0) {
    for (uint32_t wd = lo / 32 + threadIdx.x; wd < hi / 32; wd += WG1) {
        const uint32_t base = wd * 32;
        uint32_t bits = 0;
        for (uint32_t j = 0; j < 32; j++) {
            const uint32_t w = base + j;
            if (w < (uint32_t)n && pending[w] == 1u) {
                bits |= 1u << j;
                pending[w] = 2u;
            }
        }
        s_pb[pb_word(base, circ, off)] = bits;
    }
}

__global__ void __launch_bounds__(WG1) k_waves_1wg(uint32_t k0, uint32_t thresh, WaveArgs a,
                                                   const NodeRec *__restrict__ recs, const NodeOfl *__restrict__ ofls,
                                                   const uint32_t *__restrict__ adjY, uint32_t *kOut, int32_t n) {
    __shared__ uint32_t e_v[2][WL_FM], e_f0[2][WL_FM], e_f1[2][WL_FM], e_slot[2][WL_FM];
    __shared__ unsigned long long e_s0[2][WL_FM], e_s1[2][WL_FM];
    __shared__ uint32_t g_w[2][WL_EM], g_y[2][WL_EM];
    __shared__ uint16_t g_q[2][WL_EM];
    __shared__ unsigned long long s_rep[WL_FM];
    __shared__ unsigned long long d_key[WL_DT], d_val[WL_DT];
    __shared__ uint32_t s_cnt[2], s_ecnt[2], s_over;
    __shared__ uint32_t s_pb[PB_WORDS];
    __shared__ uint32_t s_pt[PT_WORDS];
    __shared__ uint32_t s_pbase, s_fmin[2];
    const uint32_t tid = threadIdx.x;
    const int32_t nPIs = a.nPIs;
    uint32_t k = k0;
    const uint32_t pn = a.fcnt[k - 1], ps = a.fstart[k - 1];
    uint32_t cn = a.fcnt[k], cs = a.fstart[k];
    for (uint32_t i = tid; i < WL_DT; i += WG1) { d_key[i] = 0; d_val[i] = 0; }
    if (tid == 0) {
        s_cnt[0] = 0; s_cnt[1] = 0; s_ecnt[0] = 0; s_ecnt[1] = 0; s_over = 0;
        s_fmin[0] = 0xFFFFFFFFu; s_fmin[1] = 0xFFFFFFFFu;
    }
    // the previous wave may come from the grid-wide kernel, unresolved
    for (uint32_t i = tid; i < pn; i += WG1) resolve<__HIP_MEMORY_SCOPE_WORKGROUP>(a, a.order[ps + i]);
    __syncthreads();
    if (cn > thresh || cn > (uint32_t)WL_FM) {
        if (tid == 0) *kOut = k;
        return;
    }
    // first frontier from global memory: mark positions, then load records
    for (uint32_t i = tid; i < cn; i += WG1) {
        const uint32_t v = a.order[cs + i];
        e_v[0][i] = v;
        a.rl[v] = (unsigned long long)INVALID32 | (unsigned long long)i << 32;
        atomicMin(&s_fmin[0], v);
    }
    __syncthreads();
    // counter windows: the tail of the highest ids, and a sliding window from
    // the frontier's smallest id up to the tail
    // rounded up to a whole word, so that [tailStart, n) spans at most PT_WORDS words
    const uint32_t tailStart = (((uint32_t)n > PT_WIN ? (uint32_t)n - PT_WIN : 0u) + 31u) & ~31u;
    pb_take_range(s_pt, tailStart, (uint32_t)n + 31u, a.pending, n, false, tailStart);
    const uint32_t pb0 = (s_fmin[0] > PB_MARGIN ? s_fmin[0] - PB_MARGIN : 0u) & ~31u;
    pb_take_range(s_pb, pb0, min(pb0 + PB_WIN, tailStart), a.pending, n);
    if (tid == 0) { s_pbase = pb0; s_fmin[0] = 0xFFFFFFFFu; }
    __syncthreads();
    for (uint32_t i = tid; i < cn; i += WG1) {
        const uint32_t v = e_v[0][i];
        const RecFull r = load_full(recs, v);
        OflFull o{};
        if (r.cnt > REC_INLINE) o = load_ofl(ofls, v);
        const uint32_t u0 = LID(r.f0), u1 = LID(r.f1);
        e_f0[0][i] = r.f0;
        e_f1[0][i] = r.f1;
        e_slot[0][i] = r.loc;
        // fanins are in earlier waves, all resolved
        e_s0[0][i] = u0 <= (uint32_t)nPIs ? 2ull * u0 : a.rl[u0];
        e_s1[0][i] = u1 <= (uint32_t)nPIs ? 2ull * u1 : a.rl[u1];
        const uint32_t eb = atomicAdd(&s_ecnt[0], r.cnt);
        if (eb + r.cnt <= (uint32_t)WL_EM) put_edges(g_w[0], g_y[0], g_q[0], eb, r, o, (uint16_t)i, a.adj, adjY);
        else s_over = 1;
    }
    __syncthreads();
    if (s_over) {
        pb_release_range(s_pb, s_pbase, min(s_pbase + PB_WIN, tailStart), a.pending, n);
        pb_release_range(s_pt, tailStart, (uint32_t)n + 31u, a.pending, n, false, tailStart);
        if (tid == 0) *kOut = k;
        return;
    }
    int b = 0;
    uint32_t last = 0, np = 0;
    while (true) {
        const int nb = b ^ 1;
        const uint32_t ne = s_ecnt[b], nbase = cs + cn;
        // ---- phase A: hash this wave's nodes
        bool have = false, tri = false, fin = false;
        uint32_t v = 0, r = 0, lev = 0, dh = 0;
        if (tid < cn) {
            have = true;
            v = e_v[b][tid];
            const uint32_t raw0 = e_f0[b][tid], raw1 = e_f1[b][tid];
            const unsigned long long s0 = e_s0[b][tid], s1 = e_s1[b][tid];
            const unsigned long long x0 = (s0 & SRC_POS) ? s_rep[(uint32_t)s0] : s0;
            const unsigned long long x1 = (s1 & SRC_POS) ? s_rep[(uint32_t)s1] : s1;
            uint32_t l0 = (uint32_t)x0 ^ LC(raw0), l1 = (uint32_t)x1 ^ LC(raw1);
            uint32_t la = (uint32_t)(x0 >> 32), lb = (uint32_t)(x1 >> 32);
            if (l0 > l1) { uint32_t t = l0; l0 = l1; l1 = t; t = la; la = lb; lb = t; }
            r = INVALID32;
            if (l0 == l1) { r = l0; lev = la; }
            else if (l0 == (l1 ^ 1u)) { r = 0; lev = 0; }
            else if (l0 == 1u) { r = l1; lev = lb; }
            else if (l0 == 0u) { r = 0; lev = 0; }
            tri = r != INVALID32;
            if (!tri) {
                lev = 1u + (la > lb ? la : lb);
                const unsigned long long val = (static_cast<unsigned long long>(k) << 32) | v;
                const uint32_t r0 = raw0 < raw1 ? raw0 : raw1, r1 = raw0 < raw1 ? raw1 : raw0;
                uint32_t loc = e_slot[b][tid];
                if (!(l0 == r0 && l1 == r1 && loc != NO_SLOT)) {
                    // fanins changed: probe for the mapped key
                    const uint64_t key = and_key(l0, l1);
                    uint64_t s = mix64(key) & a.mask;
                    loc = NO_SLOT;
                    for (uint64_t probe = 0; probe <= a.mask; ++probe) {
                        unsigned long long prev = KEY_EMPTY;
                        __hip_atomic_compare_exchange_strong(&a.ht[2 * s], &prev, (unsigned long long)key,
                                                             __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP);
                        ++np;
                        if (prev == KEY_EMPTY || prev == key) {
                            loc = loc_of_slot<__HIP_MEMORY_SCOPE_WORKGROUP>(a, s);
                            break;
                        }
                        s = (s + 1) & a.mask;
                    }
                    if (loc == NO_SLOT) { atomicOr(a.err, 1u); loc = LOC_HT; }
                    a.locOf[v] = loc;
                }
                const unsigned long long prev = __hip_atomic_fetch_min(vptr(a, loc), val, __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                ++np;
                if (prev != KEY_EMPTY && (uint32_t)(prev >> 32) < k) {
                    fin = true;                    // an earlier wave holds the slot
                    r = 2u * (uint32_t)prev;
                } else {
                    // smallest id of this wave on the value location
                    const unsigned long long dkey = (static_cast<unsigned long long>(k) << 32) | loc;
                    uint32_t h = (loc * 0x9E3779B1u) >> (32 - WL_DT_BITS);
                    for (int probe = 0; probe < WL_DT; probe++) {
                        unsigned long long cur = __hip_atomic_load(&d_key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if ((uint32_t)(cur >> 32) != k) {
                            unsigned long long exp = cur;
                            __hip_atomic_compare_exchange_strong(&d_key[h], &exp, dkey, __ATOMIC_RELAXED,
                                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            cur = exp == cur ? dkey : exp;
                        }
                        if (cur == dkey) break;
                        h = (h + 1) & (WL_DT - 1);
                    }
                    atomicMax(&d_val[h], (static_cast<unsigned long long>(k) << 32) | (uint32_t)~v);
                    dh = h;
                }
            }
        }
        // ---- phase A: release the fanout edges.  Edge t runs on thread
        // (t + cn rounded up to a wavefront) mod WG1: no wavefront both hashes
        // a node (and waits for its table atomic) and releases edges, unless
        // nodes and edges together outnumber the threads.  The two passes
        // are wavefront-uniform: a wavefront reserves its ready nodes' frontier
        // positions and edge-list ranges with one LDS atomic each (lane ranks
        // and a prefix sum of fanout counts), instead of one per node on a
        // single contended address.
        const uint32_t pbase = s_pbase;
        const uint32_t lane = tid & 63u;
        const uint32_t cnw = (cn + 63u) & ~63u;
        const uint32_t t0 = (tid + WG1 - cnw) % WG1;
#pragma unroll 1
        for (uint32_t pass = 0; pass < 2; pass++) {
            const uint32_t t = t0 + pass * WG1;
            const bool valid = t < ne;
            if (__ballot(valid) == 0) continue;
            uint32_t w = 0, yw = 0, y = 0;
            bool ready = false;
            if (valid) {
                w = g_w[b][t];
                yw = g_y[b][t];
                y = yw & Y_ID;
                ready = true;
                if (y > (uint32_t)nPIs) {
                    // two non-PI fanins: ready on the second release
                    const uint32_t bit = 1u << (w % 32u);
                    if (w >= tailStart) {
                        ready = (atomicXor(&s_pt[(w - tailStart) / 32], bit) & bit) != 0;
                    } else if (w - pbase < PB_WIN) {
                        ready = (atomicXor(&s_pb[(w % PB_WIN) / 32], bit) & bit) != 0;
                    } else {
                        ready = __hip_atomic_fetch_add(&a.pending[w], 0xFFFFFFFFu, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP) == 1u;
                    }
                }
            }
            const uint64_t rm = __ballot(ready);
            if (rm == 0) continue;
            // a ready node's record, overflow record and other fanin's
            // representative, in one trip
            uint32_t q = 0, x = 0;
            RecFull rw{};
            OflFull ow{};
            unsigned long long ry = 0;
            if (ready) {
                q = g_q[b][t];
                x = e_v[b][q];
                rw = load_full(recs, w);
                if (yw & Y_OFL) ow = load_ofl(ofls, w);
                if (y != x && y > (uint32_t)nPIs) ry = a.rl[y];
            }
            const int leader = __ffsll((unsigned long long)rm) - 1;
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(rm >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)rm, 0u));
            const uint32_t c = ready ? rw.cnt : 0u;
            // inclusive prefix sum over the wavefront: shifts within rows of 16
            // lanes, then each row adds the totals of the rows before it
            uint32_t incl = c;
            incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x111, 0xF, 0xF, false);   // row_shr:1
            incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x112, 0xF, 0xF, false);   // row_shr:2
            incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x114, 0xF, 0xF, false);   // row_shr:4
            incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x118, 0xF, 0xF, false);   // row_shr:8
            incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x142, 0xA, 0xF, false);   // row_bcast:15
            incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x143, 0xC, 0xF, false);   // row_bcast:31
            const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            uint32_t pb = 0, eb0 = 0;
            if ((int)lane == leader) {
                pb = atomicAdd(&s_cnt[nb], (uint32_t)__popcll((unsigned long long)rm));
                eb0 = atomicAdd(&s_ecnt[nb], total);
                atomicMin(&s_fmin[nb], w);       // one id per wavefront guides the counter window
            }
            pb = (uint32_t)__builtin_amdgcn_readlane((int)pb, leader);
            eb0 = (uint32_t)__builtin_amdgcn_readlane((int)eb0, leader);
            if (!ready) continue;
            const uint32_t p = pb + rank, eb = eb0 + incl - c;
            a.order[nbase + p] = w;
            if (p >= (uint32_t)WL_FM) { s_over = 1; continue; }
            const unsigned long long sx = SRC_POS | q;
            const unsigned long long sy = y == x ? sx : y <= (uint32_t)nPIs ? 2ull * y : src_of(ry);
            e_v[nb][p] = w;
            e_f0[nb][p] = rw.f0;
            e_f1[nb][p] = rw.f1;
            e_slot[nb][p] = rw.loc;
            e_s0[nb][p] = LID(rw.f0) == x ? sx : sy;
            e_s1[nb][p] = LID(rw.f1) == x ? sx : sy;
            a.rl[w] = (unsigned long long)INVALID32 | (unsigned long long)p << 32;
            if (eb + c <= (uint32_t)WL_EM) put_edges(g_w[nb], g_y[nb], g_q[nb], eb, rw, ow, (uint16_t)p, a.adj, adjY);
            else s_over = 1;
        }
        __syncthreads();
        // ---- phase B: representatives of this wave
        if (have) {
            if (!tri && !fin) r = 2u * ~(uint32_t)d_val[dh];
            const unsigned long long x = pack_rl(r, lev);
            s_rep[tid] = x;
            a.rl[v] = x;
        }
        const uint32_t nn = s_cnt[nb];
        const bool over = s_over != 0;
        const uint32_t fmin = s_fmin[nb];
        if (tid == 0) {
            a.fstart[k + 1] = nbase;
            a.fcnt[k + 1] = nn;
            s_cnt[b] = 0;
            s_ecnt[b] = 0;
            s_fmin[b] = 0xFFFFFFFFu;
        }
        last = k;
        ++k;
        cs = nbase;
        cn = nn;
        b = nb;
        __syncthreads();
        if (over || cn == 0 || cn > thresh) break;
        // move the counter window one step once the next frontier is a step
        // past the margin, so that fanouts far ahead stay inside it
        if (fmin != 0xFFFFFFFFu && fmin >= pbase + PB_MARGIN + PB_STEP) {
            pb_release_range(s_pb, pbase, min(pbase + PB_STEP, tailStart), a.pending, n);
            __syncthreads();
            pb_take_range(s_pb, min(pbase + PB_WIN, tailStart), min(pbase + PB_WIN + PB_STEP, tailStart), a.pending, n);
            if (tid == 0) s_pbase = pbase + PB_STEP;
            __syncthreads();
        }
    }
    pb_release_range(s_pb, s_pbase, min(s_pbase + PB_WIN, tailStart), a.pending, n);
    pb_release_range(s_pt, tailStart, (uint32_t)n + 31u, a.pending, n, false, tailStart);
    if (tid == 0) {
        *kOut = k;
        if (last) atomicMax(a.depth, last);
    }
    block_add_u64(a.probes, np);
}

// wave of every processed node from the per-wave ranges of order[]: wave k
// holds positions [fstart[k], fstart[k + 1]) for k = 1..depth
__global__ void k_wave_of_order(uint32_t processed, uint32_t depth, const uint32_t *__restrict__ fstart,
                                const uint32_t *__restrict__ order, uint32_t *wave) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < processed; i += stride) {
        uint32_t lo = 1, hi = depth;     // largest k with fstart[k] <= i
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (fstart[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        wave[order[i]] = lo;
    }
}

__global__ void k_resolve_all(int32_t n, int32_t nPIs, const uint8_t *__restrict__ dead, unsigned long long *rl,
                              const uint32_t *__restrict__ locOf, const unsigned long long *__restrict__ ht,
                              const unsigned long long *__restrict__ val, uint8_t *surv) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < n; v += stride) {
        uint8_t s = 0;
        if (v > nPIs && !dead[v]) {
            unsigned long long x = rl[v];
            uint32_t r = (uint32_t)x;
            if (r == INVALID32) {
                const uint32_t loc = locOf[v];
                const unsigned long long w = (loc & LOC_HT) ? ht[2 * (size_t)(loc & ~LOC_HT) + 1] : val[loc];
                r = 2u * (uint32_t)(w & 0xFFFFFFFFull);
                rl[v] = (x & 0xFFFFFFFF00000000ull) | r;
            }
            s = (r == 2u * (uint32_t)v);
        }
        surv[v] = s;
    }
}

__global__ void k_gather_u32(int32_t m, const uint32_t *__restrict__ idx, const uint32_t *__restrict__ src,
                             uint32_t *dst) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tid; i < m; i += stride) dst[i] = src[idx[i]];
}

__global__ void k_new_ids(int32_t n, int32_t nPIs, int32_t m, const uint32_t *__restrict__ sorted,
                          uint32_t *newId) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v <= nPIs && v < n; v += stride) newId[v] = (uint32_t)v;
    for (int64_t p = tid; p < m; p += stride) newId[sorted[p]] = (uint32_t)(nPIs + 1 + p);
}

__device__ __forceinline__ uint32_t final_lit(uint32_t lit, int32_t nPIs, const unsigned long long *rl,
                                              const uint32_t *newId) {
    uint32_t u = LID(lit);
    uint32_t r = (u <= (uint32_t)nPIs) ? 2u * u : (uint32_t)rl[u];
    r ^= LC(lit);
    return 2u * newId[LID(r)] | LC(r);
}

__global__ void k_emit(int32_t m, int32_t nPIs, const uint32_t *__restrict__ sorted,
                       const uint32_t *__restrict__ f0, const uint32_t *__restrict__ f1,
                       const unsigned long long *__restrict__ rl, const uint32_t *__restrict__ newId,
                       uint32_t *of0, uint32_t *of1, uint32_t *olevel, const uint32_t *__restrict__ tagIn,
                       uint32_t *tagOut) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = tid; p < m; p += stride) {
        uint32_t v = sorted[p];
        uint32_t a = final_lit(f0[v], nPIs, rl, newId), b = final_lit(f1[v], nPIs, rl, newId);
        if (a > b) { uint32_t t = a; a = b; b = t; }
        int64_t o = nPIs + 1 + p;
        of0[o] = a;
        of1[o] = b;
        olevel[o] = (uint32_t)(rl[v] >> 32);
        if (tagOut) tagOut[o] = tagIn[v];
    }
}

__global__ void k_map_pos(int32_t nPOs, int32_t nPIs, const uint32_t *__restrict__ pos,
                          const unsigned long long *__restrict__ rl, const uint32_t *__restrict__ newId,
                          uint32_t *opos) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tid; i < nPOs; i += stride) opos[i] = final_lit(pos[i], nPIs, rl, newId);
}

// ---------------------------------------------------------------------------
// Strash of a graph composed by a rewrite (dev_strash_rw)
// ---------------------------------------------------------------------------
// Longest-path levels of a topologically ordered position array, by one
// workgroup: lev[p] = 0 where fa[p] == NOFAN, else 1 + max(lev[fa[p]],
// lev[fb[p]]), fanin positions below p.  Positions go in chunks of SW_WG, one
// per thread.  The levels of the last SW_RING positions stay in an LDS ring
// (one word each, SW_NOLEV until known), so a fanin at most SW_RING - 2 SW_WG
// positions back is read from LDS; older ones from lev in global memory.  A
// position whose fanin lies in its own chunk waits for that slot in LDS: every
// fanin sits at a lower position (the same or an earlier wavefront of the
// workgroup), so each wait ends; it is bounded anyway, and a fanin at or above
// its position sets *err instead of waiting.  Each thread clears the slot of
// its next chunk's position during the current chunk (a slot no position of
// the current chunk reads), so the one barrier per chunk both publishes the
// chunk's levels and leaves the next chunk's slots cleared: no waiting
// position can read a level left there by the position SW_RING earlier.
constexpr int SW_WG = 1024;
constexpr int SW_RING = 32768;              // 128 KiB of LDS
constexpr uint32_t NOFAN = 0xFFFFFFFFu;
constexpr uint32_t SW_NOLEV = 0xFFFFFFFFu;
constexpr uint32_t SW_SPIN_LIMIT = 1u << 22;
// rounds of the incremental strash before it gives way to the wave sweep
constexpr uint32_t kRwRoundLimit = 4096;

__device__ __forceinline__ uint32_t sw_lds_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__global__ void __launch_bounds__(SW_WG) k_level_sweep(uint32_t nPos, const uint32_t *__restrict__ fa,
                                                      const uint32_t *__restrict__ fb, uint32_t *lev,
                                                      uint32_t *err, uint32_t *maxLev) {
    __shared__ uint32_t ring[SW_RING];
    __shared__ uint2 s_fan[2][SW_WG];           // fanin positions of this chunk and the next
    __shared__ uint32_t s_max, s_bad;
    const uint32_t tid = threadIdx.x;
    // every wait is on a lower position, so a lower wavefront never waits on a
    // higher one: lower wavefronts issue first on their SIMD, so the spinning
    // of the waiting ones does not slow the one they wait for (wavefronts
    // w, w + 4, w + 8, w + 12 share a SIMD under round-robin placement)
    // (a wavefront-uniform value, so the branches are scalar: s_setprio runs
    // whatever the exec mask holds)
    switch (__builtin_amdgcn_readfirstlane(tid >> 8)) {
    case 0: __builtin_amdgcn_s_setprio(3); break;
    case 1: __builtin_amdgcn_s_setprio(2); break;
    case 2: __builtin_amdgcn_s_setprio(1); break;
    default: break;
    }
    if (tid == 0) { s_max = 0; s_bad = 0; }
    for (uint32_t i = tid; i < (uint32_t)SW_RING; i += SW_WG) ring[i] = SW_NOLEV;
    s_fan[0][tid] = tid < nPos ? make_uint2(fa[tid], fb[tid]) : make_uint2(NOFAN, NOFAN);
    uint32_t mx = 0;
    // the level of this thread's position in the previous chunk, stored one
    // chunk late: stores count on vmcnt like loads, so a store issued at the
    // end of a chunk would hold up the next chunk's first wait
    uint32_t pendP = 0, pendV = 0;
    bool pend = false;
    __syncthreads();
    for (uint32_t c = 0, h = 0; c < nPos; c += SW_WG, h ^= 1u) {
        const uint32_t p = c + tid;
        const bool active = p < nPos;
        const uint2 f = s_fan[h][tid];
        const uint32_t a = f.x, b = f.y;
        const bool leaf = !active || a == NOFAN;
        const bool bad = !leaf && (a >= p || b >= p);
        // ring slots of this chunk and the next held positions SW_RING earlier,
        // which this chunk reads from global memory
        const uint32_t oldest = c > (uint32_t)(SW_RING - 2 * SW_WG) ? c - (uint32_t)(SW_RING - 2 * SW_WG) : 0u;
        uint32_t la = 0, lb = 0;
        bool needA = false, needB = false;
        if (!leaf && !bad) {
            if (a < oldest) la = lev[a]; else needA = true;
            if (b < oldest) lb = lev[b]; else needB = true;
        }
        // wait for those (rare) loads here, then load the next chunk's fanins
        // (staged in LDS at the end of the chunk): their latency overlaps this
        // chunk's work, and nothing inside the wait loop below waits on them
        __builtin_amdgcn_s_waitcnt(0x0F70);     // vmcnt(0)
        if (pend) lev[pendP] = pendV;
        const uint32_t q = p + SW_WG;
        const uint32_t na = q < nPos ? fa[q] : NOFAN;
        const uint32_t nb = q < nPos ? fb[q] : NOFAN;
        // a leaf publishes at once; the other slots of this chunk were cleared
        // during the previous one (or before the first); clear the next one's
        if (active && leaf) __hip_atomic_store(&ring[p & (SW_RING - 1)], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        ring[(p + SW_WG) & (SW_RING - 1)] = SW_NOLEV;
        uint32_t val = 0;
        // the value is published inside the loop, in the iteration that
        // completes it: a lane of the same wavefront may be waiting for it
        bool done = leaf || bad;
        uint32_t spins = 0;
        // both slots are read every round (a and b are positions below p here,
        // so the masked slots are in the ring), so the two LDS reads are in
        // flight together instead of one after the other
        while (!done) {
            const uint32_t x = sw_lds_load(&ring[a & (SW_RING - 1)]);
            const uint32_t y = sw_lds_load(&ring[b & (SW_RING - 1)]);
            if (needA && x != SW_NOLEV) { la = x; needA = false; }
            if (needB && y != SW_NOLEV) { lb = y; needB = false; }
            if (!needA && !needB) {
                val = 1u + (la > lb ? la : lb);
                __hip_atomic_store(&ring[p & (SW_RING - 1)], val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                done = true;
            } else if (++spins > SW_SPIN_LIMIT) {
                atomicOr(&s_bad, 1u);
                done = true;
            }
        }
        if (bad) atomicOr(&s_bad, 1u);
        pend = active;
        pendP = p;
        pendV = val;
        mx = val > mx ? val : mx;
        s_fan[h ^ 1u][tid] = make_uint2(na, nb);
        __syncthreads();
        if (s_bad) break;
    }
    if (pend) lev[pendP] = pendV;
    atomicMax(&s_max, mx);
    __syncthreads();
    if (tid == 0) {
        atomicMax(maxLev, s_max);
        if (s_bad) atomicOr(err, 1u);
    }
}

// A topological key of the composed graph: starting from the input graph's
// levels (new nodes from 0), every live AND is raised to one above its fanins'
// keys until nothing moves.  Keys only grow and are bounded, so the loop
// ends; the fixpoint orders every edge (key of a fanin < key of its fanout).
// Most old nodes keep their old level; only nodes above a replaced root whose
// replacement sits higher, and the new nodes, move, a few steps each.
__global__ void k_rw_key_init(int32_t n, int32_t N, const uint32_t *__restrict__ glevel, uint32_t *key) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < N; v += stride) key[v] = v < n ? glevel[v] : 0u;
}

__global__ void k_rw_key_relax(int32_t N, int32_t nPIs, const uint32_t *__restrict__ f0,
                               const uint32_t *__restrict__ f1, const uint8_t *__restrict__ dead, uint32_t *key,
                               uint32_t *moved) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    uint32_t mv = 0;
    for (int64_t v = nPIs + 1 + tid; v < N; v += stride) {
        if (dead[v]) continue;
        const uint32_t ka = __hip_atomic_load(&key[LID(f0[v])], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t kb = __hip_atomic_load(&key[LID(f1[v])], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t need = 1u + (ka > kb ? ka : kb);
        if (__hip_atomic_load(&key[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            atomicMax(&key[v], need);
            ++mv;
        }
    }
    if (mv) atomicAdd(moved, mv);
}

// Sets *loose if some live AND's key is not one more than the larger key of
// its fanins, and raises *maxKey to the largest live key.  If no key is loose
// the keys are the composed graph's levels: the PIs and the constant have key
// 0 (their input level), and 0 there with 1 + the larger fanin value at every
// live AND defines the longest-path level uniquely.
__global__ void k_rw_key_loose(int32_t N, int32_t nPIs, const uint32_t *__restrict__ f0,
                               const uint32_t *__restrict__ f1, const uint8_t *__restrict__ dead,
                               const uint32_t *__restrict__ key, uint32_t *loose, uint32_t *maxKey) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    bool any = false;
    uint32_t mx = 0;
    for (int64_t v = nPIs + 1 + tid; v < N; v += stride) {
        if (dead[v]) continue;
        const uint32_t ka = key[LID(f0[v])], kb = key[LID(f1[v])], k = key[v];
        any |= k != 1u + (ka > kb ? ka : kb);
        mx = k > mx ? k : mx;
    }
    if (__ballot(any) != 0 && (threadIdx.x & 63) == 0) atomicOr(loose, 1u);
    if (mx) atomicMax(maxKey, mx);
}

// The level of every position when the keys are the levels: 0 for the PIs,
// the constant and dangling nodes.
__global__ void k_rw_levpos(int32_t N, int32_t nPIs, const uint8_t *__restrict__ dead,
                            const uint32_t *__restrict__ key, const uint32_t *__restrict__ posOf, uint32_t *levPos) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < N; v += stride) levPos[posOf[v]] = (v <= nPIs || dead[v]) ? 0u : key[v];
}

__global__ void k_iota_u32(int32_t N, uint32_t *x) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < N; v += stride) x[v] = (uint32_t)v;
}

// position of every object in the key order
__global__ void k_rw_pos(int32_t N, const uint32_t *__restrict__ order, uint32_t *posOf) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = tid; p < N; p += stride) posOf[order[p]] = (uint32_t)p;
}

// Positions of a whole graph's level sweep: the constant, the PIs and the
// live ANDs keep one each, in id order (keep[v] = 1), dangling ANDs none.
__global__ void k_keep_flags(int32_t N, int32_t nPIs, const uint8_t *__restrict__ dead, uint32_t *keep) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < N; v += stride) keep[v] = (v <= nPIs || !dead[v]) ? 1u : 0u;
}

// Fanin positions at those positions (posOf: exclusive scan of keep); the
// constant and the PIs have none.
__global__ void k_full_fanpos(int32_t N, int32_t nPIs, const uint32_t *__restrict__ f0,
                              const uint32_t *__restrict__ f1, const uint8_t *__restrict__ dead,
                              const uint32_t *__restrict__ posOf, uint32_t *fa, uint32_t *fb) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < N; v += stride) {
        if (v <= nPIs) {
            fa[posOf[v]] = NOFAN;
            fb[posOf[v]] = NOFAN;
        } else if (!dead[v]) {
            fa[posOf[v]] = posOf[LID(f0[v])];   // fanins of a live node are live
            fb[posOf[v]] = posOf[LID(f1[v])];
        }
    }
}

// Fanin positions of every live AND at its sweep position; PIs, the constant
// and dangling nodes have none.
__global__ void k_rw_fanpos(int32_t N, int32_t nPIs, const uint32_t *__restrict__ f0, const uint32_t *__restrict__ f1,
                            const uint8_t *__restrict__ dead, const uint32_t *__restrict__ posOf, uint32_t *fa,
                            uint32_t *fb, uint32_t *err) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < N; v += stride) {
        const uint32_t p = posOf[v];
        if (p >= (uint32_t)N) { atomicOr(err, 8u); continue; }   // not a permutation: no write out of range
        if (v <= nPIs || dead[v]) { fa[p] = NOFAN; fb[p] = NOFAN; continue; }
        fa[p] = posOf[LID(f0[v])];
        fb[p] = posOf[LID(f1[v])];
    }
}

// Wave of every object (0 for PIs, the constant and dangling nodes), the
// initial representative literal (itself), and the first dirty set as flags
// (inD): live new nodes and live old nodes with a redirected fanin.  (The list
// is a compaction of the flags: appends to one counter serialise on its
// address, about 1.8 ms for the 156k wavefronts of a 10M-node graph.)
__global__ void k_rw_seed(int32_t n, int32_t N, int32_t nPIs, const uint32_t *__restrict__ f0,
                          const uint32_t *__restrict__ f1, const uint32_t *__restrict__ g0,
                          const uint32_t *__restrict__ g1, const uint8_t *__restrict__ dead,
                          const uint32_t *__restrict__ posOf, const uint32_t *__restrict__ levPos, uint32_t *wave,
                          uint32_t *rep, uint32_t *inD) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < N; v += stride) {
        const bool live = v > nPIs && !dead[v];
        wave[v] = live ? levPos[posOf[v]] : 0u;
        rep[v] = 2u * (uint32_t)v;
        const bool d = live && (v >= n || f0[v] != g0[v] || f1[v] != g1[v]);
        inD[v] = d ? 1u : 0u;
    }
}

__device__ __forceinline__ uint32_t rep_lit(const uint32_t *rep, int32_t nPIs, uint32_t lit) {
    const uint32_t u = LID(lit);
    return (u <= (uint32_t)nPIs ? 2u * u : rep[u]) ^ LC(lit);
}

// Round, step 1: every dirty node's key over the representatives of the
// previous round; trivial ANDs resolve at once (hash_table.h:435-449 rules),
// the others enter this round's table with value (wave << 32 | id), kept by
// minimum.  dkey keeps the key, dslot its slot (NO_SLOT: resolved, nrep).
__global__ void k_rw_insert(uint32_t nD, const uint32_t *__restrict__ dlist, int32_t nPIs,
                            const uint32_t *__restrict__ f0, const uint32_t *__restrict__ f1,
                            const uint32_t *__restrict__ rep, const uint32_t *__restrict__ wave,
                            unsigned long long *ht, uint64_t mask, uint2 *dkey, uint32_t *dslot, uint32_t *nrep,
                            uint32_t *err, unsigned long long *probes) {
    uint32_t np = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nD; i += stride) {
        const uint32_t v = dlist[i];
        uint32_t l0 = rep_lit(rep, nPIs, f0[v]), l1 = rep_lit(rep, nPIs, f1[v]);
        if (l0 > l1) { const uint32_t t = l0; l0 = l1; l1 = t; }
        uint32_t r = INVALID32;
        if (l0 == l1) r = l0;
        else if (l0 == (l1 ^ 1u)) r = 0;
        else if (l0 == 1u) r = l1;
        else if (l0 == 0u) r = 0;
        dkey[i] = make_uint2(l0, l1);
        if (r != INVALID32) { nrep[i] = r; dslot[i] = NO_SLOT; continue; }
        const uint64_t key = and_key(l0, l1);
        uint64_t s = mix64(key) & mask;
        uint32_t slot = NO_SLOT;
        for (uint64_t probe = 0; probe <= mask; ++probe) {
            unsigned long long prev = KEY_EMPTY;
            __hip_atomic_compare_exchange_strong(&ht[2 * s], &prev, (unsigned long long)key, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ++np;
            if (prev == KEY_EMPTY || prev == key) { slot = (uint32_t)s; break; }
            s = (s + 1) & mask;
        }
        nrep[i] = INVALID32;
        dslot[i] = slot;
        if (slot == NO_SLOT) { atomicOr(err, 2u); continue; }
        atomicMin(&ht[2 * (size_t)slot + 1], ((unsigned long long)wave[v] << 32) | v);
        ++np;
    }
    block_add_u64(probes, np);
}

// The node of the rewrite's input graph with this key, if any: a scan of its
// structural table bucket of the fanin of larger id (entries {first fanin
// literal, node | complement of the second << 31}).  Keys over a new node
// have none.
__device__ __forceinline__ int32_t old_holder(int32_t n, const uint32_t *__restrict__ tOff,
                                              const uint2 *__restrict__ tEnt, uint32_t l0, uint32_t l1,
                                              uint32_t &np) {
    if (LID(l1) >= (uint32_t)n || LID(l0) >= (uint32_t)n) return -1;
    const uint32_t hc = LC(l1) << 31, e = tOff[LID(l1) + 1];
    for (uint32_t k = tOff[LID(l1)]; k < e; k++) {
        const uint2 x = tEnt[k];
        ++np;
        if (x.x == l0 && (x.y & 0x80000000u) == hc) return (int32_t)(x.y & 0x7FFFFFFFu);
    }
    return -1;
}

__device__ __forceinline__ void join_dirty(uint32_t *inD, uint32_t v, uint32_t *dlist, uint32_t *dcount,
                                           uint32_t dcap, uint32_t *err) {
    if (atomicCAS(&inD[v], 0u, 1u) != 0u) return;
    const uint32_t k = atomicAdd(dcount, 1u);
    if (k < dcap) dlist[k] = v;
    else atomicOr(err, 4u);
}

// Round, step 2: a non-trivial dirty node takes the smallest (wave, id) among
// this round's dirty nodes of its key and the key's clean holder (a live old
// node outside the dirty set, which still has its input fanins); a clean
// holder that loses joins the dirty set.  changed counts representatives that
// changed.
__global__ void k_rw_resolve(uint32_t nD, const uint32_t *__restrict__ dlistIn, int32_t n,
                             const uint8_t *__restrict__ dead, const uint32_t *__restrict__ wave, uint32_t *rep,
                             uint32_t *inD, const unsigned long long *__restrict__ ht, const uint2 *__restrict__ dkey,
                             const uint32_t *__restrict__ dslot, const uint32_t *__restrict__ nrep,
                             const uint32_t *__restrict__ tOff, const uint2 *__restrict__ tEnt, uint32_t *dlist,
                             uint32_t *dcount, uint32_t dcap, uint32_t *changed, uint32_t *err,
                             unsigned long long *probes) {
    uint32_t np = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nD; i += stride) {
        const uint32_t v = dlistIn[i];
        uint32_t r = nrep[i];
        if (r == INVALID32) {
            const uint32_t slot = dslot[i];
            if (slot == NO_SLOT) continue;   // table overflow, reported
            const unsigned long long dmin = ht[2 * (size_t)slot + 1];
            ++np;
            unsigned long long win = dmin;
            const uint2 k = dkey[i];
            const int32_t c = old_holder(n, tOff, tEnt, k.x, k.y, np);
            if (c >= 0 && !dead[c] && !__hip_atomic_load(&inD[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                const unsigned long long cv = ((unsigned long long)wave[c] << 32) | (uint32_t)c;
                if (cv < dmin) win = cv;
                else join_dirty(inD, (uint32_t)c, dlist, dcount, dcap, err);
            }
            r = 2u * (uint32_t)(win & 0xFFFFFFFFull);
        }
        if (r != rep[v]) {
            rep[v] = r;
            atomicAdd(changed, 1u);
        }
    }
    block_add_u64(probes, np);
}

// Round, step 3: a live node outside the dirty set whose fanin's
// representative is no longer the fanin itself joins the dirty set.
__global__ void k_rw_pull(int32_t N, int32_t nPIs, const uint32_t *__restrict__ f0, const uint32_t *__restrict__ f1,
                          const uint8_t *__restrict__ dead, const uint32_t *__restrict__ rep, uint32_t *inD,
                          uint32_t *dlist, uint32_t *dcount, uint32_t dcap, uint32_t *err) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + tid; v < N; v += stride) {
        if (dead[v] || inD[v]) continue;
        const uint32_t a = LID(f0[v]), b = LID(f1[v]);
        const bool moved = (a > (uint32_t)nPIs && rep[a] != 2u * a) || (b > (uint32_t)nPIs && rep[b] != 2u * b);
        if (moved) join_dirty(inD, (uint32_t)v, dlist, dcount, dcap, err);
    }
}

__global__ void k_rw_survivors(int32_t N, int32_t nPIs, const uint8_t *__restrict__ dead,
                               const uint32_t *__restrict__ rep, uint8_t *surv) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < N; v += stride) surv[v] = v > nPIs && !dead[v] && rep[v] == 2u * (uint32_t)v;
}

// Sets *shorter if some live node's representative lies on a lower wave than
// the node (a merge into a node of a lower wave, or a fold to a fanin, a PI or
// a constant).  If none does, every survivor's output level is its wave: its
// output fanins are representatives on the waves of its composed fanins.
__global__ void k_rw_shorter(int32_t N, int32_t nPIs, const uint8_t *__restrict__ dead,
                             const uint32_t *__restrict__ rep, const uint32_t *__restrict__ wave, uint32_t *shorter) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    bool any = false;
    for (int64_t v = nPIs + 1 + tid; v < N; v += stride) {
        if (dead[v]) continue;
        const uint32_t r = LID(rep[v]);
        if (r == (uint32_t)v) continue;
        const uint32_t wr = r <= (uint32_t)nPIs ? 0u : wave[r];
        any |= wr != wave[v];
    }
    if (__ballot(any) != 0 && (threadIdx.x & 63) == 0) atomicOr(shorter, 1u);
}

__device__ __forceinline__ uint32_t out_lit(uint32_t lit, int32_t nPIs, const uint32_t *rep, const uint32_t *newId) {
    const uint32_t r = rep_lit(rep, nPIs, lit);
    return 2u * newId[LID(r)] | LC(r);
}

// Output fanins and either the output levels (olevel non-null: every
// survivor's level is its wave) or, for the level sweep, the fanin positions
// (= output ids).
__global__ void k_rw_emit(int32_t m, int32_t nPIs, const uint32_t *__restrict__ sorted, const uint32_t *__restrict__ f0,
                          const uint32_t *__restrict__ f1, const uint32_t *__restrict__ rep,
                          const uint32_t *__restrict__ newId, uint32_t *of0, uint32_t *of1, uint32_t *fa,
                          uint32_t *fb, const uint32_t *__restrict__ wave, uint32_t *olevel,
                          const uint32_t *__restrict__ tagIn, uint32_t *tagOut) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tid; i <= nPIs; i += stride) {
        if (olevel) olevel[i] = 0u;
        else { fa[i] = NOFAN; fb[i] = NOFAN; }
    }
    for (int64_t p = tid; p < m; p += stride) {
        const uint32_t v = sorted[p];
        uint32_t a = out_lit(f0[v], nPIs, rep, newId), b = out_lit(f1[v], nPIs, rep, newId);
        if (a > b) { const uint32_t t = a; a = b; b = t; }
        const int64_t o = nPIs + 1 + p;
        of0[o] = a;
        of1[o] = b;
        if (olevel) {
            olevel[o] = wave[v];
        } else {
            fa[o] = LID(a);
            fb[o] = LID(b);
        }
        if (tagOut) tagOut[o] = tagIn[v];
    }
}

__global__ void k_rw_pos_out(int32_t nPOs, int32_t nPIs, const uint32_t *__restrict__ pos,
                             const uint32_t *__restrict__ rep, const uint32_t *__restrict__ newId, uint32_t *opos) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tid; i < nPOs; i += stride) opos[i] = out_lit(pos[i], nPIs, rep, newId);
}

__global__ void k_fill_u64(size_t n, unsigned long long *p, unsigned long long v) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

struct Tmp {
    gls_ctx *ctx;
    std::vector<void *> ptrs;
    template <class T> T *get(size_t n) {
        T *p = pool_get<T>(ctx, n);
        if (p) ptrs.push_back(p);
        return p;
    }
    ~Tmp() {
        for (void *p : ptrs) pool_put(ctx, p);
    }
};

// order[0 .. *count) = the ids v with flag[v] set, ascending (a compaction:
// appends to one counter serialise on its address)
static int select_ids(gls_ctx *ctx, Tmp &t, int32_t n, const uint8_t *flag, uint32_t *order, uint32_t *count) {
    hipStream_t st = ctx->stream;
    rocprim::counting_iterator<uint32_t> ids(0u);
    size_t tb = 0;
    GLS_HIP(ctx, rocprim::select(nullptr, tb, ids, flag, order, count, (size_t)n, st));
    void *tmp = t.get<uint8_t>(std::max<size_t>(tb, 1));
    if (!tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: list storage");
    GLS_HIP(ctx, rocprim::select(tmp, tb, ids, flag, order, count, (size_t)n, st));
    return GLS_OK;
}


// Grid of a sweep round: sized from the frontier seen at the start of a batch,
// never below 8 workgroups so that a growing frontier is not serialised.
int sweep_grid(uint32_t hint) {
    int g = (int)((2ull * hint + kBlock - 1) / kBlock);
    return g < 8 ? 8 : (g > 1024 ? 1024 : g);
}

uint32_t bits_for(uint32_t x) {
    uint32_t b = 1;
    while (b < 32 && (x >> b) != 0) ++b;
    return b;
}

} // namespace

// Runs sweep rounds (kernel `launch(k)` for k = first..) in growing batches
// until the frontier is empty.  Returns the number of rounds launched.
template <class F>
static int run_sweep(gls_ctx *ctx, SweepCtl *d_ctl, uint32_t first, F launch, SweepCtl *h_ctl, uint32_t *rounds) {
    uint32_t k = first, batch = 4;
    for (;;) {
        for (uint32_t i = 0; i < batch; ++i, ++k) launch(k, h_ctl->curSize);
        GLS_HIP(ctx, hipMemcpyAsync(h_ctl, d_ctl, sizeof(SweepCtl), hipMemcpyDeviceToHost, ctx->stream));
        GLS_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (h_ctl->err) return gls_fail(ctx, GLS_ERR_INTERNAL, "strash table overflow");
        if (h_ctl->curSize == 0) break;
        if (batch < 512) batch *= 2;
    }
    *rounds = k - first;
    return GLS_OK;
}

// tagDepth[t] = the largest wave of a node tagged t (dangling nodes have wave 0)
__global__ void __launch_bounds__(kBlock) k_tag_depth(int32_t n, int32_t nPIs, const uint32_t *__restrict__ wave,
                                                      const uint32_t *__restrict__ tag, uint32_t *tagDepth) {
    __shared__ uint32_t s_bins[TAG_BINS];
    TagBins tb{s_bins};
    tb.init();
    int64_t b, e;
    tag_chunk(nPIs + 1, n, b, e);
    for (int64_t v = b + threadIdx.x; v < e; v += blockDim.x)
        if (wave[v]) tb.max(tagDepth, tag[v], wave[v]);
    tb.flush_max(tagDepth, 0xFFFFFFFFu);
}

int dev_strash(gls_ctx *ctx, const DevGraph &in, DevGraph &out, int32_t *depthOut, gls_op_info *info,
               const uint32_t *tagIn, uint32_t **tagOut, uint32_t *tagDepth) {
    const int32_t n = in.nObjs, nPIs = in.nPIs, nPOs = in.nPOs;
    hipStream_t st = ctx->stream;
    Tmp t{ctx};
    uint32_t *ref = t.get<uint32_t>(n), *order = t.get<uint32_t>(n), *pending = t.get<uint32_t>(n);
    uint32_t *cnt = t.get<uint32_t>(n + 1), *adjOff = t.get<uint32_t>(n + 1), *cursor = t.get<uint32_t>(n + 1);
    uint32_t *adj = t.get<uint32_t>(2 * (size_t)n + 1), *adjY = t.get<uint32_t>(2 * (size_t)n + 1);
    unsigned long long *rl = t.get<unsigned long long>(n);
    uint32_t *wave = t.get<uint32_t>(n), *newId = t.get<uint32_t>(n);
    uint32_t *survIds = t.get<uint32_t>(n), *survWave = t.get<uint32_t>(n);
    uint32_t *sortIds = t.get<uint32_t>(n), *sortWave = t.get<uint32_t>(n), *d_count = t.get<uint32_t>(4);
    uint8_t *dead = t.get<uint8_t>(n), *surv = t.get<uint8_t>(n);
    uint32_t *locOf = t.get<uint32_t>(n), *slotTmp = t.get<uint32_t>(n);
    uint32_t *fcnt = t.get<uint32_t>(n + 4), *fstart = t.get<uint32_t>(n + 4);
    unsigned long long *val = t.get<unsigned long long>(n);
    uint32_t *d_misc = t.get<uint32_t>(8);   // [0] depth, [1] err, [2] next wave, [4..5] probe count
    NodeRec *recs = t.get<NodeRec>(n);
    NodeOfl *ofls = t.get<NodeOfl>(n);
    SweepCtl *ctl = t.get<SweepCtl>(1);
    // room for every live node's own key plus a changed key per node
    const gls_tuning &tune = tuning_for(ctx, n);
    uint64_t cap = 1024;
    while (cap < (uint64_t)tune.strash_table_factor * (uint64_t)n) cap <<= 1;
    if (cap > (1ull << 31)) return gls_fail(ctx, GLS_ERR_ARG, "strash: %d objects exceed the table's slot range", n);
    unsigned long long *ht = t.get<unsigned long long>(2 * cap);
    if (!ref || !order || !pending || !cnt || !adjOff || !cursor || !adj || !rl || !wave ||
        !newId || !survIds || !survWave || !sortIds || !sortWave || !d_count || !dead || !surv || !locOf || !slotTmp || !val ||
        !fcnt || !fstart || !d_misc || !ctl || !ht || !recs || !ofls || !adjY)
        return gls_fail(ctx, GLS_ERR_NOMEM, "strash: device allocation failed (n=%d)", n);

    GLS_HIP(ctx, hipMemsetAsync(ref, 0, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(dead, 0, n, st));
    GLS_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(uint32_t) * (n + 1), st));
    GLS_HIP(ctx, hipMemsetAsync(rl, 0xFF, sizeof(unsigned long long) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(wave, 0, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(ht, 0xFF, sizeof(unsigned long long) * 2 * cap, st));
    GLS_HIP(ctx, hipMemsetAsync(val, 0xFF, sizeof(unsigned long long) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(ctl, 0, sizeof(SweepCtl), st));
    const int gN = grid_for(n);
    SweepCtl h_ctl{};
    uint32_t rounds = 0;

    // 1. fanout counts and dangling cascade
    uint8_t *zflag = t.get<uint8_t>((size_t)n);
    if (!zflag) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: frontier flags");
    k_count_refs<<<gN, kBlock, 0, st>>>(n, nPIs, in.f0, in.f1, nPOs, in.pos, ref);
    k_zero_flags<<<gN, kBlock, 0, st>>>(n, nPIs, ref, nullptr, zflag);
    GLS_HIP(ctx, hipGetLastError());
    GLS_TRY(select_ids(ctx, t, n, zflag, order, &ctl->curSize));
    GLS_HIP(ctx, hipMemcpyAsync(&h_ctl, ctl, sizeof(SweepCtl), hipMemcpyDeviceToHost, st));
    GLS_HIP(ctx, hipStreamSynchronize(st));
    if (h_ctl.curSize > 0) {
        GLS_TRY(run_sweep(ctx, ctl, 1, [&](uint32_t k, uint32_t hint) {
            k_dangle_round<<<sweep_grid(hint), kBlock, 0, st>>>(k, ctl, nPIs, in.f0, in.f1, ref, dead, order);
        }, &h_ctl, &rounds));
    }

    const uint32_t nDead = h_ctl.processed;
    // 2. fanout CSR over live AND nodes and the first frontier (wave 1)
    GLS_HIP(ctx, hipMemsetAsync(fcnt, 0, sizeof(uint32_t) * (n + 4), st));
    GLS_HIP(ctx, hipMemsetAsync(fstart, 0, sizeof(uint32_t) * (n + 4), st));
    GLS_HIP(ctx, hipMemsetAsync(d_misc, 0, sizeof(uint32_t) * 8, st));
    GLS_HIP(ctx, hipMemsetAsync(pending, 0, sizeof(uint32_t) * n, st));
    k_kahn_count<<<gN, kBlock, 0, st>>>(n, nPIs, in.f0, in.f1, dead, cnt, pending);
    size_t tb = 0;
    GLS_HIP(ctx, rocprim::exclusive_scan(nullptr, tb, cnt, adjOff, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(), st));
    void *tmp = t.get<uint8_t>(tb);
    if (!tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: scan storage");
    GLS_HIP(ctx, rocprim::exclusive_scan(tmp, tb, cnt, adjOff, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(), st));
    GLS_HIP(ctx, hipMemcpyAsync(cursor, adjOff, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToDevice, st));
    k_kahn_fill<<<gN, kBlock, 0, st>>>(n, nPIs, in.f0, in.f1, dead, cnt, cursor, adj, adjY);
    k_zero_flags<<<gN, kBlock, 0, st>>>(n, nPIs, pending, dead, zflag);
    GLS_TRY(select_ids(ctx, t, n, zflag, order, &fcnt[1]));
    unsigned long long *d_probes = reinterpret_cast<unsigned long long *>(&d_misc[4]);
    k_preinsert<<<gN, kBlock, 0, st>>>(n, nPIs, in.f0, in.f1, dead, ht, cap - 1, slotTmp, &d_misc[1], d_probes);
    k_canon<<<gN, kBlock, 0, st>>>(n, nPIs, dead, ht, slotTmp, locOf);
    k_node_recs<<<gN, kBlock, 0, st>>>(n, nPIs, in.f0, in.f1, dead, locOf, adjOff, adj, adjY, recs, ofls);
    GLS_HIP(ctx, hipGetLastError());

    // 3. wave sweep: narrow waves run back to back inside one workgroup, wide
    //    ones as one grid-wide kernel each
    WaveArgs a{in.f0, in.f1, nPIs, rl, locOf, wave, ht, val, cap - 1, adjOff, adj, pending, order,
               fcnt, fstart, &d_misc[0], &d_misc[1], d_probes};
    const uint32_t kThresh = WL_FM;
    uint32_t k = 1, waves = 0, cur = 0;
    for (;;) {
        GLS_TRY(stream_copy(ctx, &cur, &fcnt[k], sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (cur == 0) break;
        if (cur <= kThresh) {
            k_waves_1wg<<<1, WG1, 0, st>>>(k, kThresh, a, recs, ofls, adjY, &d_misc[2], n);
            GLS_HIP(ctx, hipGetLastError());
            uint32_t kn = k;
            GLS_TRY(stream_copy(ctx, &kn, &d_misc[2], sizeof(uint32_t), hipMemcpyDeviceToHost));
            if (kn == k) {
                // the frontier's fanout edges do not fit the workgroup: one grid-wide wave
                k_wave_multi<<<sweep_grid(cur), kBlock, 0, st>>>(k, a);
                GLS_HIP(ctx, hipGetLastError());
                ++k;
                ++waves;
                continue;
            }
            waves += kn - k;
            k = kn;
        } else {
            for (int b = 0; b < tune.strash_wave_batch; b++, k++, waves++)
                k_wave_multi<<<sweep_grid(cur), kBlock, 0, st>>>(k, a);
            GLS_HIP(ctx, hipGetLastError());
        }
    }
    uint32_t misc[8] = {};
    GLS_TRY(stream_copy(ctx, misc, d_misc, sizeof(misc), hipMemcpyDeviceToHost));
    if (misc[1]) return gls_fail(ctx, GLS_ERR_INTERNAL, "strash table overflow");
    const int32_t depth = (int32_t)misc[0];
    uint32_t processed = 0;
    GLS_TRY(stream_copy(ctx, &processed, &fstart[depth + 1], sizeof(uint32_t), hipMemcpyDeviceToHost));
    const uint32_t nAnds = n > nPIs + 1 ? (uint32_t)(n - nPIs - 1) : 0u;
    if (processed + nDead != nAnds)
        return gls_fail(ctx, GLS_ERR_ARG, "strash: %u of %u AND nodes unreachable in level order (cycle)",
                        nAnds - nDead - processed, nAnds - nDead);

    if (processed > 0)
        k_wave_of_order<<<grid_for(processed), kBlock, 0, st>>>(processed, (uint32_t)depth, fstart, order, wave);
    if (tagIn && tagDepth) k_tag_depth<<<tag_grid(n), kBlock, 0, st>>>(n, nPIs, wave, tagIn, tagDepth);
    // 4. survivors in (wave, id) order
    k_resolve_all<<<gN, kBlock, 0, st>>>(n, nPIs, dead, rl, locOf, ht, val, surv);
    GLS_HIP(ctx, hipGetLastError());
    size_t tb2 = 0;
    rocprim::counting_iterator<uint32_t> ids(0u);
    GLS_HIP(ctx, rocprim::select(nullptr, tb2, ids, surv, survIds, d_count, (size_t)n, st));
    void *tmp2 = t.get<uint8_t>(tb2);
    if (!tmp2) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: select storage");
    GLS_HIP(ctx, rocprim::select(tmp2, tb2, ids, surv, survIds, d_count, (size_t)n, st));
    uint32_t m = 0;
    GLS_HIP(ctx, hipMemcpyAsync(&m, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    GLS_HIP(ctx, hipStreamSynchronize(st));
    const int32_t nNew = nPIs + 1 + (int32_t)m;
    GLS_TRY(dev_graph_alloc(ctx, out, nNew, nPIs, nPOs));
    GLS_HIP(ctx, hipMemsetAsync(out.f0, 0, sizeof(uint32_t) * (nPIs + 1), st));
    GLS_HIP(ctx, hipMemsetAsync(out.f1, 0, sizeof(uint32_t) * (nPIs + 1), st));
    GLS_HIP(ctx, hipMemsetAsync(out.level, 0, sizeof(uint32_t) * (nPIs + 1), st));
    const uint32_t *sorted = survIds;
    if (m > 0) {
        k_gather_u32<<<grid_for(m), kBlock, 0, st>>>((int32_t)m, survIds, wave, survWave);
        size_t tb3 = 0;
        uint32_t endBit = bits_for((uint32_t)depth);
        GLS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb3, survWave, sortWave, survIds, sortIds, (size_t)m, 0u, endBit, st));
        void *tmp3 = t.get<uint8_t>(tb3);
        if (!tmp3) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: sort storage");
        GLS_HIP(ctx, rocprim::radix_sort_pairs(tmp3, tb3, survWave, sortWave, survIds, sortIds, (size_t)m, 0u, endBit, st));
        sorted = sortIds;
    }
    uint32_t *tags = nullptr;
    if (tagIn && tagOut) {
        tags = pool_get<uint32_t>(ctx, (size_t)nNew);
        if (!tags) { dev_graph_free(ctx, out); return gls_fail(ctx, GLS_ERR_NOMEM, "strash: tags"); }
        GLS_HIP(ctx, hipMemcpyAsync(tags, tagIn, sizeof(uint32_t) * (nPIs + 1), hipMemcpyDeviceToDevice, st));
        *tagOut = tags;
    }
    k_new_ids<<<grid_for(m + nPIs + 1), kBlock, 0, st>>>(n, nPIs, (int32_t)m, sorted, newId);
    if (m > 0)
        k_emit<<<grid_for(m), kBlock, 0, st>>>((int32_t)m, nPIs, sorted, in.f0, in.f1, rl, newId,
                                                out.f0, out.f1, out.level, tagIn, tags);
    if (nPOs > 0) k_map_pos<<<grid_for(nPOs), kBlock, 0, st>>>(nPOs, nPIs, in.pos, rl, newId, out.pos);
    GLS_HIP(ctx, hipGetLastError());
    GLS_HIP(ctx, hipStreamSynchronize(st));
    if (depthOut) *depthOut = depth;
    if (info) {
        info->levels_enumerated = waves;
        info->hash_probes += (int64_t)((uint64_t)misc[4] | (uint64_t)misc[5] << 32);
    }
    return GLS_OK;
}

// Strash of the graph a rewrite composed (comb: the rewrite's input graph g,
// ids 0..n-1, with the selected roots' fanouts redirected, plus the new nodes
// n..N-1).  The result equals dev_strash(comb) array for array; it is computed
// from what changed instead of a wave-by-wave sweep:
//  1. dangling nodes, as in dev_strash;
//  2. waves (levels over the live nodes) by one chunked sweep over a
//     topological order of comb, sorted by a key relaxed from g's levels;
//  3. representatives: g is strashed, so a live old node whose fanins are
//     unchanged and are their own representatives keeps its key, which no
//     other such node has.  Only the other nodes (new nodes, redirected old
//     nodes, and what their merges reach) are hashed, in rounds: each round
//     hashes the dirty set over the previous round's representatives into a
//     fresh table and compares each key with its clean holder (found in g's
//     structural table); a representative that changes pulls its fanouts into
//     the dirty set, a clean holder that loses on (wave, id) joins it.  The
//     rounds stop when nothing changes: then every node's representative is
//     the smallest (wave, id) over the nodes of its key with fanins over
//     representatives, which is the fixpoint the wave sweep reaches (it is
//     unique: by induction over waves, each wave's keys and minima follow
//     from the waves before);
//  4. survivors in (wave, id) order, fanins over representatives, and the
//     output levels by the same chunked sweep over the output.
// Where a step cannot run (a position order that is not topological, a table
// or dirty-set overflow, too many rounds) it falls back to dev_strash(comb);
// info->final_strash tells which path ran.
int dev_strash_rw(gls_ctx *ctx, const DevGraph &in, const RwDelta &d, DevGraph &out, int32_t *depthOut,
                  gls_op_info *info, const uint32_t *tagIn, uint32_t **tagOut, uint32_t *tagDepth) {
    const int32_t N = in.nObjs, nPIs = in.nPIs, nPOs = in.nPOs, n = d.n;
    hipStream_t st = ctx->stream;
    // n == 0: a whole graph, every live node dirty, its levels by one sweep
    // over id order (ids must be topological, as AIGER's are)
    const bool full = n == 0;
    if (!full && (!d.g0 || !d.g1 || !d.glevel || !d.tOff || !d.tEnt || n > N))
        return gls_fail(ctx, GLS_ERR_INTERNAL, "strash: the rewrite's input graph is not available");
    auto fallback = [&](const char *why) {
        (void)why;
        const int rc = dev_strash(ctx, in, out, depthOut, info, tagIn, tagOut, tagDepth);
        if (info) info->final_strash = 2;
        return rc;
    };
    // The level sweeps below run in one workgroup, about 3 ns per object for
    // the two of them; the wave sweep costs about 2.6 us per level plus its
    // per-node work.  On wide graphs (many objects per level, as in a batched
    // arena) the wave sweep is the faster one.
    if (!full && (int64_t)N > (int64_t)tuning_for(ctx, d.n).rw_sweep_width * (int64_t)std::max(d.depth, 1)) {
        const int rc = dev_strash(ctx, in, out, depthOut, info, tagIn, tagOut, tagDepth);
        if (info) info->final_strash = 0;
        return rc;
    }
    {
        Tmp t{ctx};
        uint32_t *ref = t.get<uint32_t>(N), *order = t.get<uint32_t>(N);
        uint8_t *dead = t.get<uint8_t>(N), *surv = t.get<uint8_t>(N);
        uint32_t *posOf = t.get<uint32_t>(N), *fa = t.get<uint32_t>(N), *fb = t.get<uint32_t>(N);
        uint32_t *levPos = t.get<uint32_t>(N), *wave = t.get<uint32_t>(N), *rep = t.get<uint32_t>(N);
        uint32_t *inD = t.get<uint32_t>(N), *dlist = t.get<uint32_t>(N);
        uint32_t *survIds = t.get<uint32_t>(N), *survWave = t.get<uint32_t>(N);
        uint32_t *sortIds = t.get<uint32_t>(N), *sortWave = t.get<uint32_t>(N), *newId = t.get<uint32_t>(N);
        uint32_t *d_misc = t.get<uint32_t>(16);   // [0] err, [1] max level, [2] dirty count, [3] changed,
                                                  // [4] survivors, [6..7] probes, [8] a path got shorter,
                                                  // [9] a key is loose, [10] largest key
        SweepCtl *ctl = t.get<SweepCtl>(1);
        if (!ref || !order || !dead || !surv || !posOf || !fa || !fb || !levPos || !wave || !rep || !inD || !dlist ||
            !survIds || !survWave || !sortIds || !sortWave || !newId || !d_misc || !ctl)
            return gls_fail(ctx, GLS_ERR_NOMEM, "strash: device allocation failed (n=%d)", N);
        unsigned long long *d_probes = reinterpret_cast<unsigned long long *>(&d_misc[6]);
        GLS_HIP(ctx, hipMemsetAsync(ref, 0, sizeof(uint32_t) * N, st));
        GLS_HIP(ctx, hipMemsetAsync(dead, 0, N, st));
        GLS_HIP(ctx, hipMemsetAsync(ctl, 0, sizeof(SweepCtl), st));
        GLS_HIP(ctx, hipMemsetAsync(d_misc, 0, sizeof(uint32_t) * 16, st));
        const int gN = grid_for(N);
        SweepCtl h_ctl{};
        uint32_t rounds = 0;
        // 1. dangling cascade (its first frontier, the ANDs without fanout, by
        //    a compaction of flags)
        k_count_refs<<<gN, kBlock, 0, st>>>(N, nPIs, in.f0, in.f1, nPOs, in.pos, ref);
        k_zero_flags<<<gN, kBlock, 0, st>>>(N, nPIs, ref, nullptr, surv);
        GLS_HIP(ctx, hipGetLastError());
        GLS_TRY(select_ids(ctx, t, N, surv, order, &ctl->curSize));
        GLS_HIP(ctx, hipMemcpyAsync(&h_ctl, ctl, sizeof(SweepCtl), hipMemcpyDeviceToHost, st));
        GLS_HIP(ctx, hipStreamSynchronize(st));
        if (h_ctl.curSize > 0) {
            GLS_TRY(run_sweep(ctx, ctl, 1, [&](uint32_t k, uint32_t hint) {
                k_dangle_round<<<sweep_grid(hint), kBlock, 0, st>>>(k, ctl, nPIs, in.f0, in.f1, ref, dead, order);
            }, &h_ctl, &rounds));
        }
        // 2. a topological order (keys relaxed from the input levels, then a
        //    stable sort), and the waves by one chunked sweep over it; for a
        //    whole graph the ids themselves
        uint32_t misc[16] = {};
        bool sweepIn = true;
        if (full) {
            // positions: the constant and the PIs, then the live ANDs in id
            // order (inD is free until the dirty set is seeded)
            k_keep_flags<<<gN, kBlock, 0, st>>>(N, nPIs, dead, inD);
            size_t tbp = 0;
            GLS_HIP(ctx, rocprim::exclusive_scan(nullptr, tbp, inD, posOf, 0u, (size_t)N, rocprim::plus<uint32_t>(), st));
            void *tmpp = t.get<uint8_t>(std::max<size_t>(tbp, 1));
            if (!tmpp) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: position scan storage");
            GLS_HIP(ctx, rocprim::exclusive_scan(tmpp, tbp, inD, posOf, 0u, (size_t)N, rocprim::plus<uint32_t>(), st));
            k_full_fanpos<<<gN, kBlock, 0, st>>>(N, nPIs, in.f0, in.f1, dead, posOf, fa, fb);
            GLS_HIP(ctx, hipGetLastError());
            uint32_t lastPos = 0, lastKeep = 0;
            GLS_TRY(stream_copy(ctx, &lastPos, posOf + (N - 1), sizeof(uint32_t), hipMemcpyDeviceToHost));
            GLS_TRY(stream_copy(ctx, &lastKeep, inD + (N - 1), sizeof(uint32_t), hipMemcpyDeviceToHost));
            const uint32_t nPos = lastPos + lastKeep;
            k_level_sweep<<<1, SW_WG, 0, st>>>(nPos, fa, fb, levPos, &d_misc[0], &d_misc[1]);
            GLS_HIP(ctx, hipGetLastError());
            GLS_TRY(stream_copy(ctx, misc, d_misc, sizeof(misc), hipMemcpyDeviceToHost));
            if (misc[0]) return fallback("ids not topological");
            // wide graphs are faster by the grid-wide wave sweep
            if ((int64_t)N > (int64_t)tuning_for(ctx, N).rw_sweep_width * (int64_t)std::max<uint32_t>(misc[1], 1u)) {
                const int rc = dev_strash(ctx, in, out, depthOut, info, tagIn, tagOut, tagDepth);
                if (info) info->final_strash = 0;
                return rc;
            }
        } else {
        k_rw_key_init<<<gN, kBlock, 0, st>>>(n, N, d.glevel, fa);
        uint32_t keyRounds = 0;
        for (;; ++keyRounds) {
            if (keyRounds >= kRwRoundLimit) return fallback("key rounds");
            GLS_HIP(ctx, hipMemsetAsync(&d_misc[5], 0, sizeof(uint32_t), st));
            k_rw_key_relax<<<gN, kBlock, 0, st>>>(N, nPIs, in.f0, in.f1, dead, fa, &d_misc[5]);
            GLS_HIP(ctx, hipGetLastError());
            uint32_t moved = 0;
            GLS_TRY(stream_copy(ctx, &moved, &d_misc[5], sizeof(uint32_t), hipMemcpyDeviceToHost));
            if (moved == 0) break;
        }
        {
            // order: ids by key, stable (old ids, then new ids, within a key)
            k_iota_u32<<<gN, kBlock, 0, st>>>(N, levPos);
            size_t tbk = 0;
            GLS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tbk, fa, fb, levPos, order, (size_t)N, 0u, 32u, st));
            void *tmpk = t.get<uint8_t>(std::max<size_t>(tbk, 1));
            if (!tmpk) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: key sort storage");
            GLS_HIP(ctx, rocprim::radix_sort_pairs(tmpk, tbk, fa, fb, levPos, order, (size_t)N, 0u, 32u, st));
        }
        k_rw_pos<<<gN, kBlock, 0, st>>>(N, order, posOf);
        // the waves: the keys themselves where every live key is tight, else a
        // level sweep over the positions
        k_rw_key_loose<<<gN, kBlock, 0, st>>>(N, nPIs, in.f0, in.f1, dead, fa, &d_misc[9], &d_misc[10]);
        GLS_HIP(ctx, hipGetLastError());
        GLS_TRY(stream_copy(ctx, misc, d_misc, sizeof(misc), hipMemcpyDeviceToHost));
        sweepIn = misc[9] != 0u;
        if (sweepIn) {
            k_rw_fanpos<<<gN, kBlock, 0, st>>>(N, nPIs, in.f0, in.f1, dead, posOf, fa, fb, &d_misc[0]);
            GLS_HIP(ctx, hipGetLastError());
            GLS_TRY(stream_copy(ctx, misc, d_misc, sizeof(misc), hipMemcpyDeviceToHost));
            if (misc[0]) return fallback("positions");
            k_level_sweep<<<1, SW_WG, 0, st>>>((uint32_t)N, fa, fb, levPos, &d_misc[0], &d_misc[1]);
            GLS_HIP(ctx, hipGetLastError());
            GLS_TRY(stream_copy(ctx, misc, d_misc, sizeof(misc), hipMemcpyDeviceToHost));
            if (misc[0]) return fallback("sweep order");
        } else {
            k_rw_levpos<<<gN, kBlock, 0, st>>>(N, nPIs, dead, fa, posOf, levPos);
            GLS_HIP(ctx, hipGetLastError());
            misc[1] = misc[10];
        }
        }
        const int32_t depth = (int32_t)misc[1];
        k_rw_seed<<<gN, kBlock, 0, st>>>(n, N, nPIs, in.f0, in.f1, d.g0, d.g1, dead, posOf, levPos, wave, rep, inD);
        GLS_HIP(ctx, hipGetLastError());
        {
            rocprim::counting_iterator<uint32_t> ids(0u);
            size_t tbd = 0;
            GLS_HIP(ctx, rocprim::select(nullptr, tbd, ids, inD, dlist, &d_misc[2], (size_t)N, st));
            void *tmpd = t.get<uint8_t>(std::max<size_t>(tbd, 1));
            if (!tmpd) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: dirty list storage");
            GLS_HIP(ctx, rocprim::select(tmpd, tbd, ids, inD, dlist, &d_misc[2], (size_t)N, st));
        }
        if (tagIn && tagDepth) k_tag_depth<<<tag_grid(N), kBlock, 0, st>>>(N, nPIs, wave, tagIn, tagDepth);
        // 3. representatives of the dirty set, in rounds
        uint32_t nD = 0;
        GLS_TRY(stream_copy(ctx, &nD, &d_misc[2], sizeof(uint32_t), hipMemcpyDeviceToHost));
        uint64_t cap = 0;
        unsigned long long *ht = nullptr;
        uint2 *dkey = nullptr;
        uint32_t *dslot = nullptr, *nrep = nullptr;
        uint32_t prevD = 0, r = 0;
        for (; nD > 0; ++r) {
            if (r >= kRwRoundLimit) return fallback("rounds");
            uint64_t want = 1024;
            while (want < 4ull * nD) want <<= 1;
            if (want > cap) {
                cap = want;
                ht = t.get<unsigned long long>(2 * cap);
                dkey = t.get<uint2>(cap / 2);
                dslot = t.get<uint32_t>(cap / 2);
                nrep = t.get<uint32_t>(cap / 2);
                if (!ht || !dkey || !dslot || !nrep) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: dirty table");
            }
            k_fill_u64<<<grid_for((int64_t)2 * cap), kBlock, 0, st>>>(2 * cap, ht, KEY_EMPTY);
            GLS_HIP(ctx, hipMemsetAsync(&d_misc[3], 0, sizeof(uint32_t), st));
            const int gD = grid_for(nD);
            k_rw_insert<<<gD, kBlock, 0, st>>>(nD, dlist, nPIs, in.f0, in.f1, rep, wave, ht, cap - 1, dkey, dslot,
                                               nrep, &d_misc[0], d_probes);
            k_rw_resolve<<<gD, kBlock, 0, st>>>(nD, dlist, n, dead, wave, rep, inD, ht, dkey, dslot, nrep, d.tOff,
                                                d.tEnt, dlist, &d_misc[2], (uint32_t)N, &d_misc[3], &d_misc[0],
                                                d_probes);
            k_rw_pull<<<gN, kBlock, 0, st>>>(N, nPIs, in.f0, in.f1, dead, rep, inD, dlist, &d_misc[2], (uint32_t)N,
                                             &d_misc[0]);
            GLS_HIP(ctx, hipGetLastError());
            GLS_TRY(stream_copy(ctx, misc, d_misc, sizeof(uint32_t) * 4, hipMemcpyDeviceToHost));
            if (misc[0]) return fallback("dirty table");
            prevD = nD;
            nD = misc[2];
            if (misc[3] == 0 && nD == prevD) break;
        }
        // 4. survivors in (wave, id) order
        k_rw_shorter<<<gN, kBlock, 0, st>>>(N, nPIs, dead, rep, wave, &d_misc[8]);
        k_rw_survivors<<<gN, kBlock, 0, st>>>(N, nPIs, dead, rep, surv);
        size_t tb = 0;
        rocprim::counting_iterator<uint32_t> ids(0u);
        GLS_HIP(ctx, rocprim::select(nullptr, tb, ids, surv, survIds, &d_misc[4], (size_t)N, st));
        void *tmp = t.get<uint8_t>(tb);
        if (!tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: select storage");
        GLS_HIP(ctx, rocprim::select(tmp, tb, ids, surv, survIds, &d_misc[4], (size_t)N, st));
        GLS_TRY(stream_copy(ctx, misc, d_misc, sizeof(misc), hipMemcpyDeviceToHost));
        const uint32_t m = misc[4];
        const int32_t nNew = nPIs + 1 + (int32_t)m;
        GLS_TRY(dev_graph_alloc(ctx, out, nNew, nPIs, nPOs));
        const uint32_t *sorted = survIds;
        if (m > 0) {
            k_gather_u32<<<grid_for(m), kBlock, 0, st>>>((int32_t)m, survIds, wave, survWave);
            size_t tb3 = 0;
            const uint32_t endBit = bits_for((uint32_t)depth);
            GLS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb3, survWave, sortWave, survIds, sortIds, (size_t)m, 0u,
                                                   endBit, st));
            void *tmp3 = t.get<uint8_t>(tb3);
            if (!tmp3) { dev_graph_free(ctx, out); return gls_fail(ctx, GLS_ERR_NOMEM, "strash: sort storage"); }
            GLS_HIP(ctx, rocprim::radix_sort_pairs(tmp3, tb3, survWave, sortWave, survIds, sortIds, (size_t)m, 0u,
                                                   endBit, st));
            sorted = sortIds;
        }
        uint32_t *tags = nullptr;
        if (tagIn && tagOut) {
            tags = pool_get<uint32_t>(ctx, (size_t)nNew);
            if (!tags) { dev_graph_free(ctx, out); return gls_fail(ctx, GLS_ERR_NOMEM, "strash: tags"); }
            GLS_HIP(ctx, hipMemcpyAsync(tags, tagIn, sizeof(uint32_t) * (nPIs + 1), hipMemcpyDeviceToDevice, st));
            *tagOut = tags;
        }
        GLS_HIP(ctx, hipMemsetAsync(out.f0, 0, sizeof(uint32_t) * (nPIs + 1), st));
        GLS_HIP(ctx, hipMemsetAsync(out.f1, 0, sizeof(uint32_t) * (nPIs + 1), st));
        k_new_ids<<<grid_for(m + nPIs + 1), kBlock, 0, st>>>(N, nPIs, (int32_t)m, sorted, newId);
        // 5. output levels: the survivors' waves when no representative lies on
        // a lower wave than its node, else a level sweep over the output (fa /
        // fb are free again: they take the output's fanin ids)
        const bool sweepOut = misc[8] != 0u;
        k_rw_emit<<<grid_for(nNew), kBlock, 0, st>>>((int32_t)m, nPIs, sorted, in.f0, in.f1, rep, newId, out.f0,
                                                     out.f1, fa, fb, wave, sweepOut ? nullptr : out.level, tagIn,
                                                     tags);
        if (nPOs > 0) k_rw_pos_out<<<grid_for(nPOs), kBlock, 0, st>>>(nPOs, nPIs, in.pos, rep, newId, out.pos);
        GLS_HIP(ctx, hipGetLastError());
        if (sweepOut) {
            GLS_HIP(ctx, hipMemsetAsync(&d_misc[0], 0, sizeof(uint32_t) * 2, st));
            k_level_sweep<<<1, SW_WG, 0, st>>>((uint32_t)nNew, fa, fb, out.level, &d_misc[0], &d_misc[1]);
            GLS_HIP(ctx, hipGetLastError());
            GLS_TRY(stream_copy(ctx, misc, d_misc, sizeof(misc), hipMemcpyDeviceToHost));
            if (misc[0]) {
                dev_graph_free(ctx, out);
                if (tags) { pool_put(ctx, tags); *tagOut = nullptr; }
                return gls_fail(ctx, GLS_ERR_INTERNAL, "strash: output not topological");
            }
        }
        // the depth reported is the number of waves of the input (its live
        // nodes' largest level), as the reference's strash counts it, not
        // the output's largest level
        if (depthOut) *depthOut = depth;
        if (info) {
            info->levels_enumerated = depth;
            info->hash_probes += (int64_t)((uint64_t)misc[6] | (uint64_t)misc[7] << 32);
            info->final_strash = 1;
            info->strash_rounds = (int64_t)r + 1;
            info->strash_sweeps = (sweepIn ? 1 : 0) + (sweepOut ? 1 : 0);
        }
    }
    return GLS_OK;
}
