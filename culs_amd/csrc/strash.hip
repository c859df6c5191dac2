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

__global__ void k_seed_zero(int32_t n, int32_t nPIs, const uint32_t *__restrict__ cnt, SweepCtl *ctl,
                            uint32_t *order, const uint8_t *__restrict__ dead) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + tid; v < n; v += stride) {
        if (dead && dead[v]) continue;
        if (cnt[v] == 0) order[atomicAdd(&ctl->curSize, 1u)] = (uint32_t)v;
    }
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

__global__ void k_kahn_count(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0,
                             const uint32_t *__restrict__ f1, const uint8_t *__restrict__ dead,
                             uint32_t *cnt, uint32_t *pending) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + tid; v < n; v += stride) {
        if (dead[v]) continue;
        uint32_t a = LID(f0[v]), b = LID(f1[v]);
        uint32_t p = 0;
        if (a > (uint32_t)nPIs) { atomicAdd(&cnt[a], 1u); ++p; }
        if (b > (uint32_t)nPIs) { atomicAdd(&cnt[b], 1u); ++p; }
        pending[v] = p;
    }
}

__global__ void k_kahn_fill(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0,
                            const uint32_t *__restrict__ f1, const uint8_t *__restrict__ dead,
                            uint32_t *cursor, uint32_t *adj) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + tid; v < n; v += stride) {
        if (dead[v]) continue;
        uint32_t a = LID(f0[v]), b = LID(f1[v]);
        if (a > (uint32_t)nPIs) adj[atomicAdd(&cursor[a], 1u)] = (uint32_t)v;
        if (b > (uint32_t)nPIs) adj[atomicAdd(&cursor[b], 1u)] = (uint32_t)v;
    }
}

struct StrashArgs {
    const uint32_t *f0, *f1;
    int32_t nPIs;
    uint32_t *res;                  // literal over representative ids, INVALID until resolved
    unsigned long long *keyOf;      // key of non-trivial nodes
    uint32_t *level, *wave;
    unsigned long long *htK, *htV;
    uint64_t mask;
    const uint32_t *adjOff, *adj;
    uint32_t *pending, *order;
};

__device__ __forceinline__ uint32_t map_lit(const StrashArgs &a, uint32_t lit) {
    uint32_t u = LID(lit);
    if (u <= (uint32_t)a.nPIs) return lit;
    uint32_t r = a.res[u];
    if (r == INVALID32) {
        // u belongs to the previous wave: its representative is final in the table
        unsigned long long val = 0;
        ht_find(a.htK, a.htV, a.mask, a.keyOf[u], &val);
        r = 2u * (uint32_t)(val & 0xFFFFFFFFull);
    }
    return r ^ LC(lit);
}

// One wave: resolve the previous wave's representatives, hash the current
// wave, and release the fanouts whose last fanin was in this wave.
__global__ void k_strash_round(uint32_t k, SweepCtl *ctl, StrashArgs a) {
    uint32_t ps = ctl->prevStart, pn = ctl->prevSize, cs = ctl->curStart, cn = ctl->curSize;
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (uint32_t i = tid; i < pn; i += stride) {
        uint32_t v = a.order[ps + i];
        if (a.res[v] == INVALID32) {
            unsigned long long val = 0;
            ht_find(a.htK, a.htV, a.mask, a.keyOf[v], &val);
            a.res[v] = 2u * (uint32_t)(val & 0xFFFFFFFFull);
        }
    }
    for (uint32_t i = tid; i < cn; i += stride) {
        uint32_t v = a.order[cs + i];
        uint32_t l0 = map_lit(a, a.f0[v]), l1 = map_lit(a, a.f1[v]);
        if (l0 > l1) { uint32_t t = l0; l0 = l1; l1 = t; }
        a.wave[v] = k;
        uint32_t r = INVALID32;
        if (l0 == l1) r = l0;
        else if (l0 == (l1 ^ 1u)) r = 0;
        else if (l0 == 1u) r = l1;
        else if (l0 == 0u) r = 0;
        if (r != INVALID32) {
            a.res[v] = r;
            uint32_t u = LID(r);
            a.level[v] = u <= (uint32_t)a.nPIs ? 0u : a.level[u];
        } else {
            uint64_t key = and_key(l0, l1);
            a.keyOf[v] = key;
            uint32_t la = LID(l0) <= (uint32_t)a.nPIs ? 0u : a.level[LID(l0)];
            uint32_t lb = LID(l1) <= (uint32_t)a.nPIs ? 0u : a.level[LID(l1)];
            a.level[v] = 1u + (la > lb ? la : lb);
            ht_insert_min(a.htK, a.htV, a.mask, key, (static_cast<uint64_t>(k) << 32) | v, &ctl->err);
        }
        for (uint32_t e = a.adjOff[v], ee = a.adjOff[v + 1]; e < ee; ++e) {
            uint32_t w = a.adj[e];
            if (atomicSub(&a.pending[w], 1u) == 1u) a.order[cs + cn + atomicAdd(&ctl->nextSize, 1u)] = w;
        }
    }
    sweep_round_end(ctl, k);
}

__global__ void k_resolve_all(int32_t n, int32_t nPIs, const uint8_t *__restrict__ dead, uint32_t *res,
                              const unsigned long long *__restrict__ keyOf,
                              const unsigned long long *__restrict__ htK,
                              const unsigned long long *__restrict__ htV, uint64_t mask, uint8_t *surv) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < n; v += stride) {
        uint8_t s = 0;
        if (v > nPIs && !dead[v]) {
            uint32_t r = res[v];
            if (r == INVALID32) {
                unsigned long long val = 0;
                ht_find(htK, htV, mask, keyOf[v], &val);
                r = 2u * (uint32_t)(val & 0xFFFFFFFFull);
                res[v] = r;
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

__device__ __forceinline__ uint32_t final_lit(uint32_t lit, int32_t nPIs, const uint32_t *res,
                                              const uint32_t *newId) {
    uint32_t u = LID(lit);
    uint32_t r = (u <= (uint32_t)nPIs) ? 2u * u : res[u];
    r ^= LC(lit);
    return 2u * newId[LID(r)] | LC(r);
}

__global__ void k_emit(int32_t m, int32_t nPIs, const uint32_t *__restrict__ sorted,
                       const uint32_t *__restrict__ f0, const uint32_t *__restrict__ f1,
                       const uint32_t *__restrict__ res, const uint32_t *__restrict__ newId,
                       const uint32_t *__restrict__ level, uint32_t *of0, uint32_t *of1, uint32_t *olevel) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = tid; p < m; p += stride) {
        uint32_t v = sorted[p];
        uint32_t a = final_lit(f0[v], nPIs, res, newId), b = final_lit(f1[v], nPIs, res, newId);
        if (a > b) { uint32_t t = a; a = b; b = t; }
        int64_t o = nPIs + 1 + p;
        of0[o] = a;
        of1[o] = b;
        olevel[o] = level[v];
    }
}

__global__ void k_map_pos(int32_t nPOs, int32_t nPIs, const uint32_t *__restrict__ pos,
                          const uint32_t *__restrict__ res, const uint32_t *__restrict__ newId, uint32_t *opos) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tid; i < nPOs; i += stride) opos[i] = final_lit(pos[i], nPIs, res, newId);
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

int dev_strash(gls_ctx *ctx, const DevGraph &in, DevGraph &out, int32_t *depthOut, gls_op_info *info) {
    const int32_t n = in.nObjs, nPIs = in.nPIs, nPOs = in.nPOs;
    hipStream_t st = ctx->stream;
    Tmp t{ctx};
    uint32_t *ref = t.get<uint32_t>(n), *order = t.get<uint32_t>(n), *pending = t.get<uint32_t>(n);
    uint32_t *cnt = t.get<uint32_t>(n + 1), *adjOff = t.get<uint32_t>(n + 1), *cursor = t.get<uint32_t>(n + 1);
    uint32_t *adj = t.get<uint32_t>(2 * (size_t)n + 1), *res = t.get<uint32_t>(n);
    uint32_t *level = t.get<uint32_t>(n), *wave = t.get<uint32_t>(n), *newId = t.get<uint32_t>(n);
    uint32_t *survIds = t.get<uint32_t>(n), *survWave = t.get<uint32_t>(n);
    uint32_t *sortIds = t.get<uint32_t>(n), *sortWave = t.get<uint32_t>(n), *d_count = t.get<uint32_t>(4);
    uint8_t *dead = t.get<uint8_t>(n), *surv = t.get<uint8_t>(n);
    unsigned long long *keyOf = t.get<unsigned long long>(n);
    SweepCtl *ctl = t.get<SweepCtl>(1);
    uint64_t cap = 1024;
    while (cap < 2ull * (uint64_t)n) cap <<= 1;
    unsigned long long *htK = t.get<unsigned long long>(cap), *htV = t.get<unsigned long long>(cap);
    if (!ref || !order || !pending || !cnt || !adjOff || !cursor || !adj || !res || !level || !wave ||
        !newId || !survIds || !survWave || !sortIds || !sortWave || !d_count || !dead || !surv || !keyOf ||
        !ctl || !htK || !htV)
        return gls_fail(ctx, GLS_ERR_NOMEM, "strash: device allocation failed (n=%d)", n);

    GLS_HIP(ctx, hipMemsetAsync(ref, 0, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(dead, 0, n, st));
    GLS_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(uint32_t) * (n + 1), st));
    GLS_HIP(ctx, hipMemsetAsync(res, 0xFF, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(level, 0, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(wave, 0, sizeof(uint32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(htK, 0xFF, sizeof(unsigned long long) * cap, st));
    GLS_HIP(ctx, hipMemsetAsync(htV, 0xFF, sizeof(unsigned long long) * cap, st));
    GLS_HIP(ctx, hipMemsetAsync(ctl, 0, sizeof(SweepCtl), st));
    const int gN = grid_for(n);
    SweepCtl h_ctl{};
    uint32_t rounds = 0;

    // 1. fanout counts and dangling cascade
    k_count_refs<<<gN, kBlock, 0, st>>>(n, nPIs, in.f0, in.f1, nPOs, in.pos, ref);
    k_seed_zero<<<gN, kBlock, 0, st>>>(n, nPIs, ref, ctl, order, nullptr);
    GLS_HIP(ctx, hipGetLastError());
    GLS_HIP(ctx, hipMemcpyAsync(&h_ctl, ctl, sizeof(SweepCtl), hipMemcpyDeviceToHost, st));
    GLS_HIP(ctx, hipStreamSynchronize(st));
    if (h_ctl.curSize > 0) {
        GLS_TRY(run_sweep(ctx, ctl, 1, [&](uint32_t k, uint32_t hint) {
            k_dangle_round<<<sweep_grid(hint), kBlock, 0, st>>>(k, ctl, nPIs, in.f0, in.f1, ref, dead, order);
        }, &h_ctl, &rounds));
    }

    const uint32_t nDead = h_ctl.processed;
    // 2. fanout CSR over live AND nodes and the first frontier
    GLS_HIP(ctx, hipMemsetAsync(ctl, 0, sizeof(SweepCtl), st));
    k_kahn_count<<<gN, kBlock, 0, st>>>(n, nPIs, in.f0, in.f1, dead, cnt, pending);
    size_t tb = 0;
    GLS_HIP(ctx, rocprim::exclusive_scan(nullptr, tb, cnt, adjOff, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(), st));
    void *tmp = t.get<uint8_t>(tb);
    if (!tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "strash: scan storage");
    GLS_HIP(ctx, rocprim::exclusive_scan(tmp, tb, cnt, adjOff, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(), st));
    GLS_HIP(ctx, hipMemcpyAsync(cursor, adjOff, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToDevice, st));
    k_kahn_fill<<<gN, kBlock, 0, st>>>(n, nPIs, in.f0, in.f1, dead, cursor, adj);
    k_seed_zero<<<gN, kBlock, 0, st>>>(n, nPIs, pending, ctl, order, dead);
    GLS_HIP(ctx, hipGetLastError());

    // 3. wave sweep
    StrashArgs a{in.f0, in.f1, nPIs, res, keyOf, level, wave, htK, htV, cap - 1, adjOff, adj, pending, order};
    GLS_HIP(ctx, hipMemcpyAsync(&h_ctl, ctl, sizeof(SweepCtl), hipMemcpyDeviceToHost, st));
    GLS_HIP(ctx, hipStreamSynchronize(st));
    uint32_t waves = 0;
    if (h_ctl.curSize > 0) {
        GLS_TRY(run_sweep(ctx, ctl, 1, [&](uint32_t k, uint32_t hint) {
            k_strash_round<<<sweep_grid(hint), kBlock, 0, st>>>(k, ctl, a);
        }, &h_ctl, &waves));
    }
    int32_t depth = (int32_t)h_ctl.depth;
    const uint32_t nAnds = n > nPIs + 1 ? (uint32_t)(n - nPIs - 1) : 0u;
    if (h_ctl.processed + nDead != nAnds)
        return gls_fail(ctx, GLS_ERR_ARG, "strash: %u of %u AND nodes unreachable in level order (cycle)",
                        nAnds - nDead - h_ctl.processed, nAnds - nDead);

    // 4. survivors in (wave, id) order
    k_resolve_all<<<gN, kBlock, 0, st>>>(n, nPIs, dead, res, keyOf, htK, htV, cap - 1, surv);
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
    k_new_ids<<<grid_for(m + nPIs + 1), kBlock, 0, st>>>(n, nPIs, (int32_t)m, sorted, newId);
    if (m > 0)
        k_emit<<<grid_for(m), kBlock, 0, st>>>((int32_t)m, nPIs, sorted, in.f0, in.f1, res, newId, level,
                                                out.f0, out.f1, out.level);
    if (nPOs > 0) k_map_pos<<<grid_for(nPOs), kBlock, 0, st>>>(nPOs, nPIs, in.pos, res, newId, out.pos);
    GLS_HIP(ctx, hipGetLastError());
    GLS_HIP(ctx, hipStreamSynchronize(st));
    if (depthOut) *depthOut = depth;
    if (info) info->levels_enumerated = waves;
    return GLS_OK;
}
