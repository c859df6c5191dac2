// Batched passes over many independent AIGs.
//
// The AIGs are packed into one device arena: a disjoint union whose ids are,
// in order, the constant, every AIG's PIs (AIG by AIG), then every AIG's AND
// nodes (AIG by AIG); CSR offsets (piOff, andOff, poOff) locate each AIG.  One
// pass (strash or rw) runs over the union, carrying every node's AIG index as
// a tag, and the result is split back.  Per AIG the result is bit-identical
// to running the pass on that AIG alone: every stage is either local to a
// node's own cone (cuts, evaluation, plans), or orders nodes by keys whose
// relative order inside one AIG the packing preserves (levels, ids, gains),
// and no structural key (fanin-literal pair) can be shared by two AIGs.
//
// The union has the depth of its deepest member, so a batch takes about as
// many dependent steps as its deepest AIG alone, with every level holding the
// nodes of all AIGs: that is where the throughput comes from.
#include "gls_internal.h"
#include "dev_common.h"
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <algorithm>
#include <chrono>
#include <vector>

namespace {

constexpr int kB = 256;

struct Member {                  // one AIG of the batch, device view
    const uint32_t *f0, *f1, *pos, *level;
    uint32_t *of0, *of1, *opos, *olevel;   // split output
    int32_t nPIs, nObjs, nPOs;
    int32_t piOff, andOff, poOff;          // prefix offsets in the union
    int32_t outAnds, outAndOff;            // split: AND count and offset in the sorted order
};

inline int nblocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kB - 1) / kB, 1 << 20)); }

// index k of the member whose [off[k], off[k+1]) range holds i (off ascending, off[0] = 0)
__device__ __forceinline__ int find_member(const int32_t *off, int count, int64_t i) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int64_t)off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// member literal -> union literal
__device__ __forceinline__ uint32_t to_union(uint32_t lit, const Member &m, int32_t P) {
    const uint32_t id = lit >> 1, c = lit & 1u;
    if (id == 0) return lit;
    if ((int32_t)id <= m.nPIs) return 2u * (uint32_t)(1 + m.piOff + (int32_t)id - 1) | c;
    return 2u * (uint32_t)(1 + P + m.andOff + (int32_t)id - m.nPIs - 1) | c;
}

__global__ void k_pack_ands(int64_t A, int count, int32_t P, const Member *__restrict__ mem,
                            const int32_t *__restrict__ andOffs, uint32_t *uf0, uint32_t *uf1, uint32_t *ulevel,
                            uint32_t *utag) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < A; i += stride) {
        const int k = find_member(andOffs, count, i);
        const Member &m = mem[k];
        const int32_t v = m.nPIs + 1 + (int32_t)(i - m.andOff);
        const int64_t u = 1 + P + i;
        uf0[u] = to_union(m.f0[v], m, P);
        uf1[u] = to_union(m.f1[v], m, P);
        ulevel[u] = m.level ? m.level[v] : 0u;
        utag[u] = (uint32_t)k;
    }
}

__global__ void k_pack_pis(int32_t P, int count, const int32_t *__restrict__ piOffs, uint32_t *uf0, uint32_t *uf1,
                           uint32_t *ulevel, uint32_t *utag) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= P; i += stride) {
        uf0[i] = 0;
        uf1[i] = 0;
        ulevel[i] = 0;
        utag[i] = i == 0 ? 0u : (uint32_t)find_member(piOffs, count, i - 1);
    }
}

__global__ void k_pack_pos(int64_t O, int count, int32_t P, const Member *__restrict__ mem,
                           const int32_t *__restrict__ poOffs, uint32_t *upos) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < O; j += stride) {
        const int k = find_member(poOffs, count, j);
        const Member &m = mem[k];
        upos[j] = to_union(m.pos[j - m.poOff], m, P);
    }
}

// split, step 1: AND counts per member
__global__ void __launch_bounds__(kBlock) k_split_count(int32_t P, int32_t N, const uint32_t *__restrict__ tag,
                                                        uint32_t *cnt, uint32_t *keys, uint32_t *ids) {
    __shared__ uint32_t s_bins[TAG_BINS];
    TagBins tb{s_bins};
    tb.init();
    int64_t b, e;
    tag_chunk(P + 1, N, b, e);
    for (int64_t u = b + threadIdx.x; u < e; u += blockDim.x) {
        const uint32_t k = tag[u];
        tb.add(cnt, k, 1u);
        keys[u - P - 1] = k;
        ids[u - P - 1] = (uint32_t)u;
    }
    tb.flush_add(cnt, 0xFFFFFFFFu);
}

// split, step 2: local ids (ANDs: rank inside their member, in union id order)
__global__ void k_split_ids(int32_t P, int64_t A, int count, const Member *__restrict__ mem,
                            const int32_t *__restrict__ piOffs, const uint32_t *__restrict__ sortedKeys,
                            const uint32_t *__restrict__ sortedIds, uint32_t *localId) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < A; q += stride) {
        const Member &m = mem[sortedKeys[q]];
        localId[sortedIds[q]] = (uint32_t)(m.nPIs + 1 + (int32_t)(q - m.outAndOff));
    }
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u <= P; u += stride)
        localId[u] = u == 0 ? 0u : (uint32_t)(u - mem[find_member(piOffs, count, u - 1)].piOff);
}

__device__ __forceinline__ uint32_t to_local(uint32_t lit, const uint32_t *__restrict__ localId) {
    return 2u * localId[lit >> 1] | (lit & 1u);
}

// split, step 3: emit every member's arrays
__global__ void k_split_emit(int32_t P, int64_t A, int count, const Member *__restrict__ mem,
                             const int32_t *__restrict__ piOffs, const uint32_t *__restrict__ sortedKeys,
                             const uint32_t *__restrict__ sortedIds, const uint32_t *__restrict__ localId,
                             const uint32_t *__restrict__ uf0, const uint32_t *__restrict__ uf1,
                             const uint32_t *__restrict__ ulevel) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < A; q += stride) {
        const Member &m = mem[sortedKeys[q]];
        const uint32_t u = sortedIds[q], li = localId[u];
        uint32_t a = to_local(uf0[u], localId), b = to_local(uf1[u], localId);
        if (a > b) { const uint32_t t = a; a = b; b = t; }
        m.of0[li] = a;
        m.of1[li] = b;
        m.olevel[li] = ulevel[u];
    }
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u <= P + count; u += stride) {
        // PI entries, then each member's constant entry
        if (u >= 1 && u <= P) {
            const Member &m = mem[find_member(piOffs, count, u - 1)];
            const uint32_t li = localId[u];
            m.of0[li] = 0;
            m.of1[li] = 0;
            m.olevel[li] = 0;
        } else if (u > P) {
            const Member &m = mem[u - P - 1];
            m.of0[0] = 0;
            m.of1[0] = 0;
            m.olevel[0] = 0;
        }
    }
}

__global__ void k_split_pos(int64_t O, int count, const Member *__restrict__ mem, const int32_t *__restrict__ poOffs,
                            const uint32_t *__restrict__ upos, const uint32_t *__restrict__ localId) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < O; j += stride) {
        const Member &m = mem[find_member(poOffs, count, j)];
        m.opos[j - m.poOff] = to_local(upos[j], localId);
    }
}

uint32_t key_bits(uint32_t x) {
    uint32_t b = 1;
    while (b < 32 && (x >> b) != 0) ++b;
    return b;
}

struct PoolBlock {
    gls_ctx *ctx;
    std::vector<void *> ptrs;
    template <class T> T *get(size_t n) {
        T *p = pool_get<T>(ctx, n);
        if (p) ptrs.push_back(p);
        return p;
    }
    ~PoolBlock() {
        for (void *p : ptrs) pool_put(ctx, p);
    }
};

enum class Pass { Strash, Rewrite, Balance };

int batch_pass(gls_ctx *ctx, gls_aig **aigs, int32_t count, Pass pass, int32_t useZeros, int32_t flags) {
    if (!ctx || (count > 0 && !aigs) || count < 0) return gls_fail(ctx, GLS_ERR_ARG, "batch: bad arguments");
    if (count == 0) return GLS_OK;
    if (ctx->device < 0) return gls_fail(ctx, GLS_ERR_NODEV, "host-only context");
    for (int32_t k = 0; k < count; k++) {
        if (!aigs[k]) return gls_fail(ctx, GLS_ERR_ARG, "batch: AIG %d is NULL", k);
        if (aigs[k]->ctx != ctx) return gls_fail(ctx, GLS_ERR_ARG, "batch: AIG %d belongs to another context", k);
    }
    {
        std::vector<gls_aig *> sorted(aigs, aigs + count);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return gls_fail(ctx, GLS_ERR_ARG, "batch: an AIG appears twice");
    }
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = ctx->stream;
    // members on the device; rw needs strashed members with levels
    if (pass == Pass::Rewrite) {
        std::vector<gls_aig *> todo;
        for (int32_t k = 0; k < count; k++) {
            GLS_TRY(aig_ensure_device(aigs[k]));
            if (!aigs[k]->strashed || !aigs[k]->dev_level_valid) todo.push_back(aigs[k]);
        }
        if (!todo.empty()) GLS_TRY(batch_pass(ctx, todo.data(), (int32_t)todo.size(), Pass::Strash, 0, 0));
    }
    std::vector<Member> hm(count);
    std::vector<int32_t> piOffs(count), andOffs(count), poOffs(count);
    int64_t P = 0, A = 0, O = 0;
    for (int32_t k = 0; k < count; k++) {
        gls_aig *a = aigs[k];
        GLS_TRY(aig_ensure_device(a));
        Member &m = hm[k];
        m = Member{};
        m.f0 = a->dev.f0;
        m.f1 = a->dev.f1;
        m.pos = a->dev.pos;
        m.level = a->dev_level_valid ? a->dev.level : nullptr;
        m.nPIs = a->nPIs;
        m.nObjs = a->nObjs;
        m.nPOs = a->nPOs;
        m.piOff = (int32_t)P;
        m.andOff = (int32_t)A;
        m.poOff = (int32_t)O;
        piOffs[k] = (int32_t)P;
        andOffs[k] = (int32_t)A;
        poOffs[k] = (int32_t)O;
        P += a->nPIs;
        A += a->nObjs - a->nPIs - 1;
        O += a->nPOs;
        if (1 + P + A > 0x3FFFFFF0LL) return gls_fail(ctx, GLS_ERR_ARG, "batch: union too large (%lld objects)",
                                                      (long long)(1 + P + A));
    }
    const int32_t N = (int32_t)(1 + P + A);
    PoolBlock pb{ctx};
    Member *d_mem = pb.get<Member>(count);
    int32_t *d_piOffs = pb.get<int32_t>(count), *d_andOffs = pb.get<int32_t>(count), *d_poOffs = pb.get<int32_t>(count);
    uint32_t *utag = pb.get<uint32_t>(N);
    if (!d_mem || !d_piOffs || !d_andOffs || !d_poOffs || !utag) return gls_fail(ctx, GLS_ERR_NOMEM, "batch: tables");
    GLS_TRY(stream_copy(ctx, d_mem, hm.data(), sizeof(Member) * count, hipMemcpyHostToDevice));
    GLS_TRY(stream_copy(ctx, d_piOffs, piOffs.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice));
    GLS_TRY(stream_copy(ctx, d_andOffs, andOffs.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice));
    GLS_TRY(stream_copy(ctx, d_poOffs, poOffs.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice));

    // pack
    DevGraph U;
    GLS_TRY(dev_graph_alloc(ctx, U, N, (int32_t)P, (int32_t)O));
    struct GraphGuard {
        gls_ctx *c;
        DevGraph &g;
        ~GraphGuard() { dev_graph_free(c, g); }
    } uGuard{ctx, U};
    k_pack_pis<<<nblocks(P + 1), kB, 0, st>>>((int32_t)P, count, d_piOffs, U.f0, U.f1, U.level, utag);
    if (A > 0)
        k_pack_ands<<<nblocks(A), kB, 0, st>>>(A, count, (int32_t)P, d_mem, d_andOffs, U.f0, U.f1, U.level, utag);
    if (O > 0) k_pack_pos<<<nblocks(O), kB, 0, st>>>(O, count, (int32_t)P, d_mem, d_poOffs, U.pos);
    GLS_HIP(ctx, hipGetLastError());

    // the pass
    DevGraph R;
    int32_t depth = 0;
    uint32_t *rtag = nullptr;
    uint32_t *dep = pb.get<uint32_t>(count);   // per-member depth, as the pass alone reports it
    if (!dep) return gls_fail(ctx, GLS_ERR_NOMEM, "batch: depth");
    GLS_HIP(ctx, hipMemsetAsync(dep, 0, sizeof(uint32_t) * count, st));
    gls_op_info info = {};
    std::vector<uint32_t> splitTags;
    int rc = pass == Pass::Strash    ? dev_strash(ctx, U, R, &depth, &info, utag, &rtag, dep)
             : pass == Pass::Rewrite ? dev_rewrite(ctx, U, useZeros, flags, R, &depth, &info, utag, &rtag, dep)
                                     : dev_balance(ctx, U, R, &info, utag, &rtag, &splitTags);
    if (rtag) pb.ptrs.push_back(rtag);
    if (rc == GLS_ERR_BATCH_SPLIT) {
        // members with a level wider than one balance batch go alone (their
        // batches are then the ones they get alone); the others stay together
        std::vector<char> alone(count, splitTags.empty() ? 1 : 0);
        for (uint32_t tg : splitTags)
            if (tg < (uint32_t)count) alone[tg] = 1;
        std::vector<gls_aig *> rest;
        for (int32_t k = 0; k < count; k++) {
            if (alone[k]) GLS_TRY(gls_balance(aigs[k], 1, 0));
            else rest.push_back(aigs[k]);
        }
        if (!rest.empty()) GLS_TRY(batch_pass(ctx, rest.data(), (int32_t)rest.size(), Pass::Balance, 0, 0));
        return GLS_OK;
    }
    if (rc != GLS_OK) return rc;
    struct GraphGuard rGuard{ctx, R};

    // split: counts, stable order by member, local ids
    const int64_t Ar = (int64_t)R.nObjs - P - 1;
    uint32_t *cnt = pb.get<uint32_t>(count);
    uint32_t *keys = pb.get<uint32_t>(std::max<int64_t>(Ar, 1)), *ids = pb.get<uint32_t>(std::max<int64_t>(Ar, 1));
    uint32_t *skeys = pb.get<uint32_t>(std::max<int64_t>(Ar, 1)), *sids = pb.get<uint32_t>(std::max<int64_t>(Ar, 1));
    uint32_t *localId = pb.get<uint32_t>(R.nObjs);
    if (!cnt || !keys || !ids || !skeys || !sids || !localId) return gls_fail(ctx, GLS_ERR_NOMEM, "batch: split");
    GLS_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(uint32_t) * count, st));
    if (Ar > 0) {
        k_split_count<<<tag_grid(Ar), kBlock, 0, st>>>((int32_t)P, R.nObjs, rtag, cnt, keys, ids);
        size_t tb = 0;
        const uint32_t eb = key_bits((uint32_t)count);
        GLS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, keys, skeys, ids, sids, (size_t)Ar, 0u, eb, st));
        void *tmp = pb.get<uint8_t>(tb);
        if (!tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "batch: sort storage");
        GLS_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, keys, skeys, ids, sids, (size_t)Ar, 0u, eb, st));
    }
    std::vector<uint32_t> hcnt(count), hdep(count);
    GLS_TRY(stream_copy(ctx, hcnt.data(), cnt, sizeof(uint32_t) * count, hipMemcpyDeviceToHost));
    GLS_TRY(stream_copy(ctx, hdep.data(), dep, sizeof(uint32_t) * count, hipMemcpyDeviceToHost));

    // outputs
    std::vector<DevGraph> outs(count);
    struct OutGuard {
        gls_ctx *c;
        std::vector<DevGraph> &v;
        bool keep = false;
        ~OutGuard() {
            if (!keep)
                for (auto &g : v) dev_graph_free(c, g);
        }
    } oGuard{ctx, outs};
    int64_t off = 0;
    for (int32_t k = 0; k < count; k++) {
        Member &m = hm[k];
        GLS_TRY(dev_graph_alloc(ctx, outs[k], 1 + m.nPIs + (int32_t)hcnt[k], m.nPIs, m.nPOs));
        m.of0 = outs[k].f0;
        m.of1 = outs[k].f1;
        m.opos = outs[k].pos;
        m.olevel = outs[k].level;
        m.outAnds = (int32_t)hcnt[k];
        m.outAndOff = (int32_t)off;
        off += hcnt[k];
    }
    GLS_TRY(stream_copy(ctx, d_mem, hm.data(), sizeof(Member) * count, hipMemcpyHostToDevice));
    k_split_ids<<<nblocks(std::max<int64_t>(Ar, P + 1)), kB, 0, st>>>((int32_t)P, Ar, count, d_mem, d_piOffs, skeys,
                                                                     sids, localId);
    k_split_emit<<<nblocks(std::max<int64_t>(Ar, P + 1 + count)), kB, 0, st>>>((int32_t)P, Ar, count, d_mem, d_piOffs,
                                                                              skeys, sids, localId, R.f0, R.f1,
                                                                              R.level);
    if (O > 0) k_split_pos<<<nblocks(O), kB, 0, st>>>(O, count, d_mem, d_poOffs, R.pos, localId);
    GLS_HIP(ctx, hipGetLastError());
    GLS_HIP(ctx, hipStreamSynchronize(st));

    info.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int32_t k = 0; k < count; k++) {
        gls_aig *a = aigs[k];
        aig_drop_device(a);
        a->dev = outs[k];
        a->nObjs = outs[k].nObjs;
        a->dev_valid = true;
        a->dev_level_valid = pass != Pass::Balance;
        a->host_valid = false;
        a->host_level_valid = false;
        a->levels = pass != Pass::Balance ? (int32_t)hdep[k] : -1;
        a->strashed = pass != Pass::Balance;
        a->last = info;
    }
    oGuard.keep = true;
    (void)depth;
    return GLS_OK;
}

} // namespace

// objects of the batch arena (for the size bucket of its settings)
static int64_t batch_objects(gls_aig **aigs, int32_t count) {
    int64_t n = 1;
    for (int32_t k = 0; aigs && k < count; k++)
        if (aigs[k]) n += aigs[k]->nObjs;
    return n;
}

extern "C" int gls_batch_strash(gls_ctx *ctx, gls_aig **aigs, int32_t count, int32_t flags) {
    if (!ctx) return GLS_ERR_ARG;
    if (flags != 0) return gls_fail(ctx, GLS_ERR_ARG, "gls_batch_strash: unknown flags 0x%x", flags);
    TuneScope tune(ctx, batch_objects(aigs, count));
    return batch_pass(ctx, aigs, count, Pass::Strash, 0, 0);
}

extern "C" int gls_batch_rewrite(gls_ctx *ctx, gls_aig **aigs, int32_t count, int32_t use_zeros, int32_t flags) {
    if (!ctx) return GLS_ERR_ARG;
    TuneScope tune(ctx, batch_objects(aigs, count));
    return batch_pass(ctx, aigs, count, Pass::Rewrite, use_zeros ? 1 : 0, flags);
}

extern "C" int gls_batch_balance(gls_ctx *ctx, gls_aig **aigs, int32_t count, int32_t sort_dec_id,
                                 int32_t update_level) {
    if (!ctx) return GLS_ERR_ARG;
    if (sort_dec_id != 1)
        return gls_fail(ctx, GLS_ERR_ARG, "batch balance: only sort_dec_id = 1 (level, then literal) is supported");
    TuneScope tune(ctx, batch_objects(aigs, count));
    GLS_TRY(batch_pass(ctx, aigs, count, Pass::Balance, 0, 0));
    if (update_level) GLS_TRY(batch_pass(ctx, aigs, count, Pass::Strash, 0, 0));
    return GLS_OK;
}
