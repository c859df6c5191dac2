// Internal declarations shared by the host C++ and the HIP translation units of
// libgls.so.  Not part of the C ABI.
#pragma once
#include "../../include/gls.h"
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#define GLS_LID(l) ((l) >> 1)
#define GLS_LC(l) ((l) & 1u)
#define GLS_INVALID 0xFFFFFFFFu

// Records the first failure in the context and returns its code.
int gls_fail(gls_ctx *ctx, int code, const char *fmt, ...);

#define GLS_HIP(ctx, call)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return gls_fail((ctx), GLS_ERR_HIP, "%s failed: %s (%s:%d)", #call,            \
                            hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

#define GLS_TRY(expr)                                                                       \
    do {                                                                                    \
        int r_ = (expr);                                                                    \
        if (r_ != GLS_OK) return r_;                                                        \
    } while (0)

// Device memory with a size-bucketed free list: passes of rw/strash allocate
// and release many temporaries of similar size; caching them removes
// hipMalloc/hipFree from the steady state.
struct DevPool {
    std::multimap<size_t, void *> free_blocks;
    std::map<void *, size_t> live;
    size_t cached_bytes = 0;
    void *get(size_t bytes);
    void put(void *p);
    void release_all();
};

struct gls_ctx {
    int device = -1;
    hipStream_t stream = nullptr;
    std::string err;
    DevPool pool;
    // library tables resident on the device
    uint32_t *d_npn = nullptr;          // 65536 x (phase | perm << 5 | cls << 10)
    uint8_t *d_perms4 = nullptr;        // 24 x 4
    uint16_t *d_lib_nodes = nullptr;    // packed forest nodes
    uint16_t *d_lib_node_off = nullptr; // 223
    uint16_t *d_lib_root_off = nullptr; // 223
    uint8_t *d_lib_roots = nullptr;
    hipEvent_t ev[16] = {};
    bool lib_ready = false;
    // sampled launch timing: one event pair per window of consecutive dispatches
    std::vector<hipEvent_t> kev;      // 2 per sample, created on demand
    std::vector<int64_t> ksamp_units; // units (nodes) of each sampled window
    std::vector<int64_t> ksamp_launches; // launches in each sampled window
    int ksamp_n = 0;
    // host time of the last cut-graph capture + instantiation, and its launches
    double cut_graph_ms = 0;
    int64_t cut_graph_nodes = 0;
    // settings per size bucket (floor(log2(objects))); `active` is the set an
    // entry point chose for the pass it runs (TuneScope), else null
    gls_tuning tune[64];
    const gls_tuning *active = nullptr;
};

// the defaults of gls_tuning
constexpr gls_tuning kTuningDefault{2048, 1024, 4, 4, 4096, 3};
int tuning_bucket(int64_t objects);
// the settings of the running pass (the active set, or the bucket of `objects`)
inline const gls_tuning &tuning_for(const gls_ctx *ctx, int64_t objects) {
    return ctx->active ? *ctx->active : ctx->tune[tuning_bucket(objects)];
}
// an entry point's pass uses the settings of its input size throughout
struct TuneScope {
    gls_ctx *ctx;
    const gls_tuning *prev;
    TuneScope(gls_ctx *c, int64_t objects) : ctx(c), prev(c->active) {
        if (!prev) c->active = &c->tune[tuning_bucket(objects)];
    }
    ~TuneScope() { ctx->active = prev; }
};

// A device-resident AIG in the AIGER literal convention.
struct DevGraph {
    int32_t nObjs = 0, nPIs = 0, nPOs = 0;
    uint32_t *f0 = nullptr, *f1 = nullptr, *pos = nullptr;
    uint32_t *level = nullptr; // optional: node levels (nObjs)
};

struct gls_aig {
    gls_ctx *ctx = nullptr;
    int32_t nPIs = 0, nPOs = 0, nObjs = 0;
    int32_t levels = -1;
    bool strashed = false;
    bool host_valid = false, dev_valid = false;
    std::vector<uint32_t> h_f0, h_f1, h_pos;  // AIGER convention
    DevGraph dev;                               // device copy (pool memory)
    bool dev_level_valid = false;
    std::vector<uint32_t> h_level;              // node levels of the host copy
    bool host_level_valid = false;              // (uploaded with it when valid)
    gls_op_info last = {};
};

// pool helpers
template <class T> T *pool_get(gls_ctx *ctx, size_t n) {
    return static_cast<T *>(ctx->pool.get(n * sizeof(T) + 16));
}
void pool_put(gls_ctx *ctx, void *p);
int stream_copy(gls_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind);

// host strash of `st -c` (cpu_strash.cpp)
int host_strash_cpu(int32_t nPIs, const std::vector<uint32_t> &f0, const std::vector<uint32_t> &f1,
                    const std::vector<uint32_t> &pos, std::vector<uint32_t> &of0, std::vector<uint32_t> &of1,
                    std::vector<uint32_t> &opos, std::vector<uint32_t> &olevel, int32_t &depth,
                    int32_t &nDangling);

// host <-> device
int aig_ensure_device(gls_aig *aig);
gls_aig *aig_adopt(gls_ctx *ctx, int32_t nPIs, std::vector<uint32_t> &f0, std::vector<uint32_t> &f1,
                   std::vector<uint32_t> &pos);
int aig_upload_async(gls_aig *aig);
int aig_ensure_host(gls_aig *aig);
void aig_drop_device(gls_aig *aig);
int ctx_ensure_library(gls_ctx *ctx);

// host-side helpers
void npn_generate(uint8_t *phase, uint8_t *perm, uint8_t *cls, uint8_t *perms4);
int aiger_parse(const char *buf, int64_t len, std::vector<uint32_t> &f0, std::vector<uint32_t> &f1,
                std::vector<uint32_t> &pos, int32_t &nPIs, std::string &err);
int aiger_write(const char *path, int32_t nPIs, const std::vector<uint32_t> &f0,
                const std::vector<uint32_t> &f1, const std::vector<uint32_t> &pos);

// device passes (strash.hip / rewrite.hip).  Both consume `in` (the caller
// releases it) and fill `out` with fresh pool memory, out.level included.
// tagIn / tagOut (optional): a per-node tag carried to the surviving node
// (*tagOut is allocated from the pool, out.nObjs entries; the caller frees it).
// tagDepth (optional, zeroed by the caller, indexed by tag): the depth each
// tag's nodes alone would report, i.e. the largest wave among its live nodes.
int dev_strash(gls_ctx *ctx, const DevGraph &in, DevGraph &out, int32_t *depth, gls_op_info *info,
               const uint32_t *tagIn = nullptr, uint32_t **tagOut = nullptr, uint32_t *tagDepth = nullptr);
// What a rewrite started from, for the strash of the graph it composed
// (dev_strash_rw): its input graph g (ids 0..n-1, fanins g0/g1, levels
// glevel, structural table tOff/tEnt: bucket x holds the ANDs whose fanin of
// larger id is x, as {first fanin literal, node | second literal's complement
// << 31}).
struct RwDelta {
    int32_t n, depth;   // depth: g's largest level
    const uint32_t *g0, *g1, *glevel;
    const uint32_t *tOff;
    const uint2 *tEnt;
};
int dev_strash_rw(gls_ctx *ctx, const DevGraph &in, const RwDelta &d, DevGraph &out, int32_t *depth,
                  gls_op_info *info, const uint32_t *tagIn = nullptr, uint32_t **tagOut = nullptr,
                  uint32_t *tagDepth = nullptr);
int dev_reorder(gls_ctx *ctx, const DevGraph &in, DevGraph &out, std::vector<int32_t> &levelStart,
                const uint32_t *tagIn = nullptr, uint32_t **tagOut = nullptr);
int dev_rewrite(gls_ctx *ctx, const DevGraph &in, int useZeros, int flags, DevGraph &out, int32_t *depth,
                gls_op_info *info, const uint32_t *tagIn = nullptr, uint32_t **tagOut = nullptr,
                uint32_t *tagDepth = nullptr);
// tagIn / tagOut as in dev_strash.  Returns GLS_ERR_BATCH_SPLIT (without a
// message) when a tag has more than 32768 covers on a level: that tag would
// be rebuilt in other batches than alone.  splitTags (optional) then lists
// those tags (empty when there are more than 1024 of them).
#define GLS_ERR_BATCH_SPLIT (-100)
// sortDecId: 1 = `b` (cover inputs by level, then literal), 0 = `b -s` (by
// level, ties in thrust's shuffle order; not with tags)
int dev_balance(gls_ctx *ctx, const DevGraph &in, DevGraph &out, gls_op_info *info, const uint32_t *tagIn = nullptr,
                uint32_t **tagOut = nullptr, std::vector<uint32_t> *splitTags = nullptr, int sortDecId = 1);
int dev_enumerate_cuts(gls_ctx *ctx, const DevGraph &g, const std::vector<int32_t> &levelStart,
                       uint32_t *d_cuts, int32_t *d_nref, int mode, const uint8_t *sigBits = nullptr,
                       int *ranMode = nullptr);
int dev_evaluate(gls_ctx *ctx, const DevGraph &g, const uint32_t *d_cuts, const int32_t *d_nref,
                 int useZeros, int32_t *d_best);
void dev_graph_free(gls_ctx *ctx, DevGraph &g);
int dev_graph_alloc(gls_ctx *ctx, DevGraph &g, int32_t nObjs, int32_t nPIs, int32_t nPOs);
// simulation (sim.hip).  dev_sim_check compares the POs of a and b with rows
// (rowsBytes bytes), d_first (one u64) and d_err (one u32) supplied by the
// caller; dev_simulate copies PO words to the host ([po][nWords]).
int dev_sim_check(gls_ctx *ctx, const DevGraph &a, const DevGraph &b, int32_t nWords, uint64_t seed, uint64_t *rows,
                  size_t rowsBytes, unsigned long long *d_first, uint32_t *d_err, gls_cec_info *out);
int dev_simulate(gls_ctx *ctx, const DevGraph &a, int32_t nWords, uint64_t seed, int32_t exhaustive, uint64_t *h_out);
// gls_cec_noalloc: the same check on the static per-device buffer
// (sim_scratch_init caches its address for a device, with that device current)
int sim_scratch_init(int device);
int dev_sim_check_noalloc(gls_ctx *ctx, const DevGraph &a, const DevGraph &b, int32_t nWords, uint64_t seed,
                          gls_cec_info *out);
// Row bytes dev_sim_check needs for the first batch of words.
size_t sim_rows_bytes(int64_t objs, int32_t nPIs, int32_t nWords, size_t budget);
