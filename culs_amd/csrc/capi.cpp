// C ABI of libgls.so: contexts, AIG handles and the entry points of gls.h.
#include "gls_internal.h"
#include <chrono>
#include "rwlib_forest.h"
#include <cstdarg>
#include <cstring>
#include <fstream>
#include <sstream>

int gen_random_arrays(int32_t, int32_t, int32_t, int32_t, uint64_t, std::vector<uint32_t> &, std::vector<uint32_t> &,
                      std::vector<uint32_t> &);
int gen_geometric_arrays(int32_t, int32_t, int32_t, int32_t, double, uint64_t, std::vector<uint32_t> &,
                         std::vector<uint32_t> &, std::vector<uint32_t> &);
int gen_multiplier_arrays(int32_t, double, uint64_t, std::vector<uint32_t> &, std::vector<uint32_t> &,
                          std::vector<uint32_t> &);

int gls_fail(gls_ctx *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

// ---------------------------------------------------------------------------
// device pool
// ---------------------------------------------------------------------------
static size_t round_size(size_t b) {
    size_t r = 256;
    while (r < b) r <<= 1;
    return r;
}

void *DevPool::get(size_t bytes) {
    size_t sz = round_size(bytes);
    auto it = free_blocks.find(sz);
    void *p = nullptr;
    if (it != free_blocks.end()) {
        p = it->second;
        free_blocks.erase(it);
        cached_bytes -= sz;
    } else if (hipMalloc(&p, sz) != hipSuccess) {
        // drop the cache and retry once
        release_all();
        if (hipMalloc(&p, sz) != hipSuccess) return nullptr;
    }
    live[p] = sz;
    return p;
}

void DevPool::put(void *p) {
    if (!p) return;
    auto it = live.find(p);
    if (it == live.end()) return;
    free_blocks.emplace(it->second, p);
    cached_bytes += it->second;
    live.erase(it);
}

void DevPool::release_all() {
    for (auto &kv : free_blocks) (void)hipFree(kv.second);
    free_blocks.clear();
    cached_bytes = 0;
}

void pool_put(gls_ctx *ctx, void *p) { ctx->pool.put(p); }

// Every copy is ordered on the context stream: the stream is non-blocking, so
// copies on the legacy default stream would not wait for its kernels.
int stream_copy(gls_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    GLS_HIP(ctx, hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    GLS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GLS_OK;
}

int dev_graph_alloc(gls_ctx *ctx, DevGraph &g, int32_t nObjs, int32_t nPIs, int32_t nPOs) {
    g.nObjs = nObjs;
    g.nPIs = nPIs;
    g.nPOs = nPOs;
    g.f0 = pool_get<uint32_t>(ctx, nObjs);
    g.f1 = pool_get<uint32_t>(ctx, nObjs);
    g.pos = pool_get<uint32_t>(ctx, nPOs + 1);
    g.level = pool_get<uint32_t>(ctx, nObjs);
    if (!g.f0 || !g.f1 || !g.pos || !g.level) {
        dev_graph_free(ctx, g);
        return gls_fail(ctx, GLS_ERR_NOMEM, "device allocation of an AIG with %d objects failed", nObjs);
    }
    return GLS_OK;
}

void dev_graph_free(gls_ctx *ctx, DevGraph &g) {
    pool_put(ctx, g.f0);
    pool_put(ctx, g.f1);
    pool_put(ctx, g.pos);
    pool_put(ctx, g.level);
    g.f0 = g.f1 = g.pos = g.level = nullptr;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int gls_ctx_create(int device, gls_ctx **out) {
    if (!out) return GLS_ERR_ARG;
    *out = nullptr;
    gls_ctx *ctx = new gls_ctx();
    for (auto &t : ctx->tune) t = kTuningDefault;
    if (device >= 0) {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= device) {
            delete ctx;
            return GLS_ERR_NODEV;
        }
        if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
            delete ctx;
            return GLS_ERR_HIP;
        }
        for (auto &e : ctx->ev)
            if (hipEventCreate(&e) != hipSuccess) { delete ctx; return GLS_ERR_HIP; }
        ctx->device = device;
        // the static buffer of gls_cec_noalloc: its address on this device
        // (retried by the entry point if this fails)
        (void)sim_scratch_init(device);
    }
    *out = ctx;
    return GLS_OK;
}

extern "C" void gls_ctx_destroy(gls_ctx *ctx) {
    if (!ctx) return;
    if (ctx->device >= 0) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (void *p : {(void *)ctx->d_npn, (void *)ctx->d_perms4, (void *)ctx->d_lib_nodes, (void *)ctx->d_lib_node_off,
                        (void *)ctx->d_lib_root_off, (void *)ctx->d_lib_roots})
            if (p) (void)hipFree(p);
        for (auto &kv : ctx->pool.live) (void)hipFree(kv.first);
        ctx->pool.live.clear();
        ctx->pool.release_all();
        for (auto &e : ctx->ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : ctx->kev)
            if (e) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

extern "C" const char *gls_last_error(const gls_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int gls_ctx_has_device(const gls_ctx *ctx) { return ctx && ctx->device >= 0 ? 1 : 0; }

int ctx_ensure_library(gls_ctx *ctx) {
    if (ctx->device < 0) return gls_fail(ctx, GLS_ERR_NODEV, "host-only context");
    if (ctx->lib_ready) return GLS_OK;
    std::vector<uint8_t> ph(65536), pe(65536), cl(65536), p4(96);
    npn_generate(ph.data(), pe.data(), cl.data(), p4.data());
    std::vector<uint32_t> packed(65536);
    for (int t = 0; t < 65536; t++) packed[t] = (uint32_t)ph[t] | (uint32_t)pe[t] << 5 | (uint32_t)cl[t] << 10;
    GLS_HIP(ctx, hipMalloc(&ctx->d_npn, 65536 * 4));
    GLS_HIP(ctx, hipMalloc(&ctx->d_perms4, 96));
    GLS_HIP(ctx, hipMalloc(&ctx->d_lib_nodes, sizeof(gls_forest_nodes)));
    GLS_HIP(ctx, hipMalloc(&ctx->d_lib_node_off, sizeof(gls_forest_node_off)));
    GLS_HIP(ctx, hipMalloc(&ctx->d_lib_root_off, sizeof(gls_forest_root_off)));
    GLS_HIP(ctx, hipMalloc(&ctx->d_lib_roots, sizeof(gls_forest_roots)));
    GLS_TRY(stream_copy(ctx, ctx->d_npn, packed.data(), 65536 * 4, hipMemcpyHostToDevice));
    GLS_TRY(stream_copy(ctx, ctx->d_perms4, p4.data(), 96, hipMemcpyHostToDevice));
    GLS_TRY(stream_copy(ctx, ctx->d_lib_nodes, gls_forest_nodes, sizeof(gls_forest_nodes), hipMemcpyHostToDevice));
    GLS_TRY(stream_copy(ctx, ctx->d_lib_node_off, gls_forest_node_off, sizeof(gls_forest_node_off), hipMemcpyHostToDevice));
    GLS_TRY(stream_copy(ctx, ctx->d_lib_root_off, gls_forest_root_off, sizeof(gls_forest_root_off), hipMemcpyHostToDevice));
    GLS_TRY(stream_copy(ctx, ctx->d_lib_roots, gls_forest_roots, sizeof(gls_forest_roots), hipMemcpyHostToDevice));
    ctx->lib_ready = true;
    return GLS_OK;
}

// ---------------------------------------------------------------------------
// AIG handles
// ---------------------------------------------------------------------------
static int validate(gls_ctx *ctx, int32_t nPIs, const std::vector<uint32_t> &f0, const std::vector<uint32_t> &f1,
                    const std::vector<uint32_t> &pos) {
    const int64_t n = (int64_t)f0.size();
    for (int64_t v = nPIs + 1; v < n; v++) {
        if ((int64_t)GLS_LID(f0[v]) >= v || (int64_t)GLS_LID(f1[v]) >= v)
            return gls_fail(ctx, GLS_ERR_ARG, "AND %lld has a fanin that is not smaller than itself", (long long)v);
    }
    for (size_t i = 0; i < pos.size(); i++)
        if ((int64_t)GLS_LID(pos[i]) >= n) return gls_fail(ctx, GLS_ERR_ARG, "PO %zu out of range", i);
    return GLS_OK;
}

static gls_aig *new_aig(gls_ctx *ctx, int32_t nPIs, std::vector<uint32_t> &f0, std::vector<uint32_t> &f1,
                        std::vector<uint32_t> &pos) {
    gls_aig *a = new gls_aig();
    a->ctx = ctx;
    a->nPIs = nPIs;
    a->nPOs = (int32_t)pos.size();
    a->nObjs = (int32_t)f0.size();
    a->h_f0.swap(f0);
    a->h_f1.swap(f1);
    a->h_pos.swap(pos);
    for (int32_t v = 0; v <= nPIs && v < a->nObjs; v++) a->h_f0[v] = a->h_f1[v] = 0;
    a->host_valid = true;
    return a;
}

static uint32_t conv_lit(uint32_t l, int32_t conv) { return (conv == GLS_LIT_CULS && l < 2) ? 1u - l : l; }

extern "C" int gls_aig_from_arrays(gls_ctx *ctx, int32_t nPIs, int32_t nPOs, int32_t nAnds, const uint32_t *fanin0,
                                   const uint32_t *fanin1, const uint32_t *pos, int32_t conv, gls_aig **out) {
    if (!ctx || !out || nPIs < 0 || nPOs < 0 || nAnds < 0 || (nAnds > 0 && (!fanin0 || !fanin1)) || (nPOs > 0 && !pos))
        return gls_fail(ctx, GLS_ERR_ARG, "gls_aig_from_arrays: bad arguments");
    if ((int64_t)nPIs + nAnds + 1 > 0x7FFFFFF0) return gls_fail(ctx, GLS_ERR_ARG, "AIG too large");
    const int32_t n = nPIs + nAnds + 1;
    std::vector<uint32_t> f0(n, 0), f1(n, 0), p(nPOs);
    for (int32_t v = nPIs + 1; v < n; v++) {
        uint32_t a = conv_lit(fanin0[v], conv), b = conv_lit(fanin1[v], conv);
        if (a > b) std::swap(a, b);
        f0[v] = a;
        f1[v] = b;
    }
    for (int32_t i = 0; i < nPOs; i++) p[i] = conv_lit(pos[i], conv);
    GLS_TRY(validate(ctx, nPIs, f0, f1, p));
    *out = new_aig(ctx, nPIs, f0, f1, p);
    return GLS_OK;
}

int aig_ensure_host(gls_aig *a) {
    if (a->host_valid) return GLS_OK;
    gls_ctx *ctx = a->ctx;
    a->h_f0.resize(a->nObjs);
    a->h_f1.resize(a->nObjs);
    a->h_pos.resize(a->nPOs);
    GLS_HIP(ctx, hipMemcpyAsync(a->h_f0.data(), a->dev.f0, sizeof(uint32_t) * a->nObjs, hipMemcpyDeviceToHost, ctx->stream));
    GLS_HIP(ctx, hipMemcpyAsync(a->h_f1.data(), a->dev.f1, sizeof(uint32_t) * a->nObjs, hipMemcpyDeviceToHost, ctx->stream));
    if (a->nPOs)
        GLS_HIP(ctx, hipMemcpyAsync(a->h_pos.data(), a->dev.pos, sizeof(uint32_t) * a->nPOs, hipMemcpyDeviceToHost, ctx->stream));
    GLS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    a->host_valid = true;
    return GLS_OK;
}

// A handle over arrays that are already AIGER literals with fanins below
// their node (checked by the caller); the vectors are taken over.
gls_aig *aig_adopt(gls_ctx *ctx, int32_t nPIs, std::vector<uint32_t> &f0, std::vector<uint32_t> &f1,
                   std::vector<uint32_t> &pos) {
    return new_aig(ctx, nPIs, f0, f1, pos);
}

// Allocates the device copy and queues its upload on the context stream; the
// caller synchronises the stream before the host arrays change.
int aig_upload_async(gls_aig *a) {
    gls_ctx *ctx = a->ctx;
    if (ctx->device < 0) return gls_fail(ctx, GLS_ERR_NODEV, "host-only context");
    if (a->dev_valid) return GLS_OK;
    GLS_TRY(dev_graph_alloc(ctx, a->dev, a->nObjs, a->nPIs, a->nPOs));
    GLS_HIP(ctx, hipMemcpyAsync(a->dev.f0, a->h_f0.data(), sizeof(uint32_t) * a->nObjs, hipMemcpyHostToDevice, ctx->stream));
    GLS_HIP(ctx, hipMemcpyAsync(a->dev.f1, a->h_f1.data(), sizeof(uint32_t) * a->nObjs, hipMemcpyHostToDevice, ctx->stream));
    if (a->nPOs)
        GLS_HIP(ctx, hipMemcpyAsync(a->dev.pos, a->h_pos.data(), sizeof(uint32_t) * a->nPOs, hipMemcpyHostToDevice, ctx->stream));
    a->dev_valid = true;
    a->dev_level_valid = false;
    return GLS_OK;
}

int aig_ensure_device(gls_aig *a) {
    gls_ctx *ctx = a->ctx;
    if (ctx->device < 0) return gls_fail(ctx, GLS_ERR_NODEV, "host-only context");
    if (a->dev_valid) return GLS_OK;
    GLS_TRY(dev_graph_alloc(ctx, a->dev, a->nObjs, a->nPIs, a->nPOs));
    GLS_HIP(ctx, hipMemcpyAsync(a->dev.f0, a->h_f0.data(), sizeof(uint32_t) * a->nObjs, hipMemcpyHostToDevice, ctx->stream));
    GLS_HIP(ctx, hipMemcpyAsync(a->dev.f1, a->h_f1.data(), sizeof(uint32_t) * a->nObjs, hipMemcpyHostToDevice, ctx->stream));
    if (a->nPOs)
        GLS_HIP(ctx, hipMemcpyAsync(a->dev.pos, a->h_pos.data(), sizeof(uint32_t) * a->nPOs, hipMemcpyHostToDevice, ctx->stream));
    if (a->host_level_valid)
        GLS_HIP(ctx, hipMemcpyAsync(a->dev.level, a->h_level.data(), sizeof(uint32_t) * a->nObjs,
                                    hipMemcpyHostToDevice, ctx->stream));
    GLS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    a->dev_valid = true;
    a->dev_level_valid = a->host_level_valid;
    return GLS_OK;
}

void aig_drop_device(gls_aig *a) {
    if (a->dev_valid) dev_graph_free(a->ctx, a->dev);
    a->dev_valid = false;
    a->dev_level_valid = false;
}

extern "C" int gls_aig_to_arrays(gls_aig *a, uint32_t *fanin0, uint32_t *fanin1, uint32_t *pos, int32_t conv) {
    if (!a) return GLS_ERR_ARG;
    GLS_TRY(aig_ensure_host(a));
    for (int32_t v = 0; v < a->nObjs; v++) {
        if (fanin0) fanin0[v] = v <= a->nPIs ? 0 : conv_lit(a->h_f0[v], conv);
        if (fanin1) fanin1[v] = v <= a->nPIs ? 0 : conv_lit(a->h_f1[v], conv);
    }
    if (pos)
        for (int32_t i = 0; i < a->nPOs; i++) pos[i] = conv_lit(a->h_pos[i], conv);
    return GLS_OK;
}

extern "C" int gls_aig_stats(gls_aig *a, gls_stats *s) {
    if (!a || !s) return GLS_ERR_ARG;
    s->nPIs = a->nPIs;
    s->nPOs = a->nPOs;
    s->nObjs = a->nObjs;
    s->nAnds = a->nObjs - a->nPIs - 1;
    s->levels = a->levels;
    s->on_device = a->dev_valid && !a->host_valid ? 1 : 0;
    return GLS_OK;
}

extern "C" int gls_aig_copy(gls_aig *src, gls_aig **out) {
    if (!src || !out) return GLS_ERR_ARG;
    if (src->dev_valid) {
        // device-resident copy: stays in HBM, no host round trip
        gls_ctx *ctx = src->ctx;
        gls_aig *a = new gls_aig();
        a->ctx = ctx;
        a->nPIs = src->nPIs;
        a->nPOs = src->nPOs;
        a->nObjs = src->nObjs;
        a->levels = src->levels;
        a->strashed = src->strashed;
        int rc = dev_graph_alloc(ctx, a->dev, src->nObjs, src->nPIs, src->nPOs);
        if (rc != GLS_OK) { delete a; return rc; }
        const size_t nb = sizeof(uint32_t) * (size_t)src->nObjs;
        hipStream_t st = ctx->stream;
        if (hipMemcpyAsync(a->dev.f0, src->dev.f0, nb, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(a->dev.f1, src->dev.f1, nb, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(a->dev.level, src->dev.level, nb, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            (src->nPOs && hipMemcpyAsync(a->dev.pos, src->dev.pos, sizeof(uint32_t) * src->nPOs,
                                         hipMemcpyDeviceToDevice, st) != hipSuccess)) {
            dev_graph_free(ctx, a->dev);
            delete a;
            return gls_fail(ctx, GLS_ERR_HIP, "device copy failed");
        }
        a->dev_valid = true;
        a->dev_level_valid = src->dev_level_valid;
        a->host_valid = false;
        *out = a;
        return GLS_OK;
    }
    GLS_TRY(aig_ensure_host(src));
    std::vector<uint32_t> f0 = src->h_f0, f1 = src->h_f1, p = src->h_pos;
    gls_aig *a = new_aig(src->ctx, src->nPIs, f0, f1, p);
    a->levels = src->levels;
    a->strashed = src->strashed;
    a->h_level = src->h_level;
    a->host_level_valid = src->host_level_valid;
    *out = a;
    return GLS_OK;
}

extern "C" void gls_aig_free(gls_aig *a) {
    if (!a) return;
    aig_drop_device(a);
    delete a;
}

static int read_aiger_mem(gls_ctx *ctx, const char *buf, int64_t len, gls_aig **out) {
    std::vector<uint32_t> f0, f1, pos;
    int32_t nPIs = 0;
    std::string err;
    int rc = aiger_parse(buf, len, f0, f1, pos, nPIs, err);
    if (rc != GLS_OK) return gls_fail(ctx, rc, "%s", err.c_str());
    GLS_TRY(validate(ctx, nPIs, f0, f1, pos));
    *out = new_aig(ctx, nPIs, f0, f1, pos);
    return GLS_OK;
}

// Host allocation failures (a huge but well-formed file) come back as
// GLS_ERR_NOMEM; no exception crosses the C ABI.
extern "C" int gls_aig_read_aiger_mem(gls_ctx *ctx, const char *buf, int64_t len, gls_aig **out) {
    if (!ctx || !buf || !out) return GLS_ERR_ARG;
    try {
        return read_aiger_mem(ctx, buf, len, out);
    } catch (const std::bad_alloc &) {
        return gls_fail(ctx, GLS_ERR_NOMEM, "out of host memory reading an AIGER buffer of %lld bytes", (long long)len);
    }
}

extern "C" int gls_aig_read_aiger(gls_ctx *ctx, const char *path, gls_aig **out) {
    if (!ctx || !path || !out) return GLS_ERR_ARG;
    try {
        std::ifstream in(path, std::ios::binary);
        if (!in) return gls_fail(ctx, GLS_ERR_IO, "cannot open %s", path);
        std::stringstream ss;
        ss << in.rdbuf();
        std::string s = ss.str();
        return read_aiger_mem(ctx, s.data(), (int64_t)s.size(), out);
    } catch (const std::bad_alloc &) {
        return gls_fail(ctx, GLS_ERR_NOMEM, "out of host memory reading %s", path);
    }
}

extern "C" int gls_aig_write_aiger(gls_aig *a, const char *path) {
    if (!a || !path) return GLS_ERR_ARG;
    GLS_TRY(aig_ensure_host(a));
    int rc = aiger_write(path, a->nPIs, a->h_f0, a->h_f1, a->h_pos);
    if (rc != GLS_OK) return gls_fail(a->ctx, rc, "cannot write %s", path);
    return GLS_OK;
}

// ---------------------------------------------------------------------------
// hot path
// ---------------------------------------------------------------------------
static void adopt(gls_aig *a, DevGraph &out, int32_t depth) {
    aig_drop_device(a);
    a->dev = out;
    a->nObjs = out.nObjs;
    a->dev_valid = true;
    a->dev_level_valid = true;
    a->host_valid = false;
    a->host_level_valid = false;
    a->levels = depth;
}

// `st -c`: host strash (AIGMan::strash(fCPU = true), strash.cu:322-466)
static int strash_cpu(gls_aig *a) {
    GLS_TRY(aig_ensure_host(a));
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> f0, f1, pos, lev;
    int32_t depth = 0, nDangling = 0;
    GLS_TRY(host_strash_cpu(a->nPIs, a->h_f0, a->h_f1, a->h_pos, f0, f1, pos, lev, depth, nDangling));
    aig_drop_device(a);
    a->nObjs = (int32_t)f0.size();
    a->h_f0.swap(f0);
    a->h_f1.swap(f1);
    a->h_pos.swap(pos);
    a->h_level.swap(lev);
    a->host_valid = true;
    a->host_level_valid = true;
    a->levels = depth;
    a->strashed = true;
    gls_op_info info = {};
    info.total_ms = info.strash_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    a->last = info;
    return GLS_OK;
}

extern "C" int gls_strash(gls_aig *a, int32_t flags) {
    if (!a) return GLS_ERR_ARG;
    if (flags & ~(GLS_ST_CPU | GLS_ST_WAVES | GLS_ST_LEVELS) || (flags & GLS_ST_WAVES && flags & GLS_ST_LEVELS))
        return gls_fail(a->ctx, GLS_ERR_ARG, "gls_strash: unknown flags 0x%x", flags);
    if (flags & GLS_ST_CPU) return strash_cpu(a);
    gls_ctx *ctx = a->ctx;
    TuneScope tune(ctx, a->nObjs);
    GLS_TRY(aig_ensure_device(a));
    GLS_HIP(ctx, hipEventRecord(ctx->ev[14], ctx->stream));
    DevGraph out;
    int32_t depth = 0;
    gls_op_info info = {};
    // large graphs: the whole-graph mode of the incremental strash (levels by
    // one sweep over id order, every node hashed in rounds), which gives way
    // to the wave sweep for wide graphs and ids that are not topological
    constexpr int32_t kLevelStrashMin = 1 << 18;
    const bool levels = (flags & GLS_ST_LEVELS) || (!(flags & GLS_ST_WAVES) && a->nObjs >= kLevelStrashMin);
    if (levels) {
        const RwDelta whole{0, 0, nullptr, nullptr, nullptr, nullptr, nullptr};
        GLS_TRY(dev_strash_rw(ctx, a->dev, whole, out, &depth, &info));
    } else {
        GLS_TRY(dev_strash(ctx, a->dev, out, &depth, &info));
    }
    GLS_HIP(ctx, hipEventRecord(ctx->ev[15], ctx->stream));
    GLS_HIP(ctx, hipEventSynchronize(ctx->ev[15]));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev[14], ctx->ev[15]);
    info.total_ms = info.strash_ms = ms;
    adopt(a, out, depth);
    a->strashed = true;
    a->last = info;
    return GLS_OK;
}

extern "C" int gls_rewrite(gls_aig *a, int32_t use_zeros, int32_t flags) {
    if (!a) return GLS_ERR_ARG;
    gls_ctx *ctx = a->ctx;
    TuneScope tune(ctx, a->nObjs);
    GLS_TRY(aig_ensure_device(a));
    if (!a->strashed || !a->dev_level_valid) GLS_TRY(gls_strash(a, 0));
    GLS_HIP(ctx, hipEventRecord(ctx->ev[14], ctx->stream));
    DevGraph out;
    int32_t depth = 0;
    gls_op_info info = {};
    GLS_TRY(dev_rewrite(ctx, a->dev, use_zeros ? 1 : 0, flags, out, &depth, &info));
    adopt(a, out, depth);
    a->strashed = true;
    a->last = info;
    return GLS_OK;
}

extern "C" int gls_balance(gls_aig *a, int32_t sort_dec_id, int32_t update_level) {
    if (!a) return GLS_ERR_ARG;
    gls_ctx *ctx = a->ctx;
    if (sort_dec_id != 0 && sort_dec_id != 1)
        return gls_fail(ctx, GLS_ERR_ARG, "balance: sort_dec_id must be 1 (`b`) or 0 (`b -s`)");
    TuneScope tune(ctx, a->nObjs);
    GLS_TRY(aig_ensure_device(a));
    DevGraph out;
    gls_op_info info = {};
    GLS_TRY(dev_balance(ctx, a->dev, out, &info, nullptr, nullptr, nullptr, sort_dec_id));
    adopt(a, out, -1);
    a->dev_level_valid = false;   // levels are not computed by the balance itself
    a->strashed = false;
    a->last = info;
    if (update_level) {
        const double ms = info.total_ms;
        GLS_TRY(gls_strash(a, 0));
        a->last.total_ms += ms;
    }
    return GLS_OK;
}

extern "C" int gls_aig_last_op(gls_aig *a, gls_op_info *out) {
    if (!a || !out) return GLS_ERR_ARG;
    *out = a->last;
    return GLS_OK;
}

extern "C" int gls_reorder(gls_aig *a, uint32_t *level_out) {
    if (!a) return GLS_ERR_ARG;
    gls_ctx *ctx = a->ctx;
    TuneScope tune(ctx, a->nObjs);
    GLS_TRY(aig_ensure_device(a));
    if (!a->strashed || !a->dev_level_valid) GLS_TRY(gls_strash(a, 0));
    DevGraph out;
    std::vector<int32_t> starts;
    GLS_TRY(dev_reorder(ctx, a->dev, out, starts));
    int32_t depth = a->levels;
    adopt(a, out, depth);
    if (level_out) {
        GLS_TRY(stream_copy(ctx, level_out, a->dev.level, sizeof(uint32_t) * a->nObjs, hipMemcpyDeviceToHost));
    }
    return GLS_OK;
}

// level starts of an AIG whose ids are level-major with a valid level array
static int level_starts_host(gls_aig *a, std::vector<int32_t> &starts) {
    gls_ctx *ctx = a->ctx;
    std::vector<uint32_t> lev(a->nObjs);
    GLS_TRY(stream_copy(ctx, lev.data(), a->dev.level, sizeof(uint32_t) * a->nObjs, hipMemcpyDeviceToHost));
    uint32_t maxL = 0;
    for (int32_t v = 1; v < a->nObjs; v++) {
        if (lev[v] < lev[v - 1] && v > 1) return gls_fail(ctx, GLS_ERR_ARG, "AIG is not level-major (call gls_reorder)");
        maxL = lev[v] > maxL ? lev[v] : maxL;
    }
    starts.assign(maxL + 2, a->nObjs);
    for (int32_t v = a->nObjs - 1; v >= 1; v--) starts[lev[v]] = v;
    return GLS_OK;
}

extern "C" int gls_enumerate_cuts(gls_aig *a, uint32_t *cuts) { return gls_enumerate_cuts_sched(a, 0, cuts); }

extern "C" int gls_enumerate_cuts_sched(gls_aig *a, int32_t flags, uint32_t *cuts) {
    if (!a || !cuts) return GLS_ERR_ARG;
    if (flags & ~(GLS_RW_COOP_CUTS | GLS_RW_DATAFLOW_CUTS | GLS_RW_XCD_CUTS | GLS_RW_LEVEL_CUTS | GLS_RW_PAIR_CUTS |
                  GLS_RW_GRAPH_CUTS | GLS_RW_PLAN_CUTS | GLS_RW_QUAD_CUTS))
        return gls_fail(a->ctx, GLS_ERR_ARG, "gls_enumerate_cuts_sched: unknown flags 0x%x", flags);
    gls_ctx *ctx = a->ctx;
    GLS_TRY(aig_ensure_device(a));
    if (!a->dev_level_valid) return gls_fail(ctx, GLS_ERR_ARG, "levels unknown: call gls_reorder first");
    std::vector<int32_t> starts;
    GLS_TRY(level_starts_host(a, starts));
    const size_t n = (size_t)a->nObjs;
    uint32_t *d_cuts = pool_get<uint32_t>(ctx, n * 40);
    int32_t *d_nref = pool_get<int32_t>(ctx, n);
    if (!d_cuts || !d_nref) { pool_put(ctx, d_cuts); pool_put(ctx, d_nref); return gls_fail(ctx, GLS_ERR_NOMEM, "cuts"); }
    int rc = GLS_OK;
    {
        TuneScope ts(ctx, a->nObjs);
        rc = dev_enumerate_cuts(ctx, a->dev, starts, d_cuts, d_nref, flags);
    }
    if (rc == GLS_OK) rc = stream_copy(ctx, cuts, d_cuts, n * 40 * sizeof(uint32_t), hipMemcpyDeviceToHost);
    pool_put(ctx, d_cuts);
    pool_put(ctx, d_nref);
    return rc;
}

extern "C" int gls_evaluate(gls_aig *a, int32_t use_zeros, int32_t *best) {
    if (!a || !best) return GLS_ERR_ARG;
    gls_ctx *ctx = a->ctx;
    GLS_TRY(aig_ensure_device(a));
    if (!a->dev_level_valid) return gls_fail(ctx, GLS_ERR_ARG, "levels unknown: call gls_reorder first");
    std::vector<int32_t> starts;
    GLS_TRY(level_starts_host(a, starts));
    const size_t n = (size_t)a->nObjs;
    uint32_t *d_cuts = pool_get<uint32_t>(ctx, n * 40);
    int32_t *d_nref = pool_get<int32_t>(ctx, n);
    int32_t *d_best = pool_get<int32_t>(ctx, n * 4);
    int rc = GLS_OK;
    if (!d_cuts || !d_nref || !d_best) rc = gls_fail(ctx, GLS_ERR_NOMEM, "evaluate");
    if (rc == GLS_OK) rc = dev_enumerate_cuts(ctx, a->dev, starts, d_cuts, d_nref, 0);
    if (rc == GLS_OK) rc = dev_evaluate(ctx, a->dev, d_cuts, d_nref, use_zeros, d_best);
    if (rc == GLS_OK) rc = stream_copy(ctx, best, d_best, n * 4 * sizeof(int32_t), hipMemcpyDeviceToHost);
    pool_put(ctx, d_cuts);
    pool_put(ctx, d_nref);
    pool_put(ctx, d_best);
    return rc;
}

extern "C" int gls_npn_tables(uint8_t *phase, uint8_t *perm, uint8_t *cls, uint8_t *perms4) {
    if (!phase || !perm || !cls || !perms4) return GLS_ERR_ARG;
    npn_generate(phase, perm, cls, perms4);
    return GLS_OK;
}

// ---------------------------------------------------------------------------
// generators
// ---------------------------------------------------------------------------
static int finish_gen(gls_ctx *ctx, int rc, int32_t nPIs, std::vector<uint32_t> &f0, std::vector<uint32_t> &f1,
                      std::vector<uint32_t> &pos, gls_aig **out) {
    if (rc != GLS_OK) return gls_fail(ctx, rc, "generator: bad arguments");
    *out = new_aig(ctx, nPIs, f0, f1, pos);
    return GLS_OK;
}

extern "C" int gls_gen_random(gls_ctx *ctx, int32_t nPIs, int32_t nAnds, int32_t nPOs, int32_t window, uint64_t seed,
                              gls_aig **out) {
    if (!ctx || !out) return GLS_ERR_ARG;
    std::vector<uint32_t> f0, f1, pos;
    int rc = gen_random_arrays(nPIs, nAnds, nPOs, window, seed, f0, f1, pos);
    return finish_gen(ctx, rc, nPIs, f0, f1, pos, out);
}

extern "C" int gls_gen_geometric(gls_ctx *ctx, int32_t nPIs, int32_t nAnds, int32_t nPOs, int32_t window, double p,
                                 uint64_t seed, gls_aig **out) {
    if (!ctx || !out) return GLS_ERR_ARG;
    std::vector<uint32_t> f0, f1, pos;
    int rc = gen_geometric_arrays(nPIs, nAnds, nPOs, window, p, seed, f0, f1, pos);
    return finish_gen(ctx, rc, nPIs, f0, f1, pos, out);
}

extern "C" int gls_gen_multiplier(gls_ctx *ctx, int32_t nbits, double redundant, uint64_t seed, gls_aig **out) {
    if (!ctx || !out) return GLS_ERR_ARG;
    std::vector<uint32_t> f0, f1, pos;
    int rc = gen_multiplier_arrays(nbits, redundant, seed, f0, f1, pos);
    return finish_gen(ctx, rc, 2 * nbits, f0, f1, pos, out);
}

// ---------------------------------------------------------------------------
// equivalence checking by simulation
// ---------------------------------------------------------------------------
// rows of simulation values held at once by gls_cec (both AIGs together)
static constexpr size_t kCecRowBudget = (size_t)8 << 30;

extern "C" int gls_cec(gls_aig *a, gls_aig *b, int32_t nWords, uint64_t seed, gls_cec_info *out) {
    if (!a || !b || !out) return GLS_ERR_ARG;
    gls_ctx *ctx = a->ctx;
    if (b->ctx != ctx) return gls_fail(ctx, GLS_ERR_ARG, "cec: the AIGs belong to different contexts");
    if (a->nPIs != b->nPIs || a->nPOs != b->nPOs)
        return gls_fail(ctx, GLS_ERR_ARG, "cec: interfaces differ (%d vs %d PIs, %d vs %d POs)", a->nPIs, b->nPIs,
                        a->nPOs, b->nPOs);
    if (a->nPIs > GLS_SIM_EXHAUSTIVE_PIS && nWords < 1)
        return gls_fail(ctx, GLS_ERR_ARG, "cec: nWords must be positive");
    GLS_TRY(aig_ensure_device(a));
    GLS_TRY(aig_ensure_device(b));
    const size_t bytes = sim_rows_bytes((int64_t)a->nObjs + b->nObjs, a->nPIs, nWords, kCecRowBudget);
    uint64_t *rows = static_cast<uint64_t *>(ctx->pool.get(bytes));
    unsigned long long *d_first = pool_get<unsigned long long>(ctx, 1);
    uint32_t *d_err = pool_get<uint32_t>(ctx, 1);
    int rc = GLS_OK;
    if (!rows || !d_first || !d_err) rc = gls_fail(ctx, GLS_ERR_NOMEM, "cec: device allocation of %zu bytes", bytes);
    if (rc == GLS_OK) rc = dev_sim_check(ctx, a->dev, b->dev, nWords, seed, rows, bytes, d_first, d_err, out);
    pool_put(ctx, rows);
    pool_put(ctx, d_first);
    pool_put(ctx, d_err);
    return rc;
}

extern "C" int gls_cec_noalloc(gls_aig *a, gls_aig *b, int32_t nWords, uint64_t seed, gls_cec_info *out) {
    if (!a || !b || !out) return GLS_ERR_ARG;
    gls_ctx *ctx = a->ctx;
    if (b->ctx != ctx) return gls_fail(ctx, GLS_ERR_ARG, "cec: the AIGs belong to different contexts");
    if (ctx->device < 0) return gls_fail(ctx, GLS_ERR_NODEV, "host-only context");
    if (a->nPIs != b->nPIs || a->nPOs != b->nPOs)
        return gls_fail(ctx, GLS_ERR_ARG, "cec: interfaces differ (%d vs %d PIs, %d vs %d POs)", a->nPIs, b->nPIs,
                        a->nPOs, b->nPOs);
    if (a->nPIs > GLS_SIM_EXHAUSTIVE_PIS && nWords < 1)
        return gls_fail(ctx, GLS_ERR_ARG, "cec: nWords must be positive");
    // putting a handle on the device would allocate: it must be there already
    if (!a->dev_valid || !b->dev_valid)
        return gls_fail(ctx, GLS_ERR_ARG, "cec (no allocation): both AIGs must already be on the device");
    return dev_sim_check_noalloc(ctx, a->dev, b->dev, nWords, seed, out);
}

extern "C" int gls_simulate(gls_aig *a, int32_t nWords, uint64_t seed, int32_t exhaustive, uint64_t *po_words) {
    if (!a || (a->nPOs > 0 && !po_words) || nWords < 1) return a ? gls_fail(a->ctx, GLS_ERR_ARG, "simulate: bad arguments") : GLS_ERR_ARG;
    GLS_TRY(aig_ensure_device(a));
    return dev_simulate(a->ctx, a->dev, nWords, seed, exhaustive ? 1 : 0, po_words);
}
