// Bit-parallel simulation and simulation-based equivalence checking on the
// device.
//
// The reference's GPU simulation-based CEC is not part of its source release
// (README.md:102-111 lists it among the closed-source algorithms); this is an
// own design that checks rw / rwz / b / st results at full size without the host.
//
// Layout: one 64-bit pattern word per node and word, rows [word][node] in HBM.
// Mapping to MI355X: pattern words are independent, so each workgroup owns one
// word of one AIG and walks the whole AIG alone — no inter-workgroup
// synchronisation at all; a launch covers both AIGs and up to 256 words, so
// every CU runs its own sweep.  A sweep takes the AIG in chunks of 1024
// consecutive ids (one node per thread): a node whose fanins lie below the
// chunk reads them from its row (L2-resident: fanins are mostly recent ids),
// the others wait for their in-chunk fanins in LDS, one barrier per round.
// Ids are topological, so each round completes at least the smallest pending
// node and a chunk needs at most 1024 rounds; a chunk that does not complete
// (a fanin at or above its own id, which the handles never hold) stops the
// sweep with an error instead of waiting.
#include "gls_internal.h"
#include "dev_common.h"
#include <algorithm>
#include <mutex>

namespace {

constexpr int SIM_WG = 1024;

// splitmix64's output function (Steele, Lea, Flood, "Fast splittable
// pseudorandom number generators", 2014) over seed + (key + 1) * golden gamma
__device__ __forceinline__ uint64_t sim_mix(uint64_t seed, uint64_t key) {
    uint64_t z = seed + (key + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Pattern word w of PI j (0-based).  Exhaustive: pattern p = 64w + b assigns
// PI j the bit j of p.  Random: a hash of (j, w) under the seed.
__device__ __forceinline__ uint64_t pi_word(uint32_t j, uint32_t w, uint64_t seed, bool exhaustive) {
    if (!exhaustive) return sim_mix(seed, ((uint64_t)j << 32) | w);
    const uint64_t el[6] = {0xAAAAAAAAAAAAAAAAull, 0xCCCCCCCCCCCCCCCCull, 0xF0F0F0F0F0F0F0F0ull,
                            0xFF00FF00FF00FF00ull, 0xFFFF0000FFFF0000ull, 0xFFFFFFFF00000000ull};
    if (j < 6) return el[j];
    if (j - 6 >= 32) return 0ull;   // pattern indexes stay below 2^38
    return ((w >> (j - 6)) & 1u) ? ~0ull : 0ull;
}

struct SimGraph {
    const uint32_t *f0, *f1;
    int32_t n, nPIs;
    uint64_t *rows;   // [words][n]
};

// grid: nGraphs * words workgroups; workgroup b simulates word (b % words) of
// graph (b / words), pattern word index wordBase + (b % words).
__global__ void __launch_bounds__(SIM_WG) k_sim(SimGraph g0, SimGraph g1, int32_t words, uint32_t wordBase,
                                               uint64_t seed, int32_t exhaustive, uint32_t *err) {
    __shared__ uint64_t s_val[SIM_WG];
    __shared__ uint32_t s_done[SIM_WG / 32];
    const int gi = blockIdx.x / words, w = blockIdx.x % words;
    const SimGraph &g = gi == 0 ? g0 : g1;
    const int32_t n = g.n, nPIs = g.nPIs;
    uint64_t *row = g.rows + (size_t)w * (size_t)n;
    const int tid = threadIdx.x;
    for (int32_t v = tid; v <= nPIs && v < n; v += SIM_WG)
        row[v] = v == 0 ? 0ull : pi_word((uint32_t)(v - 1), wordBase + (uint32_t)w, seed, exhaustive != 0);
    __syncthreads();
    for (int32_t c = nPIs + 1; c < n; c += SIM_WG) {
        const int32_t v = c + tid;
        const bool active = v < n;
        uint32_t l0 = 0, l1 = 0;
        if (active) {
            l0 = g.f0[v];
            l1 = g.f1[v];
        }
        const uint32_t a = LID(l0), b = LID(l1);
        bool bad = active && ((int64_t)a >= (int64_t)c + SIM_WG || (int64_t)b >= (int64_t)c + SIM_WG ||
                              (int64_t)a >= n || (int64_t)b >= n);
        bool needA = active && !bad && (int32_t)a >= c, needB = active && !bad && (int32_t)b >= c;
        uint64_t va = 0, vb = 0;
        if (active && !bad && !needA) va = row[a];
        if (active && !bad && !needB) vb = row[b];
        if (tid < SIM_WG / 32) s_done[tid] = 0;
        __syncthreads();
        bool done = !active || bad, fresh = false;
        uint64_t val = 0;
        int round = 0;
        for (;;) {
            if (!done) {
                if (needA && ((s_done[(a - c) >> 5] >> ((a - c) & 31)) & 1u)) {
                    va = s_val[a - c];
                    needA = false;
                }
                if (needB && ((s_done[(b - c) >> 5] >> ((b - c) & 31)) & 1u)) {
                    vb = s_val[b - c];
                    needB = false;
                }
                if (!needA && !needB) {
                    val = (va ^ (LC(l0) ? ~0ull : 0ull)) & (vb ^ (LC(l1) ? ~0ull : 0ull));
                    s_val[tid] = val;
                    done = fresh = true;
                }
            }
            // values written this round become visible to the next one
            __syncthreads();
            if (fresh) atomicOr(&s_done[tid >> 5], 1u << (tid & 31));
            fresh = false;
            const int pending = __syncthreads_or(!done);
            if (!pending) break;
            if (++round > SIM_WG) {   // cannot happen for topological ids
                bad = true;
                break;
            }
        }
        if (__syncthreads_or(bad)) {
            if (tid == 0) atomicOr(err, 1u);
            return;
        }
        if (active) row[v] = val;
        __syncthreads();
    }
}

// first differing (PO, pattern) over the words of this batch, as po << 32 | pattern
__global__ void k_sim_cmp(int32_t nPOs, const uint32_t *__restrict__ posA, const uint32_t *__restrict__ posB,
                          const uint64_t *__restrict__ rowsA, int32_t nA, const uint64_t *__restrict__ rowsB,
                          int32_t nB, int32_t words, uint32_t wordBase, uint64_t validMask,
                          unsigned long long *first) {
    const int64_t total = (int64_t)nPOs * words;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t po = (int32_t)(t / words), w = (int32_t)(t % words);
        const uint32_t la = posA[po], lb = posB[po];
        const uint64_t x = rowsA[(size_t)w * nA + LID(la)] ^ (LC(la) ? ~0ull : 0ull);
        const uint64_t y = rowsB[(size_t)w * nB + LID(lb)] ^ (LC(lb) ? ~0ull : 0ull);
        const uint64_t d = (x ^ y) & validMask;
        if (d) {
            const uint64_t pat = (uint64_t)(wordBase + (uint32_t)w) * 64u + (uint64_t)__builtin_ctzll(d);
            atomicMin(first, ((unsigned long long)po << 32) | pat);
        }
    }
}

// PO words of one batch to a compact [po][batchWord] array
__global__ void k_sim_pos(int32_t nPOs, const uint32_t *__restrict__ pos, const uint64_t *__restrict__ rows, int32_t n,
                          int32_t words, uint64_t *out) {
    const int64_t total = (int64_t)nPOs * words;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t po = (int32_t)(t / words), w = (int32_t)(t % words);
        const uint32_t l = pos[po];
        out[t] = rows[(size_t)w * n + LID(l)] ^ (LC(l) ? ~0ull : 0ull);
    }
}

__global__ void k_sim_init(unsigned long long *first, uint32_t *err) {
    if (threadIdx.x == 0) {
        *first = ~0ull;
        *err = 0;
    }
}

}  // namespace

// Words simulated per launch: at most 256 (one workgroup per word and graph)
// and at most `budget` bytes of rows for the graphs together.
static int32_t sim_batch_words(int64_t objs, int32_t words, size_t budget) {
    int64_t b = (int64_t)(budget / (8 * (size_t)std::max<int64_t>(objs, 1)));
    b = std::min<int64_t>(std::min<int64_t>(b, 256), words);
    return (int32_t)std::max<int64_t>(b, 1);
}

static bool sim_exhaustive(int32_t nPIs) { return nPIs <= GLS_SIM_EXHAUSTIVE_PIS; }

static int32_t sim_words_for(int32_t nPIs, int32_t nWords) {
    if (!sim_exhaustive(nPIs)) return nWords;
    return nPIs <= 6 ? 1 : (int32_t)(1u << (nPIs - 6));
}

size_t sim_rows_bytes(int64_t objs, int32_t nPIs, int32_t nWords, size_t budget) {
    return (size_t)sim_batch_words(objs, sim_words_for(nPIs, nWords), budget) * 8 * (size_t)std::max<int64_t>(objs, 1);
}

int dev_sim_check(gls_ctx *ctx, const DevGraph &a, const DevGraph &b, int32_t nWords, uint64_t seed, uint64_t *rows,
                  size_t rowsBytes, unsigned long long *d_first, uint32_t *d_err, gls_cec_info *out) {
    hipStream_t st = ctx->stream;
    const bool exh = sim_exhaustive(a.nPIs);
    const int32_t words = sim_words_for(a.nPIs, nWords);
    const int64_t objs = (int64_t)a.nObjs + b.nObjs;
    const int32_t bw = sim_batch_words(objs, words, rowsBytes);
    if ((size_t)bw * 8 * (size_t)objs > rowsBytes)
        return gls_fail(ctx, GLS_ERR_ARG, "cec: %lld objects exceed the %zu-byte simulation buffer", (long long)objs,
                        rowsBytes);
    const uint64_t validMask = (exh && a.nPIs < 6) ? ((1ull << (1u << a.nPIs)) - 1ull) : ~0ull;
    k_sim_init<<<1, 64, 0, st>>>(d_first, d_err);
    GLS_HIP(ctx, hipGetLastError());
    for (int32_t w0 = 0; w0 < words; w0 += bw) {
        const int32_t cw = std::min(bw, words - w0);
        SimGraph ga{a.f0, a.f1, a.nObjs, a.nPIs, rows};
        SimGraph gb{b.f0, b.f1, b.nObjs, b.nPIs, rows + (size_t)cw * a.nObjs};
        k_sim<<<2 * cw, SIM_WG, 0, st>>>(ga, gb, cw, (uint32_t)w0, seed, exh ? 1 : 0, d_err);
        GLS_HIP(ctx, hipGetLastError());
        if (a.nPOs > 0) {
            k_sim_cmp<<<grid_for((int64_t)a.nPOs * cw), kBlock, 0, st>>>(a.nPOs, a.pos, b.pos, ga.rows, a.nObjs,
                                                                         gb.rows, b.nObjs, cw, (uint32_t)w0,
                                                                         validMask, d_first);
            GLS_HIP(ctx, hipGetLastError());
        }
    }
    unsigned long long first = ~0ull;
    uint32_t err = 0;
    GLS_HIP(ctx, hipMemcpyAsync(&first, d_first, sizeof first, hipMemcpyDeviceToHost, st));
    GLS_HIP(ctx, hipMemcpyAsync(&err, d_err, sizeof err, hipMemcpyDeviceToHost, st));
    GLS_HIP(ctx, hipStreamSynchronize(st));
    if (err) return gls_fail(ctx, GLS_ERR_ARG, "cec: an AND has a fanin that is not below it");
    out->exhaustive = exh ? 1 : 0;
    out->patterns = exh ? ((int64_t)1 << a.nPIs) : (int64_t)words * 64;
    out->equivalent = first == ~0ull ? 1 : 0;
    out->po = first == ~0ull ? -1 : (int32_t)(first >> 32);
    out->pattern = first == ~0ull ? -1 : (int64_t)(first & 0xFFFFFFFFull);
    out->pad = 0;
    return GLS_OK;
}

int dev_simulate(gls_ctx *ctx, const DevGraph &a, int32_t nWords, uint64_t seed, int32_t exhaustive, uint64_t *h_out) {
    hipStream_t st = ctx->stream;
    const int32_t words = nWords;
    const size_t budget = (size_t)2 << 30;
    const int32_t bw = sim_batch_words(a.nObjs, words, budget);
    uint64_t *rows = pool_get<uint64_t>(ctx, (size_t)bw * a.nObjs);
    uint64_t *d_po = pool_get<uint64_t>(ctx, (size_t)std::max(a.nPOs, 1) * bw);
    uint32_t *d_err = pool_get<uint32_t>(ctx, 1);
    struct Free {
        gls_ctx *c;
        void *p[3];
        ~Free() {
            for (void *q : p) pool_put(c, q);
        }
    } fr{ctx, {rows, d_po, d_err}};
    if (!rows || !d_po || !d_err) return gls_fail(ctx, GLS_ERR_NOMEM, "simulate: device allocation failed");
    std::vector<uint64_t> buf((size_t)std::max(a.nPOs, 1) * bw);
    for (int32_t w0 = 0; w0 < words; w0 += bw) {
        const int32_t cw = std::min(bw, words - w0);
        GLS_HIP(ctx, hipMemsetAsync(d_err, 0, sizeof(uint32_t), st));
        SimGraph ga{a.f0, a.f1, a.nObjs, a.nPIs, rows};
        k_sim<<<cw, SIM_WG, 0, st>>>(ga, ga, cw, (uint32_t)w0, seed, exhaustive, d_err);
        GLS_HIP(ctx, hipGetLastError());
        if (a.nPOs > 0) {
            k_sim_pos<<<grid_for((int64_t)a.nPOs * cw), kBlock, 0, st>>>(a.nPOs, a.pos, rows, a.nObjs, cw, d_po);
            GLS_HIP(ctx, hipGetLastError());
            GLS_HIP(ctx, hipMemcpyAsync(buf.data(), d_po, sizeof(uint64_t) * (size_t)a.nPOs * cw,
                                        hipMemcpyDeviceToHost, st));
        }
        uint32_t err = 0;
        GLS_HIP(ctx, hipMemcpyAsync(&err, d_err, sizeof err, hipMemcpyDeviceToHost, st));
        GLS_HIP(ctx, hipStreamSynchronize(st));
        if (err) return gls_fail(ctx, GLS_ERR_ARG, "simulate: an AND has a fanin that is not below it");
        for (int32_t po = 0; po < a.nPOs; po++)
            for (int32_t w = 0; w < cw; w++) h_out[(size_t)po * words + w0 + w] = buf[(size_t)po * cw + w];
    }
    return GLS_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free check (gls_cec_noalloc): a static device buffer, one copy
// per GPU, held until the process ends.  Its host-side name is not a device
// address: host code uses only the address hipGetSymbolAddress returns for
// the device that is current, cached per device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned long long scratch[GLS_CEC_NOALLOC_BYTES / 8];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
void *g_scratch[kMaxDevices] = {};
}  // namespace

// Called once per device with the device current (gls_ctx_create; again from
// the entry point if that one failed).
int sim_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return GLS_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (g_scratch[device]) return GLS_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GLS_ERR_HIP;
    g_scratch[device] = p;
    return GLS_OK;
}

int dev_sim_check_noalloc(gls_ctx *ctx, const DevGraph &a, const DevGraph &b, int32_t nWords, uint64_t seed,
                          gls_cec_info *out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return gls_fail(ctx, GLS_ERR_NODEV, "cec: no device");
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratchMu);
        base = g_scratch[ctx->device];
    }
    if (!base) {
        GLS_HIP(ctx, hipSetDevice(ctx->device));
        if (sim_scratch_init(ctx->device) != GLS_OK)
            return gls_fail(ctx, GLS_ERR_HIP, "cec: the static simulation buffer has no device address");
        std::lock_guard<std::mutex> lk(g_scratchMu);
        base = g_scratch[ctx->device];
    }
    // [0, 256): first difference and error flag; the rows after
    unsigned long long *d_first = static_cast<unsigned long long *>(base);
    uint32_t *d_err = reinterpret_cast<uint32_t *>(static_cast<char *>(base) + 64);
    uint64_t *rows = reinterpret_cast<uint64_t *>(static_cast<char *>(base) + 256);
    const size_t rowsBytes = (size_t)GLS_CEC_NOALLOC_BYTES - 256;
    const int64_t objs = (int64_t)a.nObjs + b.nObjs;
    if ((size_t)objs * 8 > rowsBytes)
        return gls_fail(ctx, GLS_ERR_ARG,
                        "cec (no allocation): %lld objects need %lld bytes per pattern word, more than the "
                        "%zu-byte static buffer holds", (long long)objs, (long long)objs * 8, rowsBytes);
    return dev_sim_check(ctx, a, b, nWords, seed, rows, rowsBytes, d_first, d_err, out);
}
