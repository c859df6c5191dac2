// Host-side helpers: closed-form NPN tables, binary AIGER I/O and the seeded
// synthetic-input generators.  Nothing here touches the GPU.
#include "gls_internal.h"
#include <cmath>
#include <cstring>
#include <fstream>

// ---------------------------------------------------------------------------
// NPN tables.  The reference reads pPhases/pPerms/pMap/pPerms4 from its data
// file (src/algorithms/rewrite.cu:1281-1307); they are a pure function of the
// 16-bit truth table and are generated here: classes are numbered in order of
// their smallest member (the representative), and every member records the
// first (input phase, permutation) — phases outer, permutations inner — that
// maps the representative onto it, with bit 4 set when the complement is
// reached first.  The permutation order is the reference's 24-entry table.
// ---------------------------------------------------------------------------
static const uint8_t kPerms4[24][4] = {
    {0, 1, 2, 3}, {1, 0, 2, 3}, {0, 2, 1, 3}, {2, 0, 1, 3}, {2, 1, 0, 3}, {1, 2, 0, 3},
    {0, 1, 3, 2}, {1, 0, 3, 2}, {0, 3, 1, 2}, {3, 0, 1, 2}, {3, 1, 0, 2}, {1, 3, 0, 2},
    {0, 3, 2, 1}, {3, 0, 2, 1}, {0, 2, 3, 1}, {2, 0, 3, 1}, {2, 3, 0, 1}, {3, 2, 0, 1},
    {3, 1, 2, 0}, {1, 3, 2, 0}, {3, 2, 1, 0}, {2, 3, 1, 0}, {2, 1, 3, 0}, {1, 2, 3, 0}};

static uint32_t apply_phase_perm(uint32_t t, int phase, const uint8_t *p) {
    uint32_t r = 0;
    for (int m = 0; m < 16; m++) {
        // minterm m of the result reads minterm mo of the input: input variable
        // j lands on output variable p[j], then input phases are applied
        int mo = 0;
        for (int j = 0; j < 4; j++)
            if ((m >> p[j]) & 1) mo |= 1 << j;
        if ((t >> (mo ^ phase)) & 1) r |= 1u << m;
    }
    return r;
}

void npn_generate(uint8_t *phase, uint8_t *perm, uint8_t *cls, uint8_t *perms4) {
    std::vector<uint8_t> seen(65536, 0);
    memset(phase, 0, 65536);
    memset(perm, 0, 65536);
    memset(cls, 0, 65536);
    seen[0] = seen[0xFFFF] = 1;
    phase[0xFFFF] = 16;
    int next = 1;
    for (uint32_t u = 1; u < 0xFFFF; u++) {
        if (seen[u]) continue;
        const uint8_t c = (uint8_t)next++;
        for (int i = 0; i < 16; i++)
            for (int k = 0; k < 24; k++) {
                uint32_t v = apply_phase_perm(u, i, kPerms4[k]);
                if (!seen[v]) { seen[v] = 1; phase[v] = (uint8_t)i; perm[v] = (uint8_t)k; cls[v] = c; }
                uint32_t nv = ~v & 0xFFFF;
                if (!seen[nv]) { seen[nv] = 1; phase[nv] = (uint8_t)(i | 16); perm[nv] = (uint8_t)k; cls[nv] = c; }
            }
    }
    for (int k = 0; k < 24; k++)
        for (int j = 0; j < 4; j++) perms4[4 * k + j] = kPerms4[k][j];
}

// ---------------------------------------------------------------------------
// Binary AIGER (reference reader: src/aig_manager.cu:110-229, writer :231-267).
// ---------------------------------------------------------------------------
static bool read_uint(const char *&p, const char *end, int64_t &x) {
    while (p < end && *p == ' ') ++p;
    if (p >= end || *p < '0' || *p > '9') return false;
    x = 0;
    while (p < end && *p >= '0' && *p <= '9') { x = x * 10 + (*p - '0'); ++p; if (x > (1ll << 40)) return false; }
    return true;
}

static bool decode_delta(const unsigned char *&p, const unsigned char *end, uint64_t &x) {
    x = 0;
    int shift = 0;
    for (;;) {
        if (p >= end || shift > 35) return false;
        unsigned char ch = *p++;
        x |= (uint64_t)(ch & 0x7f) << shift;
        if (!(ch & 0x80)) return true;
        shift += 7;
    }
}

int aiger_parse(const char *buf, int64_t len, std::vector<uint32_t> &f0, std::vector<uint32_t> &f1,
                std::vector<uint32_t> &pos, int32_t &nPIs, std::string &err) {
    const char *p = buf, *end = buf + len;
    if (len < 4 || strncmp(p, "aig ", 4) != 0) { err = "not a binary AIGER file (header must start with 'aig ')"; return GLS_ERR_FORMAT; }
    p += 4;
    int64_t M, I, L, O, A;
    if (!read_uint(p, end, M) || !read_uint(p, end, I) || !read_uint(p, end, L) || !read_uint(p, end, O) ||
        !read_uint(p, end, A)) { err = "malformed AIGER header"; return GLS_ERR_FORMAT; }
    while (p < end && *p == ' ') ++p;
    if (p >= end || *p != '\n') { err = "unsupported AIGER header fields (only M I L O A)"; return GLS_ERR_FORMAT; }
    ++p;
    if (L != 0) { err = "latches are not supported"; return GLS_ERR_FORMAT; }
    if (M != I + A) { err = "M != I + A"; return GLS_ERR_FORMAT; }
    if (M + 1 > 0x7FFFFFFF) { err = "AIG too large"; return GLS_ERR_FORMAT; }
    // Every output line takes at least 2 bytes ("0\n") and every AND at least
    // 2 (two one-byte deltas), so a header that promises more than the bytes
    // left is rejected before anything is allocated.
    const int64_t left = (int64_t)(end - p);
    if (O > left / 2 || A > (left - 2 * O) / 2) { err = "AIGER header promises more outputs/ANDs than the file holds"; return GLS_ERR_FORMAT; }
    nPIs = (int32_t)I;
    const int64_t nObjs = M + 1;
    try {
        f0.assign(nObjs, 0);
        f1.assign(nObjs, 0);
        pos.assign(O, 0);
    } catch (const std::bad_alloc &) {
        err = "out of host memory for the AIG arrays";
        return GLS_ERR_NOMEM;
    }
    for (int64_t i = 0; i < O; i++) {
        int64_t x;
        if (!read_uint(p, end, x) || x > 2 * M + 1) { err = "bad output literal"; return GLS_ERR_FORMAT; }
        pos[i] = (uint32_t)x;
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
    const unsigned char *q = reinterpret_cast<const unsigned char *>(p), *qe = reinterpret_cast<const unsigned char *>(end);
    for (int64_t i = 0; i < A; i++) {
        uint64_t lhs = 2 * (uint64_t)(I + 1 + i), d0, d1;
        if (!decode_delta(q, qe, d0) || !decode_delta(q, qe, d1)) { err = "truncated AND section"; return GLS_ERR_FORMAT; }
        if (d0 == 0 || d0 > lhs || d1 > lhs - d0) { err = "invalid AND delta"; return GLS_ERR_FORMAT; }
        uint64_t r0 = lhs - d0, r1 = r0 - d1;   // r0 >= r1
        f0[I + 1 + i] = (uint32_t)r1;
        f1[I + 1 + i] = (uint32_t)r0;
    }
    return GLS_OK;
}

int aiger_write(const char *path, int32_t nPIs, const std::vector<uint32_t> &f0, const std::vector<uint32_t> &f1,
                const std::vector<uint32_t> &pos) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return GLS_ERR_IO;
    const int64_t nObjs = (int64_t)f0.size(), nAnds = nObjs - nPIs - 1;
    fprintf(fp, "aig %lld %d 0 %zu %lld\n", (long long)(nObjs - 1), nPIs, pos.size(), (long long)nAnds);
    for (uint32_t l : pos) fprintf(fp, "%u\n", l);
    std::vector<unsigned char> out;
    out.reserve((size_t)nAnds * 4 + 16);
    auto enc = [&](uint64_t x) {
        while (x & ~0x7full) { out.push_back((unsigned char)((x & 0x7f) | 0x80)); x >>= 7; }
        out.push_back((unsigned char)x);
    };
    for (int64_t v = nPIs + 1; v < nObjs; v++) {
        uint32_t a = f0[v], b = f1[v];
        if (a > b) std::swap(a, b);
        enc(2 * (uint64_t)v - b);
        enc((uint64_t)(b - a));
    }
    size_t w = fwrite(out.data(), 1, out.size(), fp);
    fprintf(fp, "c\n");
    int rc = (w == out.size() && fclose(fp) == 0) ? GLS_OK : GLS_ERR_IO;
    return rc;
}

// ---------------------------------------------------------------------------
// Generators: xoshiro256** seeded through splitmix64 (Blackman & Vigna).
// ---------------------------------------------------------------------------
namespace {
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; i++) {
            uint64_t z = (seed += 0x9E3779B97F4A7C15ULL);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    uint64_t below(uint64_t m) { return (uint64_t)(((unsigned __int128)next() * m) >> 64); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Builder {
    int32_t nPIs;
    std::vector<uint32_t> f0, f1;
    explicit Builder(int32_t pis) : nPIs(pis), f0(pis + 1, 0), f1(pis + 1, 0) {}
    uint32_t pi(int i) const { return 2u * (uint32_t)(i + 1); }
    uint32_t add(uint32_t a, uint32_t b) {
        if (a > b) std::swap(a, b);
        f0.push_back(a);
        f1.push_back(b);
        return 2u * (uint32_t)(f0.size() - 1);
    }
};
} // namespace

// Random AIG of the small synthetic configurations: each AND draws two
// distinct fanins uniformly from the PIs and the previous `window` ANDs.
int gen_random_arrays(int32_t nPIs, int32_t nAnds, int32_t nPOs, int32_t window, uint64_t seed,
                      std::vector<uint32_t> &f0, std::vector<uint32_t> &f1, std::vector<uint32_t> &pos) {
    if (nPIs < 2 || nAnds < 1 || nPOs < 0 || nPOs > nAnds || window < 1) return GLS_ERR_ARG;
    Rng rng(seed);
    Builder b(nPIs);
    b.f0.reserve((size_t)nPIs + nAnds + 1);
    b.f1.reserve((size_t)nPIs + nAnds + 1);
    for (int32_t i = 0; i < nAnds; i++) {
        const int64_t w = i < window ? i : window;
        const uint64_t m = (uint64_t)nPIs + (uint64_t)w;
        auto pick = [&](uint64_t r) -> uint32_t {
            if (r < (uint64_t)nPIs) return (uint32_t)(r + 1);
            return (uint32_t)(nPIs + 1 + (i - w) + (int64_t)(r - nPIs));
        };
        uint64_t r0 = rng.below(m), r1 = rng.below(m);
        while (r1 == r0) r1 = rng.below(m);
        uint32_t a = 2u * pick(r0) | (uint32_t)(rng.next() >> 63);
        uint32_t c = 2u * pick(r1) | (uint32_t)(rng.next() >> 63);
        b.add(a, c);
    }
    f0.swap(b.f0);
    f1.swap(b.f1);
    pos.resize(nPOs);
    const int64_t nObjs = (int64_t)f0.size();
    for (int32_t i = 0; i < nPOs; i++) pos[i] = 2u * (uint32_t)(nObjs - nPOs + i);
    return GLS_OK;
}

// Deep random AIG: fanin distance d ~ Geometric(p) over the previous ANDs; a
// draw beyond the window (or before the first AND) is replaced by a uniform
// draw over the PIs and the window.
int gen_geometric_arrays(int32_t nPIs, int32_t nAnds, int32_t nPOs, int32_t window, double p, uint64_t seed,
                         std::vector<uint32_t> &f0, std::vector<uint32_t> &f1, std::vector<uint32_t> &pos) {
    if (nPIs < 2 || nAnds < 1 || nPOs < 0 || nPOs > nAnds || window < 1 || !(p > 0 && p < 1)) return GLS_ERR_ARG;
    Rng rng(seed);
    Builder b(nPIs);
    b.f0.reserve((size_t)nPIs + nAnds + 1);
    b.f1.reserve((size_t)nPIs + nAnds + 1);
    const double lq = std::log1p(-p);
    for (int32_t i = 0; i < nAnds; i++) {
        const int64_t w = i < window ? i : window;
        auto draw = [&]() -> uint32_t {
            double u = rng.unit();
            int64_t d = 1 + (int64_t)std::floor(std::log1p(-u) / lq);
            if (d >= 1 && d <= w) return (uint32_t)(nPIs + 1 + i - d);
            uint64_t r = rng.below((uint64_t)nPIs + (uint64_t)w);
            if (r < (uint64_t)nPIs) return (uint32_t)(r + 1);
            return (uint32_t)(nPIs + 1 + (i - w) + (int64_t)(r - nPIs));
        };
        uint32_t x = draw(), y = draw();
        while (y == x) y = draw();
        uint32_t a = 2u * x | (uint32_t)(rng.next() >> 63);
        uint32_t c = 2u * y | (uint32_t)(rng.next() >> 63);
        b.add(a, c);
    }
    f0.swap(b.f0);
    f1.swap(b.f1);
    pos.resize(nPOs);
    const int64_t nObjs = (int64_t)f0.size();
    for (int32_t i = 0; i < nPOs; i++) pos[i] = 2u * (uint32_t)(nObjs - nPOs + i);
    return GLS_OK;
}

// n x n unsigned carry-save array multiplier with a final ripple-carry row.
// XOR(x, y) = AND(!AND(x, y), !AND(!x, !y)) (3 ANDs); the full-adder carry is
// OR(AND(a, b), AND(a ^ b, c)) reusing the XORs' inner ANDs.  A seeded fraction
// `redundant` of the ANDs g = AND(x, y) is re-expressed as AND(g, OR(x, y)).
int gen_multiplier_arrays(int32_t n, double redundant, uint64_t seed, std::vector<uint32_t> &f0,
                          std::vector<uint32_t> &f1, std::vector<uint32_t> &pos) {
    if (n < 2 || n > 256 || redundant < 0 || redundant > 1) return GLS_ERR_ARG;
    Rng rng(seed);
    Builder b(2 * n);
    const uint32_t NONE = 0xFFFFFFFFu;
    auto AND = [&](uint32_t x, uint32_t y) -> uint32_t {
        uint32_t g = b.add(x, y);
        if (redundant > 0 && rng.unit() < redundant) {
            uint32_t r = b.add(x ^ 1u, y ^ 1u);   // !x & !y
            g = b.add(g, r ^ 1u);                 // g & (x | y)
        }
        return g;
    };
    struct XorRes { uint32_t x, g1; };
    auto XOR = [&](uint32_t x, uint32_t y) -> XorRes {
        uint32_t g1 = AND(x, y), g2 = AND(x ^ 1u, y ^ 1u);
        return {AND(g1 ^ 1u, g2 ^ 1u), g1};
    };
    // returns (sum, carry); NONE inputs are absent bits
    auto ADD = [&](uint32_t x, uint32_t y, uint32_t z, uint32_t &carry) -> uint32_t {
        uint32_t in[3];
        int k = 0;
        if (x != NONE) in[k++] = x;
        if (y != NONE) in[k++] = y;
        if (z != NONE) in[k++] = z;
        if (k == 0) { carry = NONE; return NONE; }
        if (k == 1) { carry = NONE; return in[0]; }
        XorRes h = XOR(in[0], in[1]);
        if (k == 2) { carry = h.g1; return h.x; }
        XorRes s = XOR(h.x, in[2]);
        carry = AND(h.g1 ^ 1u, s.g1 ^ 1u) ^ 1u;
        return s.x;
    };
    auto A = [&](int i) { return b.pi(i); };
    auto B = [&](int j) { return b.pi(n + j); };
    std::vector<uint32_t> S(n), C(n, NONE), P(2 * n, NONE);
    for (int j = 0; j < n; j++) S[j] = AND(A(0), B(j));
    P[0] = S[0];
    for (int i = 1; i < n; i++) {
        std::vector<uint32_t> nS(n), nC(n);
        for (int j = 0; j < n; j++) {
            uint32_t pp = AND(A(i), B(j));
            uint32_t up = (j + 1 < n) ? S[j + 1] : NONE;
            nS[j] = ADD(pp, up, C[j], nC[j]);
        }
        S.swap(nS);
        C.swap(nC);
        P[i] = S[0];
    }
    // final ripple row: weight n + k gets S[k+1], C[k] and the running carry
    uint32_t r = NONE;
    for (int k = 0; k < n; k++) {
        uint32_t up = (k + 1 < n) ? S[k + 1] : NONE, c2;
        P[n + k] = ADD(up, C[k], r, c2);
        r = c2;
    }
    f0.swap(b.f0);
    f1.swap(b.f1);
    pos.assign(2 * n, 0);
    for (int i = 0; i < 2 * n; i++) pos[i] = P[i] == NONE ? 0u : P[i];
    (void)r;
    return GLS_OK;
}
