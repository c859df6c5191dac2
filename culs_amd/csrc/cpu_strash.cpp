// Host structural hashing: the `st -c` path.
//
// Follows AIGMan::strash(fCPU = true) (src/aig/strash.cu:322-466): nodes with
// no fanout (POs count as fanouts) are removed with everything that only they
// used; the rest are visited in id order, their fanins mapped to the new ids,
// ordered (lit0 <= lit1) and looked up in a hash table keyed on the literal
// pair; a node whose key is already present takes that node's id, otherwise it
// gets the next id.  Unlike the device strash there is no trivial-case folding
// (x & x, x & !x and constant fanins stay as they are).  The depth is the
// largest node delay (PIs at 0), or -1 when no AND node remains.
//
// Literals here are AIGER literals (lit 0 = FALSE).  The reference orders a
// node's fanins by its own literal values, where lit 0 = TRUE; the two orders
// differ only for a node whose fanins are both constants, which is ordered as
// the reference orders it.
#include "gls_internal.h"
#include <algorithm>
#include <vector>

namespace {

struct KeyTable {
    std::vector<uint64_t> keys;
    std::vector<uint32_t> vals;
    uint64_t mask = 0;
    static constexpr uint64_t kEmpty = ~0ull;
    explicit KeyTable(size_t n) {
        size_t cap = 1024;
        while (cap < 2 * n) cap <<= 1;
        keys.assign(cap, kEmpty);
        vals.assign(cap, 0);
        mask = cap - 1;
    }
    static uint64_t mix(uint64_t k) {   // murmur3 fmix64
        k ^= k >> 33;
        k *= 0xff51afd7ed558ccdull;
        k ^= k >> 33;
        k *= 0xc4ceb9fe1a85ec53ull;
        k ^= k >> 33;
        return k;
    }
    // returns the stored value for key, inserting `val` if absent
    uint32_t find_or_insert(uint64_t key, uint32_t val, bool &inserted) {
        for (uint64_t s = mix(key) & mask;; s = (s + 1) & mask) {
            if (keys[s] == kEmpty) {
                keys[s] = key;
                vals[s] = val;
                inserted = true;
                return val;
            }
            if (keys[s] == key) {
                inserted = false;
                return vals[s];
            }
        }
    }
};

// the reference's literal value of an AIGER literal (constant flipped)
inline uint32_t culs_order(uint32_t l) { return l < 2 ? 1u - l : l; }

} // namespace

int host_strash_cpu(int32_t nPIs, const std::vector<uint32_t> &f0, const std::vector<uint32_t> &f1,
                    const std::vector<uint32_t> &pos, std::vector<uint32_t> &of0, std::vector<uint32_t> &of1,
                    std::vector<uint32_t> &opos, std::vector<uint32_t> &olevel, int32_t &depth,
                    int32_t &nDangling) {
    const int32_t n = (int32_t)f0.size();
    std::vector<int32_t> nFanouts(n, 0);
    for (int32_t v = nPIs + 1; v < n; v++) {
        nFanouts[GLS_LID(f0[v])]++;
        nFanouts[GLS_LID(f1[v])]++;
    }
    for (uint32_t l : pos) nFanouts[GLS_LID(l)]++;

    // dangling cascade (markDeleteNode, strash.cu:360-383), with an explicit stack
    std::vector<uint8_t> dead(n, 0);
    std::vector<int32_t> stack;
    for (int32_t v = nPIs + 1; v < n; v++) {
        if (nFanouts[v] != 0 || dead[v]) continue;
        stack.push_back(v);
        while (!stack.empty()) {
            const int32_t u = stack.back();
            stack.pop_back();
            if (u <= nPIs || dead[u]) continue;
            dead[u] = 1;
            const int32_t a = (int32_t)GLS_LID(f0[u]), b = (int32_t)GLS_LID(f1[u]);
            if (--nFanouts[a] == 0) stack.push_back(a);
            if (--nFanouts[b] == 0) stack.push_back(b);
        }
    }

    // id-order hashing (strash.cu:404-447)
    std::vector<uint32_t> oldToNew(n, 0);
    for (int32_t v = 0; v <= nPIs && v < n; v++) oldToNew[v] = (uint32_t)v;
    KeyTable table((size_t)std::max<int32_t>(n - nPIs - 1, 1));
    of0.assign(n, 0);
    of1.assign(n, 0);
    olevel.assign(n, 0);
    uint32_t next = (uint32_t)nPIs + 1;
    int32_t maxDelay = -1;
    nDangling = 0;
    for (int32_t v = nPIs + 1; v < n; v++) {
        if (dead[v]) {
            nDangling++;
            continue;
        }
        uint32_t a = 2u * oldToNew[GLS_LID(f0[v])] + GLS_LC(f0[v]);
        uint32_t b = 2u * oldToNew[GLS_LID(f1[v])] + GLS_LC(f1[v]);
        if (culs_order(a) > culs_order(b)) std::swap(a, b);
        const uint64_t key = (uint64_t)a << 32 | b;
        bool inserted = false;
        const uint32_t id = table.find_or_insert(key, next, inserted);
        oldToNew[v] = id;
        if (inserted) {
            of0[id] = a;
            of1[id] = b;
            const uint32_t d = 1u + std::max(olevel[GLS_LID(a)], olevel[GLS_LID(b)]);
            olevel[id] = d;
            maxDelay = std::max(maxDelay, (int32_t)d);
            next++;
        }
    }
    of0.resize(next);
    of1.resize(next);
    olevel.resize(next);
    opos.resize(pos.size());
    for (size_t i = 0; i < pos.size(); i++) opos[i] = 2u * oldToNew[GLS_LID(pos[i])] + GLS_LC(pos[i]);
    depth = maxDelay;
    return GLS_OK;
}
