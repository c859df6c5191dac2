// DAG-aware rewriting with 4-input cuts on the device.
//
// Stages (reference in parentheses, paths relative to the CULS tree):
//   reorder   level-major renumbering, ties by id (rewrite.cu:1245-1279)
//   cuts      per-level 4-feasible cut enumeration, 16-bit truth tables, support
//             minimisation, CULS slot policy (rewrite.cu:49-294)
//   table     strash table of the current AND nodes (replaces the chained
//             table of rewrite.cu:328-367, 450-471)
//   eval      MFFC size within the cut, NPN class, library matching, nodes
//             added per subgraph root, best (gain, root level) (rewrite.cu:296-447)
//   plan      the chosen replacement of every candidate: deleted set W (root +
//             MFFC), leaves L, reused existing nodes M, nodes to create
//   select    greedy conflict-free selection by priority (gain desc, id asc),
//             computed in rounds of local maxima with atomicMax claims; it
//             replaces the reference's sequential host choice pass
//             (rewrite.cu:811-897), which cannot run in parallel
//   build     new nodes for the selected plans, ids by prefix sum, fanin
//             substitution of replaced roots, then the device strash.
//
// Everything is exact integer work and independent of thread timing: every
// shared decision goes through a prefix sum, a min/max atomic or a stable sort.
#include "gls_internal.h"
#include "dev_common.h"
#include "rwlib_forest.h"
#include "tt_place.h"
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>
#include <cstdlib>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

namespace {

constexpr int NCUT = 8;
constexpr int MFFC_TAB = 8;
constexpr int NMATCH = 54;
constexpr int EVAL_BLOCK = 64;
constexpr int PLAN_MAXM = 16;

struct Lib {
    const uint32_t *npn;     // phase | perm << 5 | cls << 10
    const uint8_t *perms4;
    const uint16_t *nodes, *nodeOff, *rootOff;
    const uint8_t *roots;
};

// complement of library input a: only the 4 cut inputs carry one (the
// reference keeps these bits in a 64-bit word, rewrite.cu:388-395)
__device__ __forceinline__ uint32_t isc_bit(uint32_t isC, int a) { return a < 4 ? (isC >> a) & 1u : 0u; }
__device__ __forceinline__ int lib_nn(const Lib &L, int c) { return L.nodeOff[c + 1] - L.nodeOff[c]; }
__device__ __forceinline__ int lib_nr(const Lib &L, int c) { return L.rootOff[c + 1] - L.rootOff[c]; }
__device__ __forceinline__ int lib_root(const Lib &L, int c, int k) { return L.roots[L.rootOff[c] + k]; }
__device__ __forceinline__ uint32_t lib_word(const Lib &L, int c, int j) { return L.nodes[L.nodeOff[c] + j]; }

// ------------------------------------------------------------------------
// cut enumeration
// ------------------------------------------------------------------------
struct DCut {
    uint32_t leaves[4];
    uint32_t tt, value;
    int32_t nLeaves, used;
};

// Leaf signature of the reference (rewrite.cu:68, 262: bit leaf & 31).  In a
// batched pass sigBits holds each node's bit as it would be in its own AIG
// (its id there & 31), so that the prefilter of rewrite.cu:283 sees the same
// signatures as when that AIG is rewritten alone.
// Without branches per leaf: the sigBits test is uniform, and in the batched
// case the four byte loads are issued at once (absent leaves read entry 0).
__device__ __forceinline__ uint32_t cut_sign_of(const DCut &c, const uint8_t *__restrict__ sigBits) {
    uint32_t b[4];
    if (sigBits) {
#pragma unroll
        for (int i = 0; i < 4; i++) b[i] = sigBits[i < c.nLeaves ? c.leaves[i] : 0u];
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) b[i] = c.leaves[i] & 31;
    }
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) s |= i < c.nLeaves ? 1u << b[i] : 0u;
    return s;
}

// the reference counts of a cut's leaves, all four loads in flight at once;
// absent leaves read node 0's entry
__device__ __forceinline__ void leaf_refs(const DCut &c, const int32_t *nRef, int (&r)[4]) {
#pragma unroll
    for (int i = 0; i < 4; i++) r[i] = nRef[i < c.nLeaves ? c.leaves[i] : 0u];
}

// CutFindValue (rewrite.cu:49-61) of a cut of nLeaves leaves whose reference
// counts are r[0..nLeaves-1]
__device__ __forceinline__ uint32_t cut_value_of(int nLeaves, const int (&r)[4]) {
    int value = 0, nOnes = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (i < nLeaves) {
            value += r[i];
            nOnes += (r[i] == 1);
        }
    if (nLeaves < 2) return 1001;
    if (value > 1000) value = 1000;
    if (nOnes > 3) value = 5 - nOnes;
    return (uint32_t)value;
}

__device__ __forceinline__ uint32_t cut_value(const DCut &c, const int32_t *nRef) {
    int r[4];
    leaf_refs(c, nRef, r);
    int value = 0, nOnes = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (i < c.nLeaves) {
            value += r[i];
            nOnes += (r[i] == 1);
        }
    if (c.nLeaves < 2) return 1001;
    if (value > 1000) value = 1000;
    if (nOnes > 3) value = 5 - nOnes;
    return (uint32_t)value;
}

__device__ __forceinline__ void unpack_cut(const uint32_t *p, DCut &c) {
    c.leaves[0] = p[0]; c.leaves[1] = p[1]; c.leaves[2] = p[2]; c.leaves[3] = p[3];
    uint32_t m = p[4];
    c.tt = m & 0xFFFF;
    c.value = (m >> 16) & 0x7FF;
    c.nLeaves = (m >> 27) & 7;
    c.used = (m >> 30) & 1;
}

// Sorted union of two sorted leaf sets (rewrite.cu:112-163), the reference's
// two-pointer walk unrolled over the four output positions; fails above 4
// leaves.  The current element of each set is selected with constant indexes,
// so the leaf arrays stay in registers.
__device__ __forceinline__ uint32_t sel4(const uint32_t *l, int i) {
    return i == 0 ? l[0] : i == 1 ? l[1] : i == 2 ? l[2] : l[3];
}

// phA / phB: the output positions that hold a's / b's leaves (the truth-table
// phases of rewrite.cu:195-201, read off the walk itself).  Selects only, no
// branches: c is written whether or not the union fits in 4 leaves (the
// return value says), so a failed merge leaves leaves of a and b in c.
__device__ __forceinline__ int merge_ordered(const DCut &a, const DCut &b, DCut &c, int &phA, int &phB) {
    int i = 0, k = 0, n = 0;
    phA = 0;
    phB = 0;
    uint32_t o[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const bool ha = i < a.nLeaves, hb = k < b.nLeaves;
        const uint32_t ai = sel4(a.leaves, i < 3 ? i : 3), bk = sel4(b.leaves, k < 3 ? k : 3);
        const bool takeA = ha & (!hb | (ai <= bk));
        const bool takeB = hb & (!ha | (bk <= ai));
        o[s] = (ha | hb) ? (takeA ? ai : bk) : 0u;
        phA |= (int)takeA << s;
        phB |= (int)takeB << s;
        n += (int)(ha | hb);
        i += (int)takeA;
        k += (int)takeB;
    }
    // positions past the count repeat the first leaf (the containment tests of
    // the slot replay rely on it; stored records zero them)
#pragma unroll
    for (int p = 0; p < 4; p++) c.leaves[p] = p < n ? o[p] : o[0];
    c.nLeaves = n;
    return (i >= a.nLeaves) & (k >= b.nLeaves);
}

// Placement of truth-table variables (tt_place.h): one entry per phase,
// held in lanes 0..15 of two registers and read with ds_bpermute, so a lane
// finds its phase's swaps without a memory access and without a loop.
__constant__ uint64_t kTtPlace[16] = {
    tt_place_entry(0),  tt_place_entry(1),  tt_place_entry(2),  tt_place_entry(3),
    tt_place_entry(4),  tt_place_entry(5),  tt_place_entry(6),  tt_place_entry(7),
    tt_place_entry(8),  tt_place_entry(9),  tt_place_entry(10), tt_place_entry(11),
    tt_place_entry(12), tt_place_entry(13), tt_place_entry(14), tt_place_entry(15)};

struct PlaceTab {
    uint32_t lo, hi;    // lane i (i < 16): kTtPlace[i]
};

__device__ __forceinline__ PlaceTab place_tab(int lane) {
    const uint64_t e = kTtPlace[lane & 15];
    return PlaceTab{(uint32_t)e, (uint32_t)(e >> 32)};
}

// the entry of `phase` (0..15); every lane of the wavefront must be active
__device__ __forceinline__ uint64_t place_of(const PlaceTab &tab, int phase) {
    const int a = phase << 2;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)tab.lo);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)tab.hi);
    return (uint64_t)hi << 32 | lo;
}

// Support reduction (rewrite.cu:223-266): the leaves the table does not
// depend on are dropped, the others keep their order, and the table's
// variables move down with them.  Without branches per lane: when every leaf
// stays, the shrink has no swaps and the compaction keeps the leaves where
// they are (positions past the count already repeat the first leaf, as after
// the merge).  The shrink and the compaction run only if some lane whose
// result is used (`used_here`) shrinks; otherwise every such lane's cut is
// already its result, and the others' results are not read.  r holds the
// leaves' reference counts and moves with them.  Returns 1 if the cut shrank
// (0 in lanes left as they were).
__device__ __forceinline__ int minimize_support(DCut &c, const PlaceTab &tab, bool used_here, int (&r)[4]) {
    const uint32_t m0[4] = {0x5555, 0x3333, 0x0F0F, 0x00FF};
    const uint32_t m1[4] = {0xAAAA, 0xCCCC, 0xF0F0, 0xFF00};
    const uint32_t t = c.tt & 0xFFFF;
    int phase = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
        phase |= (int)((i < c.nLeaves) & ((t & m0[i]) != ((t & m1[i]) >> (1 << i)))) << i;
    const int nL = __popc((uint32_t)phase);
    if (__ballot(used_here && nL != c.nLeaves) == 0) return 0;     // wave-uniform
    c.tt = tt_place_shrink(t, place_of(tab, phase)) & 0xFFFF;
    // compaction: kept leaf i moves to slot popc(phase below i); the slots past
    // the new count repeat the first leaf
    uint32_t o[4];
    int q[4];
#pragma unroll
    for (int p = 0; p < 4; p++) { o[p] = c.leaves[p]; q[p] = r[p]; }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int pos = __popc((uint32_t)phase & ((1u << i) - 1u));
#pragma unroll
        for (int p = 0; p < 4; p++)
            if (p == pos && ((phase >> i) & 1)) { o[p] = c.leaves[i]; q[p] = r[i]; }
    }
#pragma unroll
    for (int p = 0; p < 4; p++) { c.leaves[p] = p < nL ? o[p] : o[0]; r[p] = q[p]; }
    const int shrunk = nL != c.nLeaves;
    c.nLeaves = nL;
    return shrunk;
}

__device__ void store_cuts(uint32_t *out, const DCut *cs) {
    for (int s = 0; s < NCUT; s++) {
        const DCut &c = cs[s];
        uint32_t *o = out + 5 * s;
        for (int i = 0; i < 4; i++) o[i] = (c.used && i < c.nLeaves) ? c.leaves[i] : 0u;
        o[4] = c.used ? ((c.tt & 0xFFFF) | (c.value & 0x7FF) << 16 | (uint32_t)(c.nLeaves & 7) << 27 | 1u << 30) : 0u;
    }
}

__global__ void k_cuts_pi(int32_t nPIs, const int32_t *__restrict__ nRef, uint32_t *cuts) {
    int64_t v = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v > nPIs) return;
    DCut c[NCUT];
    for (int s = 0; s < NCUT; s++) { c[s].used = 0; c[s].nLeaves = 0; c[s].tt = 0; c[s].value = 0; }
    c[0].used = 1; c[0].tt = 0xAAAA; c[0].nLeaves = 1; c[0].leaves[0] = (uint32_t)v;
    c[0].value = cut_value(c[0], nRef);
    store_cuts(cuts + (size_t)v * NCUT * 5, c);
}

// One wavefront per node.  Lane l = 8*i + j owns the pair (cut i of fanin0,
// cut j of fanin1): the 64 merges, truth tables, support reductions and cut
// values are computed in parallel.  The reference inserts the candidates one
// by one in (i, j) order with an order-dependent slot policy (FindCut evicts,
// CutFilter drops dominated cuts, rewrite.cu:84-193, 268-294); that replay runs
// wave-uniformly over the surviving candidates with the 8 slots held in lanes
// 0..7, using ballots for the per-slot tests, so the result equals the
// sequential one exactly.
constexpr int CUTW_BLOCK = 256;
// the per-level kernel runs one wavefront per workgroup: a level's ~200 nodes
// then spread over as many CUs, each wavefront with its own L1 and memory
// pipeline (4 wavefronts per workgroup measured 10% slower on the 10M AIG)
constexpr int CUTL_BLOCK = 64;
enum : int { C_SKIP = 0, C_FAIL = 1, C_OK = 2 };

__device__ __forceinline__ int rl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
// x with lane l replaced by the wave-uniform v: the llvm.amdgcn.writelane
// intrinsic (v_writelane_b32), which this compiler has no builtin name for
extern "C" __device__ int gls_writelane_i32(int v, int lane, int old) __asm("llvm.amdgcn.writelane.i32");
__device__ __forceinline__ uint32_t wl(uint32_t v, int l, uint32_t x) {
    return (uint32_t)gls_writelane_i32((int)v, l, (int)x);
}

// Minimum of v over lanes 0..7, valid in lane 0: two quad_perm steps (xor 1,
// xor 2) and row rotations by 4 and 12, so lane 0 also sees lanes 4..7 in
// either rotation direction.  DPP moves, no LDS round trip.
__device__ __forceinline__ uint32_t min8_lane0(uint32_t v) {
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)v, 0xB1, 0xF, 0xF, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)v, 0x4E, 0xF, 0xF, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)v, 0x124, 0xF, 0xF, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)v, 0x12C, 0xF, 0xF, false));
    return v;
}

// One slot of the cut set under construction, held by lane = slot index; which
// slots are in use is the wave-uniform bit mask `used` of the replay.
struct Slot {
    uint32_t l0, l1, l2, l3, tt, value;
    int n;
    uint32_t sig;       // bit (leaf & 31) of every leaf: a prefilter for containment
};


// Leaf signature of a merged or support-minimized cut: its leaf list is padded
// with its first leaf, so a padded position adds no bit; an empty cut has none.
__device__ __forceinline__ uint32_t sig4(uint32_t n0, uint32_t n1, uint32_t n2, uint32_t n3, int nn) {
    const uint32_t s = (1u << (n0 & 31)) | (1u << (n1 & 31)) | (1u << (n2 & 31)) | (1u << (n3 & 31));
    return nn > 0 ? s : 0u;
}

// CutFilter (rewrite.cu:178-193) of the cut in slot `id` (leaves N, nn) against
// the other used slots.  Scanning slots in order, a larger slot that contains N
// is dropped; the first slot of no more leaves that N contains drops N and ends
// the scan.  Returns true if N was dropped.
__device__ __forceinline__ bool filter_wave(Slot &s, int lane, int id, uint32_t n0, uint32_t n1, uint32_t n2,
                                            uint32_t n3, int nn, uint32_t nsig, uint32_t &used) {
    const bool mine = lane < NCUT && lane != id && ((used >> (lane & 7)) & 1u);
    const bool bigger = s.n > nn;
    // containment implies containment of the leaf signatures (CutFilter's own
    // prefilter, rewrite.cu:181, 185): when no slot passes it, nothing changes
    const bool maybe = mine & ((s.sig & nsig) == (bigger ? nsig : s.sig));
    if (__ballot(maybe) == 0) return false;
    // exact containment both ways from one 4x4 equality matrix: leaf lists are
    // padded with their first leaf past their count, so a padded entry repeats
    // a real one and no count enters the test (bitwise, without branches)
    const bool e00 = n0 == s.l0, e01 = n0 == s.l1, e02 = n0 == s.l2, e03 = n0 == s.l3;
    const bool e10 = n1 == s.l0, e11 = n1 == s.l1, e12 = n1 == s.l2, e13 = n1 == s.l3;
    const bool e20 = n2 == s.l0, e21 = n2 == s.l1, e22 = n2 == s.l2, e23 = n2 == s.l3;
    const bool e30 = n3 == s.l0, e31 = n3 == s.l1, e32 = n3 == s.l2, e33 = n3 == s.l3;
    const bool nInS = (nn == 0) | ((e00 | e01 | e02 | e03) & (e10 | e11 | e12 | e13) & (e20 | e21 | e22 | e23) &
                                   (e30 | e31 | e32 | e33));
    const bool sInN = (s.n == 0) | ((nn != 0) & (e00 | e10 | e20 | e30) & (e01 | e11 | e21 | e31) &
                                    (e02 | e12 | e22 | e32) & (e03 | e13 | e23 | e33));
    const bool d1 = maybe & bigger & nInS;
    const bool d2 = maybe & !bigger & sInN;
    const uint64_t m2 = __ballot(d2);
    const int k = m2 ? __ffsll((unsigned long long)m2) - 1 : NCUT;
    used &= ~(uint32_t)(__ballot(d1 && lane < k) & 0xFFull);
    if (k < NCUT) {
        used &= ~(1u << id);
        return true;
    }
    return false;
}

// Fence-free record format of the dataflow schedule: every 64-bit word of a
// slot carries a valid bit, so a reader polls the words themselves with
// relaxed device-scope atomics and needs no ordering between words.
//   w0 = leaf0 | leaf1 << 31 | V,  w1 = leaf2 | leaf3 << 31 | V,  w2 = meta | V
constexpr unsigned long long TAG_V = 1ull << 63;
constexpr uint32_t DF_SPIN_LIMIT = 1u << 22;

__device__ __forceinline__ bool load_tagged(const unsigned long long *rec, DCut &c) {
    const unsigned long long w0 = __hip_atomic_load(rec + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long w1 = __hip_atomic_load(rec + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long w2 = __hip_atomic_load(rec + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!(w0 & w1 & w2 & TAG_V)) return false;
    c.leaves[0] = (uint32_t)(w0 & 0x7FFFFFFFull);
    c.leaves[1] = (uint32_t)((w0 >> 31) & 0x7FFFFFFFull);
    c.leaves[2] = (uint32_t)(w1 & 0x7FFFFFFFull);
    c.leaves[3] = (uint32_t)((w1 >> 31) & 0x7FFFFFFFull);
    const uint32_t m = (uint32_t)w2;
    c.tt = m & 0xFFFF;
    c.value = (m >> 16) & 0x7FF;
    c.nLeaves = (m >> 27) & 7;
    c.used = (m >> 30) & 1;
    return true;
}

__device__ __forceinline__ void store_tagged(unsigned long long *rec, uint32_t l0, uint32_t l1, uint32_t l2,
                                             uint32_t l3, uint32_t meta) {
    __hip_atomic_store(rec + 0, (unsigned long long)l0 | (unsigned long long)l1 << 31 | TAG_V, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(rec + 1, (unsigned long long)l2 | (unsigned long long)l3 << 31 | TAG_V, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(rec + 2, (unsigned long long)meta | TAG_V, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// cut_node record modes
enum : int { REC_PLAIN = 0, REC_TAGGED = 1 };

// Cut set of node idx (fanin literals l0, l1), computed by one whole wavefront
// (lane = thread lane).  MODE REC_PLAIN: the fanins' 8-slot records (20 bytes
// per slot) start at ra and rb, in global memory or LDS, and the result goes to
// out in the same format (and to out2 as well when it is not null).  MODE
// REC_TAGGED: fanin cut sets are polled in the fence-free format and the result
// is written in it (tcuts, 24 bytes per slot).
// Returns false if a poll exceeded its bound (*err set), true otherwise.
template <int MODE>
__device__ __forceinline__ bool cut_node(const int idx, const int lane, const uint32_t l0, const uint32_t l1,
                                         const uint32_t *ra, const uint32_t *rb, uint32_t *out,
                                         const int32_t *__restrict__ nRef, unsigned long long *tcuts, uint32_t *err,
                                         const uint8_t *__restrict__ sigBits, uint32_t *out2 = nullptr) {
    const uint32_t in0 = LID(l0), in1 = LID(l1), c0 = LC(l0), c1 = LC(l1);
    const PlaceTab tab = place_tab(lane);
    DCut a, b;
    if (MODE == REC_TAGGED) {
        const unsigned long long *ra = tcuts + ((size_t)in0 * NCUT + (lane >> 3)) * 3;
        const unsigned long long *rb = tcuts + ((size_t)in1 * NCUT + (lane & 7)) * 3;
        uint32_t spins = 0;
        for (;;) {
            const bool ok = load_tagged(ra, a) && load_tagged(rb, b);
            if (__ballot(!ok) == 0) break;
            if (++spins > DF_SPIN_LIMIT ||
                ((spins & 1023u) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                if (lane == 0) atomicOr(err, 4u);
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    } else {
        unpack_cut(ra + (lane >> 3) * 5, a);
        unpack_cut(rb + (lane & 7) * 5, b);
    }
    // ---- phase 1: every pair in parallel
    int status = C_SKIP, shrunk = 0;
    DCut m, mm;
    m.leaves[0] = m.leaves[1] = m.leaves[2] = m.leaves[3] = 0;
    m.nLeaves = 0; m.tt = 0; m.value = 0; m.used = 0;
    mm = m;
    {
        // every lane runs the whole computation (no divergent branches; a
        // failed merge leaves leaves of a and b in m, so every leaf read below
        // is a valid id) and the status says which results the replay may use
        const bool cand = a.used && b.used && __popc(cut_sign_of(a, sigBits) | cut_sign_of(b, sigBits)) <= 4;
        int phA, phB;
        const int ok = merge_ordered(a, b, m, phA, phB);
        // the merged leaves' reference counts, in flight during the truth
        // tables; a reduced cut's leaves are some of them
        int r[4];
        leaf_refs(m, nRef, r);
        uint32_t t0 = c0 ? ~a.tt : a.tt, t1 = c1 ? ~b.tt : b.tt;
        t0 = tt_place_stretch(t0, place_of(tab, phA));
        t1 = tt_place_stretch(t1, place_of(tab, phB));
        m.tt = (t0 & t1) & 0xFFFF;
        mm = m;
        shrunk = minimize_support(mm, tab, cand && ok, r);
        mm.value = cut_value_of(mm.nLeaves, r);
        status = !cand ? C_SKIP : !ok ? C_FAIL : C_OK;
    }
    // per-candidate words read by the replay: status, shrink flag, leaf counts
    // and value in one; both truth tables in one; both leaf signatures
    const uint32_t info = (uint32_t)status | (uint32_t)shrunk << 2 | (uint32_t)(m.nLeaves & 7) << 3 |
                          (uint32_t)(mm.nLeaves & 7) << 6 | (mm.value & 0x7FFu) << 9;
    const uint32_t tts = (m.tt & 0xFFFFu) | (mm.tt & 0xFFFFu) << 16;
    const uint32_t msig = sig4(m.leaves[0], m.leaves[1], m.leaves[2], m.leaves[3], m.nLeaves);
    const uint32_t mmsig = sig4(mm.leaves[0], mm.leaves[1], mm.leaves[2], mm.leaves[3], mm.nLeaves);
    // ---- phase 2: replay the insertions in (i, j) order
    Slot s;
    s.l0 = lane == 0 ? (uint32_t)idx : 0u;
    s.l1 = s.l2 = s.l3 = s.l0;          // padded with the first leaf
    s.n = lane == 0 ? 1 : 0;
    s.tt = lane == 0 ? 0xAAAAu : 0u;
    s.value = lane == 0 ? 1001u : 0u;
    s.sig = lane == 0 ? 1u << ((uint32_t)idx & 31) : 0u;
    uint32_t used = 1u;                 // slots in use (bit = slot), wave-uniform
    uint64_t todo = __ballot(status != C_SKIP);
    while (todo) {
        const int l = __ffsll((unsigned long long)todo) - 1;
        todo &= todo - 1;
        const uint32_t inf = (uint32_t)rl((int)info, l);
        const int st = (int)(inf & 3u);
        // FindCut (rewrite.cu:84-109)
        int id;
        if (used != 0xFFu) {
            id = __ffs(~used & 0xFFu) - 1;
        } else {
            // categories in order (more than 2 leaves, 2, fewer), then the
            // smallest value, then the first slot: one minimum of that key
            const uint32_t cat = s.n > 2 ? 0u : s.n == 2 ? 1u : 2u;
            const uint32_t key = lane < NCUT ? (cat << 14 | (s.value & 0x7FF) << 3 | (uint32_t)lane) : 0xFFFFFFFFu;
            id = (int)((uint32_t)rl((int)min8_lane0(key), 0) & 7u);
        }
        used &= ~(1u << id);
        if (st == C_FAIL) continue;
        // the slot takes the merged cut (selects on one lane compare)
        const int nn = (int)((inf >> 3) & 7u);
        const uint32_t n0 = (uint32_t)rl((int)m.leaves[0], l), n1 = (uint32_t)rl((int)m.leaves[1], l);
        const uint32_t n2 = (uint32_t)rl((int)m.leaves[2], l), n3 = (uint32_t)rl((int)m.leaves[3], l);
        const uint32_t nsig = (uint32_t)rl((int)msig, l);
        // slot id takes the values (all wave-uniform): one v_writelane each
        s.l0 = wl(n0, id, s.l0); s.l1 = wl(n1, id, s.l1); s.l2 = wl(n2, id, s.l2); s.l3 = wl(n3, id, s.l3);
        s.n = (int)wl((uint32_t)nn, id, (uint32_t)s.n); s.sig = wl(nsig, id, s.sig);
        used |= 1u << id;
        if (filter_wave(s, lane, id, n0, n1, n2, n3, nn, nsig, used)) continue;
        const uint32_t ttw = (uint32_t)rl((int)tts, l);
        s.tt = wl(ttw & 0xFFFFu, id, s.tt);
        int fn = nn;
        if ((inf >> 2) & 1u) {
            fn = (int)((inf >> 6) & 7u);
            const uint32_t q0 = (uint32_t)rl((int)mm.leaves[0], l), q1 = (uint32_t)rl((int)mm.leaves[1], l);
            const uint32_t q2 = (uint32_t)rl((int)mm.leaves[2], l), q3 = (uint32_t)rl((int)mm.leaves[3], l);
            const uint32_t qsig = (uint32_t)rl((int)mmsig, l);
            s.l0 = wl(q0, id, s.l0); s.l1 = wl(q1, id, s.l1); s.l2 = wl(q2, id, s.l2); s.l3 = wl(q3, id, s.l3);
            s.n = (int)wl((uint32_t)fn, id, (uint32_t)s.n); s.tt = wl(ttw >> 16, id, s.tt); s.sig = wl(qsig, id, s.sig);
            filter_wave(s, lane, id, q0, q1, q2, q3, fn, qsig, used);
        }
        s.value = wl((inf >> 9) & 0x7FFu, id, s.value);
        if (fn < 2) break;
    }
    // ---- phase 3: lanes 0..7 store the slots
    if (lane < NCUT) {
        const bool u = (used >> lane) & 1u;
        const uint32_t o0 = (u && s.n > 0) ? s.l0 : 0u, o1 = (u && s.n > 1) ? s.l1 : 0u;
        const uint32_t o2 = (u && s.n > 2) ? s.l2 : 0u, o3 = (u && s.n > 3) ? s.l3 : 0u;
        const uint32_t meta = u ? ((s.tt & 0xFFFF) | (s.value & 0x7FF) << 16 | (uint32_t)(s.n & 7) << 27 | 1u << 30) : 0u;
        if (MODE == REC_TAGGED) {
            store_tagged(tcuts + ((size_t)idx * NCUT + lane) * 3, o0, o1, o2, o3, meta);
        } else {
            uint32_t *o = out + lane * 5;
            o[0] = o0; o[1] = o1; o[2] = o2; o[3] = o3; o[4] = meta;
            if (out2) {
                uint32_t *q = out2 + lane * 5;
                q[0] = o0; q[1] = o1; q[2] = o2; q[3] = o3; q[4] = meta;
            }
        }
    }
    return true;
}

// Cut set of node idx from the records of its fanins in `cuts`, into `cuts`.
__device__ __forceinline__ void cut_node_global(const int idx, const int lane, const uint32_t *__restrict__ f0,
                                                const uint32_t *__restrict__ f1, const int32_t *__restrict__ nRef,
                                                uint32_t *cuts, const uint8_t *__restrict__ sigBits) {
    const uint32_t l0 = f0[idx], l1 = f1[idx];
    cut_node<REC_PLAIN>(idx, lane, l0, l1, cuts + (size_t)LID(l0) * NCUT * 5, cuts + (size_t)LID(l1) * NCUT * 5,
                        cuts + (size_t)idx * NCUT * 5, nRef, nullptr, nullptr, sigBits);
}

__global__ void __launch_bounds__(CUTL_BLOCK) k_cuts_level(int32_t start, int32_t count,
                                                           const uint32_t *__restrict__ f0,
                                                           const uint32_t *__restrict__ f1,
                                                           const int32_t *__restrict__ nRef, uint32_t *cuts,
                                                           const uint8_t *__restrict__ sigBits) {
    const int k = blockIdx.x * (CUTL_BLOCK / 64) + (threadIdx.x >> 6);
    if (k >= count) return;    // uniform per wavefront
    cut_node_global(start + k, threadIdx.x & 63, f0, f1, nRef, cuts, sigBits);
}

// Two levels in one launch.  Workgroups 0..n0-1 compute the nodes of the first
// level (ids s0..s0+n0-1) as k_cuts_level does.  Workgroup n0 + k takes node
// s1 + k of the second level: wavefront w first recomputes the cut set of
// fanin w if that fanin lies on the first level (its own fanins are on earlier
// levels, whose records are complete), into LDS; after the barrier wavefront 0
// computes the node from those LDS records and the global records of fanins on
// earlier levels.  Every computation is the deterministic cut_node, so the
// recomputed sets equal the stored ones, and one launch replaces two dependent
// ones (a dependent launch costs about as much as a node's whole computation).
constexpr int CUTP_BLOCK = 128;
__global__ void __launch_bounds__(CUTP_BLOCK) k_cuts_pair(int32_t s0, int32_t n0, int32_t s1,
                                                          const uint32_t *__restrict__ f0,
                                                          const uint32_t *__restrict__ f1,
                                                          const int32_t *__restrict__ nRef, uint32_t *cuts,
                                                          const uint8_t *__restrict__ sigBits) {
    __shared__ uint32_t s_rec[2][NCUT * 5];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = (int)blockIdx.x;
    if (k < n0) {                       // uniform per workgroup
        if (w == 0) cut_node_global(s0 + k, lane, f0, f1, nRef, cuts, sigBits);
        return;
    }
    const int v = s1 + (k - n0);
    const uint32_t l0 = f0[v], l1 = f1[v];
    const uint32_t in0 = LID(l0), in1 = LID(l1);
    const uint32_t mine = w == 0 ? in0 : in1;
    if (mine >= (uint32_t)s0) {         // on the first level (fanins of v are below s1)
        const uint32_t g0 = f0[mine], g1 = f1[mine];
        cut_node<REC_PLAIN>((int)mine, lane, g0, g1, cuts + (size_t)LID(g0) * NCUT * 5,
                            cuts + (size_t)LID(g1) * NCUT * 5, s_rec[w], nRef, nullptr, nullptr, sigBits);
    }
    __syncthreads();
    if (w == 0) {
        const uint32_t *ra = in0 >= (uint32_t)s0 ? s_rec[0] : cuts + (size_t)in0 * NCUT * 5;
        const uint32_t *rb = in1 >= (uint32_t)s0 ? s_rec[1] : cuts + (size_t)in1 * NCUT * 5;
        cut_node<REC_PLAIN>(v, lane, l0, l1, ra, rb, cuts + (size_t)v * NCUT * 5, nRef, nullptr, nullptr, sigBits);
    }
}

// Three levels in one launch (ids s0..s1-1, s1..s2-1, s2..e-1), one workgroup
// per node.  A node of the first level is computed as in k_cuts_level, one of
// the second as in k_cuts_pair.  A node v of the third level recomputes its
// cone in its workgroup: the bottom nodes (first-level fanins of v and of v's
// second-level fanins, at most 4, one wavefront each), then the middle nodes
// (v's second-level fanins, at most 2), then v, with a barrier between the
// steps and the recomputed sets held in LDS.  Fanins below the launch are read
// from the global records, which earlier launches completed.
constexpr int CUT3_BLOCK = 256;
__device__ __forceinline__ const uint32_t *rec_or_lds(uint32_t x, const uint32_t *cuts, uint32_t (*s_rec)[NCUT * 5],
                                                      uint32_t b0, uint32_t b1, uint32_t b2, uint32_t m0,
                                                      uint32_t m1, uint32_t s0, uint32_t s1) {
    if (x >= s1) return s_rec[x == m0 ? 4 : 5];
    if (x >= s0) return s_rec[x == b0 ? 0 : x == b1 ? 1 : x == b2 ? 2 : 3];
    return cuts + (size_t)x * NCUT * 5;
}

__global__ void __launch_bounds__(CUT3_BLOCK) k_cuts_triple(int32_t s0_, int32_t s1_, int32_t s2_,
                                                            const uint32_t *__restrict__ f0,
                                                            const uint32_t *__restrict__ f1,
                                                            const int32_t *__restrict__ nRef, uint32_t *cuts,
                                                            const uint8_t *__restrict__ sigBits) {
    __shared__ uint32_t s_rec[6][NCUT * 5];     // bottom nodes 0..3, middle nodes 4..5
    const uint32_t s0 = (uint32_t)s0_, s1 = (uint32_t)s1_, s2 = (uint32_t)s2_;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t v = s0 + blockIdx.x;
    // issue priority by chain length: a third-level node's workgroup runs three
    // dependent steps and bounds the launch, a first-level node's runs one
    if (v < s1) {                               // uniform per workgroup
        __builtin_amdgcn_s_setprio(0);
        if (w == 0) cut_node_global((int)v, lane, f0, f1, nRef, cuts, sigBits);
        return;
    }
    if (v >= s2) __builtin_amdgcn_s_setprio(2);
    else __builtin_amdgcn_s_setprio(1);
    const uint32_t l0 = f0[v], l1 = f1[v];
    const uint32_t a = LID(l0), b = LID(l1);
    const uint32_t NONE = 0xFFFFFFFFu;
    // middle nodes (second level) and bottom nodes (first level) of the cone
    uint32_t m0 = NONE, m1 = NONE, ml0 = 0u, ml1 = 0u, ml2 = 0u, ml3 = 0u;    // fanin literals of m0, m1
    uint32_t b0 = NONE, b1 = NONE, b2 = NONE, b3 = NONE;
    int nb = 0;
    // appends x if it is a first-level node not yet listed (selects, so the
    // list stays in registers)
    auto add_bot = [&](uint32_t x) {
        const bool add = x >= s0 && x != b0 && x != b1 && x != b2;
        b0 = add && nb == 0 ? x : b0;
        b1 = add && nb == 1 ? x : b1;
        b2 = add && nb == 2 ? x : b2;
        b3 = add && nb == 3 ? x : b3;
        nb += add ? 1 : 0;
    };
    if (v >= s2) {
        if (a >= s1) m0 = a;
        if (b >= s1) { if (m0 == NONE) m0 = b; else m1 = b; }
        if (m0 != NONE) { ml0 = f0[m0]; ml1 = f1[m0]; }
        if (m1 != NONE) { ml2 = f0[m1]; ml3 = f1[m1]; }
        if (a < s1) add_bot(a);
        if (b < s1) add_bot(b);
        if (m0 != NONE) { add_bot(LID(ml0)); add_bot(LID(ml1)); }
        if (m1 != NONE) { add_bot(LID(ml2)); add_bot(LID(ml3)); }
    } else {
        add_bot(a);
        add_bot(b);
    }
    // step 1: bottom nodes from the global records
    const uint32_t mine = w == 0 ? b0 : w == 1 ? b1 : w == 2 ? b2 : b3;
    if (w < nb) {
        const uint32_t g0 = f0[mine], g1 = f1[mine];
        cut_node<REC_PLAIN>((int)mine, lane, g0, g1, cuts + (size_t)LID(g0) * NCUT * 5,
                            cuts + (size_t)LID(g1) * NCUT * 5, s_rec[w], nRef, nullptr, nullptr, sigBits);
    }
    __syncthreads();
    // step 2: middle nodes
    if (v >= s2) {
        const uint32_t mm = w == 0 ? m0 : m1;
        if (w < 2 && mm != NONE) {
            const uint32_t g0 = w == 0 ? ml0 : ml2, g1 = w == 0 ? ml1 : ml3;
            cut_node<REC_PLAIN>((int)mm, lane, g0, g1, rec_or_lds(LID(g0), cuts, s_rec, b0, b1, b2, m0, m1, s0, s1),
                                rec_or_lds(LID(g1), cuts, s_rec, b0, b1, b2, m0, m1, s0, s1), s_rec[4 + w], nRef,
                                nullptr, nullptr, sigBits);
        }
        __syncthreads();
    }
    // step 3: the node itself
    if (w == 0)
        cut_node<REC_PLAIN>((int)v, lane, l0, l1, rec_or_lds(a, cuts, s_rec, b0, b1, b2, m0, m1, s0, s1),
                            rec_or_lds(b, cuts, s_rec, b0, b1, b2, m0, m1, s0, s1), cuts + (size_t)v * NCUT * 5, nRef,
                            nullptr, nullptr, sigBits);
}

// ---- planned schedule (GLS_RW_PLAN_CUTS): K levels per launch
// A launch covers K consecutive levels (positions 0..K-1).  The in-launch
// cone of a node (its transitive fanins inside the launch) is fixed by the
// graph alone, so the cones are listed once per pass before the launches:
//   - every node of the top position is a root; going down the positions, a
//     node that no higher root's cone contains becomes a root as well;
//   - a root's workgroup recomputes its cone position by position (one
//     wavefront per node, records in LDS, a barrier between positions), and
//     every cone node is stored to the global records by exactly one of the
//     workgroups that recompute it (the smallest root id), the root by its own;
//   - position-0 nodes that no cone contains are packed four to a workgroup.
// Each workgroup reads a header and its entries (node, fanin literals, where
// each fanin's record is: an LDS slot of this task or the global records)
// with scalar loads, so no fanin is looked up at run time.  Every record is
// computed by the deterministic cut_node, so the results are bit-identical to
// one launch per level.  On the 10M-AND workload with K = 3 a launch runs
// ~328 workgroups instead of k_cuts_triple's ~617.
// Header (uint4): x = first entry, y = entry count | top position << 16,
//   z, w = node count of positions 0..3 and 4..7 (8 bits each).
// Entry (uint4): x = node id | stored-globally << 31 | kept-in-LDS << 30,
//   y, z = fanin literals, w = fanin sources (8 bits each: an entry index of
//   this task, or PLAN_GLOBAL).
#include "cut_cone.h"

// Four levels in one launch (ids [s0, s1), [s1, s2), [s2, s3), [s3, e)), one
// workgroup of four wavefronts per node, as k_cuts_triple with one more
// level: a node recomputes its in-launch cone (at most 8, 4 and 2 nodes on
// the launch's three lower levels, listed by one thread into LDS) level by
// level, one wavefront per cone node (the lowest level's nodes beyond four in
// a second batch), records in LDS, a barrier between the levels; its own set
// goes to the global records.  One launch replaces four dependent ones at the
// price of larger cones.  (Eight wavefronts per workgroup, one per lowest-level
// node, put two of a workgroup's own wavefronts on every SIMD and measured
// 17.9 us per launch against 10.9 for three levels.)
constexpr int CUT4_BLOCK = 256;   // (measured 159.5 ms per pass on the 10M AIG against 138.1 for three levels)

// the record of fanin x for a cone node: its LDS slot if the cone lists it,
// else the global records (x below the launch)
__device__ __forceinline__ const uint32_t *cone4_rec(uint32_t x, const uint32_t *s, const Cone4 &c,
                                                     uint32_t (*s_rec)[NCUT * 5], const uint32_t *cuts) {
    const int p = cone4_level(x, s);
    if (p >= 0 && p <= 2) {
        const uint32_t base = p == 0 ? 0u : p == 1 ? 8u : 12u;
        for (uint32_t k = 0; k < c.n[p]; k++)
            if (c.id[p][k] == x) return s_rec[base + k];
    }
    return cuts + (size_t)x * NCUT * 5;
}

__global__ void __launch_bounds__(CUT4_BLOCK) k_cuts_quad(int32_t s0_, int32_t s1_, int32_t s2_, int32_t s3_,
                                                          const uint32_t *__restrict__ f0,
                                                          const uint32_t *__restrict__ f1,
                                                          const int32_t *__restrict__ nRef, uint32_t *cuts,
                                                          const uint8_t *__restrict__ sigBits) {
    __shared__ uint32_t s_rec[14][NCUT * 5];    // level 0: 0..7, level 1: 8..11, level 2: 12..13
    __shared__ Cone4 s_cone;
    const uint32_t s[4] = {(uint32_t)s0_, (uint32_t)s1_, (uint32_t)s2_, (uint32_t)s3_};
    const int lane = threadIdx.x & 63;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t v = s[0] + blockIdx.x;
    const int pv = cone4_level(v, s);
    // issue priority by chain length: a top-level node's workgroup runs four
    // dependent steps and bounds the launch
    if (pv == 0) {                              // uniform per workgroup
        __builtin_amdgcn_s_setprio(0);
        if (w == 0) cut_node_global((int)v, lane, f0, f1, nRef, cuts, sigBits);
        return;
    }
    if (pv == 3) __builtin_amdgcn_s_setprio(3);
    else if (pv == 2) __builtin_amdgcn_s_setprio(2);
    else __builtin_amdgcn_s_setprio(1);
    if (threadIdx.x == 0) cone4_discover(v, s, f0, f1, s_cone);
    __syncthreads();
    const uint32_t n0 = s_cone.n[0], n1 = s_cone.n[1], n2 = s_cone.n[2];
    // level 0 of the cone: fanins below the launch; nodes 4..7 in a second batch
    if (w < n0) {
        const uint32_t x = s_cone.id[0][w], g0 = s_cone.l0[0][w], g1 = s_cone.l1[0][w];
        cut_node<REC_PLAIN>((int)x, lane, g0, g1, cuts + (size_t)LID(g0) * NCUT * 5,
                            cuts + (size_t)LID(g1) * NCUT * 5, s_rec[w], nRef, nullptr, nullptr, sigBits);
    }
    if (w + 4 < n0) {
        const uint32_t x = s_cone.id[0][w + 4], g0 = s_cone.l0[0][w + 4], g1 = s_cone.l1[0][w + 4];
        cut_node<REC_PLAIN>((int)x, lane, g0, g1, cuts + (size_t)LID(g0) * NCUT * 5,
                            cuts + (size_t)LID(g1) * NCUT * 5, s_rec[w + 4], nRef, nullptr, nullptr, sigBits);
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __syncthreads();
    // level 1
    if (w < n1) {
        const uint32_t x = s_cone.id[1][w], g0 = s_cone.l0[1][w], g1 = s_cone.l1[1][w];
        cut_node<REC_PLAIN>((int)x, lane, g0, g1, cone4_rec(LID(g0), s, s_cone, s_rec, cuts),
                            cone4_rec(LID(g1), s, s_cone, s_rec, cuts), s_rec[8 + w], nRef, nullptr, nullptr,
                            sigBits);
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __syncthreads();
    // level 2
    if (w < n2) {
        const uint32_t x = s_cone.id[2][w], g0 = s_cone.l0[2][w], g1 = s_cone.l1[2][w];
        cut_node<REC_PLAIN>((int)x, lane, g0, g1, cone4_rec(LID(g0), s, s_cone, s_rec, cuts),
                            cone4_rec(LID(g1), s, s_cone, s_rec, cuts), s_rec[12 + w], nRef, nullptr, nullptr,
                            sigBits);
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __syncthreads();
    // the node itself
    if (w == 0) {
        const uint32_t l0 = f0[v], l1 = f1[v];
        cut_node<REC_PLAIN>((int)v, lane, l0, l1, cone4_rec(LID(l0), s, s_cone, s_rec, cuts),
                            cone4_rec(LID(l1), s, s_cone, s_rec, cuts), cuts + (size_t)v * NCUT * 5, nRef, nullptr,
                            nullptr, sigBits);
    }
}

#include "cut_plan.h"

// One planned task per workgroup (the launch's tasks start at `base`).
__global__ void __launch_bounds__(CUT3_BLOCK) k_cuts_plan(const uint4 *__restrict__ heads,
                                                          const uint4 *__restrict__ ents, int32_t base, int32_t K,
                                                          const int32_t *__restrict__ nRef, uint32_t *cuts,
                                                          const uint8_t *__restrict__ sigBits) {
    __shared__ uint32_t s_rec[PLAN_MAX_CONE][NCUT * 5];
    const int lane = threadIdx.x & 63;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint4 h = heads[base + blockIdx.x];
    const uint32_t top = h.y >> 16;
    // issue priority by chain length: the roots of the top position bound the launch
    if (top + 1 >= (uint32_t)K && K > 1) __builtin_amdgcn_s_setprio(2);
    else if (top > 0) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
    uint32_t start = 0;
    bool first = true;
    for (int p = 0; p < K; p++) {
        const uint32_t cnt = (p < 4 ? h.z >> (8 * p) : h.w >> (8 * (p - 4))) & 0xFFu;
        if (cnt == 0) continue;                 // uniform per workgroup
        if (!first) {
            // this wavefront's LDS records are written before any wavefront
            // passes the barrier and reads them
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            __syncthreads();
        }
        first = false;
        for (uint32_t k = w; k < cnt; k += CUT3_BLOCK / 64) {
            const uint4 e = ents[h.x + start + k];
            const uint32_t id = e.x & 0x3FFFFFFFu;
            const uint32_t s0 = e.w & 0xFFu, s1 = (e.w >> 8) & 0xFFu;
            const uint32_t *ra = s0 == PLAN_GLOBAL ? cuts + (size_t)LID(e.y) * NCUT * 5 : s_rec[s0];
            const uint32_t *rb = s1 == PLAN_GLOBAL ? cuts + (size_t)LID(e.z) * NCUT * 5 : s_rec[s1];
            uint32_t *glob = cuts + (size_t)id * NCUT * 5;
            const bool stored = (e.x >> 31) != 0u, kept = ((e.x >> 30) & 1u) != 0u;
            // separate call sites, so that the LDS record is written through an
            // LDS pointer (ds_write, which the barrier waits for; a pointer that
            // is LDS or global at run time would make it a flat store that the
            // barrier does not wait for)
            if (kept)
                cut_node<REC_PLAIN>((int)id, lane, e.y, e.z, ra, rb, s_rec[start + k], nRef, nullptr, nullptr, sigBits,
                                    stored ? glob : nullptr);
            else
                cut_node<REC_PLAIN>((int)id, lane, e.y, e.z, ra, rb, glob, nRef, nullptr, nullptr, sigBits);
        }
        start += cnt;
    }
}

// All levels in one cooperative launch: the wavefronts of the grid stride over
// each level's nodes, and a grid barrier separates the levels.
__global__ void __launch_bounds__(CUTW_BLOCK) k_cuts_levels(int32_t l0, int32_t l1, const int32_t *__restrict__ starts,
                                                            const uint32_t *__restrict__ f0,
                                                            const uint32_t *__restrict__ f1,
                                                            const int32_t *__restrict__ nRef, uint32_t *cuts,
                                                            const uint8_t *__restrict__ sigBits) {
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int nw = gridDim.x * (CUTW_BLOCK / 64);
    const int gw = blockIdx.x * (CUTW_BLOCK / 64) + (threadIdx.x >> 6);
    for (int l = l0; l < l1; l++) {
        const int s = starts[l], e = starts[l + 1];
        for (int v = s + gw; v < e; v += nw)
            cut_node_global(v, threadIdx.x & 63, f0, f1, nRef, cuts, sigBits);
        grid.sync();
    }
}

// XCD-local schedule: one persistent kernel on the 32 CUs of one XCD.
// Workgroups are dispatched round-robin over the XCDs, so workgroups 0, 8, 16,
// ... share XCD 0 and its L2; the others return at once.  Between levels a
// counter barrier (device-scope atomics, about 1 us) replaces a kernel
// boundary (about 4.7 us on this path).  A CU's L1 is not coherent with the
// other CUs' stores: after each barrier one wavefront of the workgroup runs
// buffer_inv sc1, which drops the CU's L1 lines (buffer_inv sc0 does not in
// this mode), so the fanin records written by other CUs are read from L2.
// Every workgroup records its XCD; if they are not all on one, or a barrier
// wait exceeds its bound, ctl[2] is set, every workgroup stops, and the host
// recomputes the cut sets with one launch per level.
constexpr int XCD_WG = 1024;
constexpr int XCD_COUNT = 8;
constexpr uint32_t XCD_SPIN_LIMIT = 1u << 20;

__device__ __forceinline__ uint32_t xcc_id() {
    uint32_t v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & 0xFu;
}

// ctl[0] barrier arrivals, ctl[1] mask of XCDs seen, ctl[2] error
__device__ __forceinline__ bool xcd_barrier(uint32_t *ctl, uint32_t &epoch, uint32_t nwg, uint32_t *s_stop) {
    __builtin_amdgcn_s_waitcnt(0);          // this wavefront's stores have reached L2
    __syncthreads();
    ++epoch;
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(&ctl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t target = epoch * nwg;
        uint32_t spins = 0;
        bool bad = false;
        while (__hip_atomic_load(&ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > XCD_SPIN_LIMIT || __hip_atomic_load(&ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                bad = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (bad) __hip_atomic_fetch_or(&ctl[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        asm volatile("buffer_inv sc1\n\ts_waitcnt vmcnt(0)" ::: "memory");
        *s_stop = bad ? 1u : 0u;
    }
    __syncthreads();
    return *s_stop == 0;
}

__global__ void __launch_bounds__(XCD_WG) k_cuts_xcd(int32_t l0, int32_t l1, const int32_t *__restrict__ starts,
                                                     const uint32_t *__restrict__ f0, const uint32_t *__restrict__ f1,
                                                     const int32_t *__restrict__ nRef, uint32_t *cuts,
                                                     const uint8_t *__restrict__ sigBits, uint32_t nwg, uint32_t *ctl) {
    if (blockIdx.x % XCD_COUNT != 0) return;
    __shared__ uint32_t s_stop;
    const uint32_t wg = blockIdx.x / XCD_COUNT;
    if (wg >= nwg) return;
    const int lane = threadIdx.x & 63;
    // a level's nodes go round-robin over the workgroups (one per CU) first,
    // so a level of n nodes puts about n / nwg wavefronts on each CU
    const int gw = (int)(threadIdx.x >> 6) * (int)nwg + (int)wg;
    const int nw = (int)nwg * (XCD_WG / 64);
    uint32_t epoch = 0;
    if (threadIdx.x == 0) __hip_atomic_fetch_or(&ctl[1], 1u << xcc_id(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!xcd_barrier(ctl, epoch, nwg, &s_stop)) return;
    if (__popc(__hip_atomic_load(&ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 1) {
        if (threadIdx.x == 0) __hip_atomic_fetch_or(&ctl[2], 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    for (int l = l0; l < l1; l++) {
        const int s = starts[l], e = starts[l + 1];
        for (int v = s + gw; v < e; v += nw)
            cut_node_global(v, lane, f0, f1, nRef, cuts, sigBits);
        if (!xcd_barrier(ctl, epoch, nwg, &s_stop)) return;
    }
}

// Dataflow schedule: persistent wavefronts take nodes in level-major id order
// from a ticket counter and wait only for their two fanins' cut sets, polled
// in the fence-free record format; no level waits for the whole previous
// level.  Progress: the smallest unfinished ticket has every fanin finished
// (fanin ids are smaller) and is held by a running wavefront.  Every wait is
// bounded; on overrun *err is set and all wavefronts stop.
__global__ void __launch_bounds__(CUTW_BLOCK) k_cuts_dataflow(int32_t first, int32_t last,
                                                              const uint32_t *__restrict__ f0,
                                                              const uint32_t *__restrict__ f1,
                                                              const int32_t *__restrict__ nRef,
                                                              unsigned long long *tcuts, uint32_t *ticket,
                                                              uint32_t *err, const uint8_t *__restrict__ sigBits) {
    const int lane = threadIdx.x & 63;
    for (;;) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(ticket, 1u);
        t = (uint32_t)__shfl((int)t, 0);
        const int v = first + (int)t;
        if (v >= last) break;
        if (!cut_node<REC_TAGGED>(v, lane, f0[v], f1[v], nullptr, nullptr, nullptr, nRef, tcuts, err, sigBits)) break;
    }
}

// PI cut sets in the fence-free format (slot 0 trivial, others empty but valid).
__global__ void k_cuts_pi_tagged(int32_t nPIs, const int32_t *__restrict__ nRef, unsigned long long *tcuts) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)nPIs * NCUT) return;
    const int64_t v = 1 + i / NCUT;
    const int slot = (int)(i % NCUT);
    uint32_t meta = 0;
    if (slot == 0) meta = 0xAAAAu | (1001u << 16) | (1u << 27) | (1u << 30);
    store_tagged(tcuts + ((size_t)v * NCUT + slot) * 3, slot == 0 ? (uint32_t)v : 0u, 0, 0, 0, meta);
    (void)nRef;
}

// Fence-free records back to the plain 5-word records used by evaluation.
__global__ void k_untag_cuts(int64_t nslots, const unsigned long long *__restrict__ tcuts, uint32_t *cuts) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tid; i < nslots; i += stride) {
        const unsigned long long w0 = tcuts[3 * i], w1 = tcuts[3 * i + 1], w2 = tcuts[3 * i + 2];
        uint32_t *o = cuts + 5 * i;
        if (!(w2 & TAG_V)) { o[0] = o[1] = o[2] = o[3] = o[4] = 0; continue; }
        o[0] = (uint32_t)(w0 & 0x7FFFFFFFull);
        o[1] = (uint32_t)((w0 >> 31) & 0x7FFFFFFFull);
        o[2] = (uint32_t)(w1 & 0x7FFFFFFFull);
        o[3] = (uint32_t)((w1 >> 31) & 0x7FFFFFFFull);
        o[4] = (uint32_t)w2;
    }
}

__global__ void k_nref(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0, const uint32_t *__restrict__ f1,
                       int32_t nPOs, const uint32_t *__restrict__ pos, int32_t *nRef) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + tid; v < n; v += stride) {
        atomicAdd(&nRef[LID(f0[v])], 1);
        atomicAdd(&nRef[LID(f1[v])], 1);
    }
    for (int64_t i = tid; i < nPOs; i += stride) atomicAdd(&nRef[LID(pos[i])], 1);
}

// Structural table of the graph, indexed by fanin: bucket x holds the ANDs
// whose second fanin literal (the one of larger id) is on node x, as entries
// {first fanin literal, node | second literal's complement << 31}.  Buckets
// follow node ids, so the lookups of nearby nodes touch nearby memory (a hash
// table of the same keys spreads them over the whole table, which for the
// batched arena is far larger than the caches).
__global__ void k_table_count(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f1, uint32_t *cnt) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + tid; v < n; v += stride) atomicAdd(&cnt[LID(f1[v])], 1u);
}

__global__ void k_table_fill(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0,
                             const uint32_t *__restrict__ f1, const uint32_t *__restrict__ off, uint32_t *cur,
                             uint2 *ent) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = nPIs + 1 + tid; v < n; v += stride) {
        const uint32_t b = f1[v];
        const uint32_t k = off[LID(b)] + atomicAdd(&cur[LID(b)], 1u);
        ent[k] = make_uint2(f0[v], (uint32_t)v | (b & 1u) << 31);
    }
}

// ------------------------------------------------------------------------
// evaluation
// ------------------------------------------------------------------------
struct GraphView {
    const uint32_t *f0, *f1;
    const int32_t *nRef;
    const uint32_t *level;
    const uint32_t *tOff;       // structural table: bucket offsets (n + 1)
    const uint2 *tEnt;          // and entries (see k_table_fill)
};

// per-thread column of an LDS array laid out [j][lane]
template <typename T> struct ColT {
    T *p;
    __device__ __forceinline__ T &operator[](int j) const { return p[j * EVAL_BLOCK]; }
};
using Col = ColT<int32_t>;

struct Mffc {
    int size;
    int32_t id[MFFC_TAB], num[MFFC_TAB];
};

// The table and the DFS stack below are indexed by constants only (a write at
// a run-time position selects among the entries), so they live in registers
// instead of scratch memory.
__device__ __forceinline__ int mffc_decrease(const GraphView &g, int x, const int *leaves, int nL, Mffc &t) {
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (i < nL && x == leaves[i]) return 1;
    int found = -1;
#pragma unroll
    for (int i = 0; i < MFFC_TAB; i++)
        if (found < 0 && i < t.size && t.id[i] == x) found = i;
    if (found >= 0) {
        int r = 0;
#pragma unroll
        for (int i = 0; i < MFFC_TAB; i++)
            if (i == found) r = --t.num[i];
        return r;
    }
    if (t.size == MFFC_TAB) return 1;
    const int r = g.nRef[x] - 1;
#pragma unroll
    for (int i = 0; i < MFFC_TAB; i++)
        if (i == t.size) { t.id[i] = x; t.num[i] = r; }
    t.size++;
    return r;
}

// CalcMFFC (rewrite.cu:319-326) without recursion: an explicit DFS stack that
// visits fanin0's subtree before decreasing fanin1, as the recursion does.
// At most MFFC_TAB nodes are pushed besides the root (a push needs a new
// table entry), so MFFC_TAB + 2 entries never overflow.  The stack's nodes go
// to the LDS column stk (the caller's match column, free until library
// matching fills it), one access each; their stages (0, 1: fanin to visit
// next, 2: done) are 2-bit fields of one register.
__device__ int calc_mffc(const GraphView &g, int root, const int *leaves, int nL, Mffc &t, Col stk) {
    stk[0] = root;
    uint32_t stg = 0;
    int sp = 1, ans = 1;
    t.size = 0;
#pragma unroll
    for (int i = 0; i < MFFC_TAB; i++) { t.id[i] = 0; t.num[i] = 0; }
    while (sp) {
        const int top = sp - 1;
        const int stage = (int)((stg >> (2 * top)) & 3u);
        if (stage < 2) {
            stg += 1u << (2 * top);
            const int cur = stk[top];
            const int x = (int)LID(stage == 0 ? g.f0[cur] : g.f1[cur]);
            if (mffc_decrease(g, x, leaves, nL, t) == 0) {
                stk[sp] = x;
                stg &= ~(3u << (2 * sp));
                sp++;
                ans++;
            }
        } else {
            sp--;
        }
    }
    return ans;
}

__device__ __forceinline__ int mffc_deleted(const Mffc &t, int x) {
    int r = 0;
    bool found = false;
#pragma unroll
    for (int i = 0; i < MFFC_TAB; i++)
        if (!found && i < t.size && t.id[i] == x) {
            found = true;
            r = t.num[i] == 0;
        }
    return r;
}

// The AND with fanins (in0, c0), (in1, c1) if the graph has one (the smallest
// such id), else -1: a scan of the bucket of the fanin of larger id.
__device__ __forceinline__ int table_lookup(const GraphView &g, int in0, int in1, int c0, int c1) {
    if (in0 > in1) { int x = in0; in0 = in1; in1 = x; x = c0; c0 = c1; c1 = x; }
    const uint32_t lo = 2u * (uint32_t)in0 + (uint32_t)c0, hc = (uint32_t)c1 << 31;
    const uint32_t e = g.tOff[in1 + 1];
    uint32_t best = 0xFFFFFFFFu;
    for (uint32_t k = g.tOff[in1]; k < e; k++) {
        const uint2 x = g.tEnt[k];
        if (x.x == lo && (x.y & 0x80000000u) == hc) best = min(best, x.y & 0x7FFFFFFFu);
    }
    return best == 0xFFFFFFFFu ? -1 : (int)best;
}

// library matching of one cut (rewrite.cu:391-418); returns the class
// LV: storage of the structure nodes' levels, uint16_t when every level is
// below 65535 (more workgroups fit a CU's LDS), else int32_t; the values are
// the same either way.
template <typename LV>
__device__ int match_cut(const GraphView &g, const Lib &L, int id, uint32_t tt, const int *leaves,
                         const Mffc &t, Col match, ColT<LV> mlev, uint32_t &isC, uint64_t &matched) {
    uint32_t w = L.npn[tt];
    int uPhase = w & 31, perm = (w >> 5) & 31, cls = (w >> 10) & 255;
    const uint8_t *pp = L.perms4 + 4 * perm;
    // entries 0..3 (the inputs) and 4..nn+3 (the class's forest nodes) are all
    // written below before they are read; the library's roots and children
    // index only those, so the columns need no clearing.  matched: bit j set
    // when entry j holds an existing node (the inputs always do)
    isC = 0;
    matched = 0xFull;
    for (int j = 0; j < 4; j++) {
        const int q = pp[j];
        const int leaf = q == 0 ? leaves[0] : q == 1 ? leaves[1] : q == 2 ? leaves[2] : leaves[3];
        match[j] = leaf;
        mlev[j] = (LV)g.level[leaf];
        if ((uPhase >> j) & 1) isC |= 1u << j;
    }
    const int nn = lib_nn(L, cls);
    for (int j = 0; j < nn; j++) {
        uint32_t wd = lib_word(L, cls, j);
        int a = wd & 63, b = (wd >> 6) & 63, ca = (wd >> 12) & 1, cb = (wd >> 13) & 1;
        int num = j + 4;
        const int la = (int)mlev[a], lb = (int)mlev[b];
        int lev = 1 + (la > lb ? la : lb);
        const int ma = match[a], mb = match[b];
        int found = -1;
        if (!(ma == -1 || mb == -1 || ma == id || mb == id)) {
            const int nid = table_lookup(g, ma, mb, (int)isc_bit(isC, a) ^ ca, (int)isc_bit(isC, b) ^ cb);
            if (nid != -1 && !mffc_deleted(t, nid)) {
                found = nid;
                lev = (int)g.level[nid];
            }
        }
        match[num] = found;
        mlev[num] = (LV)lev;
        if (found != -1) matched |= 1ull << num;
    }
    return cls;
}

// Forest nodes that replacing with root rt would create (Eval, rewrite.cu:369-374,
// counts them): from rt down, every unmatched node that a created node uses.  Children have
// lower indexes than their parents, so taking the highest unmatched used node
// each time visits exactly the nodes of a descending scan over all indexes,
// in the same order, with one step per created node instead of per index.
__device__ __forceinline__ uint64_t added_set(const Lib &L, int cls, int rt, uint64_t matched) {
    uint64_t used = 1ull << rt, add = 0;
    for (uint64_t cand = used & ~matched; cand; cand = used & ~matched & ~add) {
        const int j = 63 - __clzll((long long)cand);
        add |= 1ull << j;
        const uint32_t wd = lib_word(L, cls, j - 4);
        used |= 1ull << (wd & 63);
        used |= 1ull << ((wd >> 6) & 63);
    }
    return add;
}

__device__ __forceinline__ void cut_leaves4(const uint32_t *p, int *leaves, int &nL, uint32_t &tt) {
    uint32_t m = p[4];
    nL = (m >> 27) & 7;
    tt = m & 0xFFFF;
    for (int i = 0; i < 4; i++) leaves[i] = (int)p[i];
    if (nL == 3) { leaves[3] = 0; nL = 4; }
}

template <typename LV>
__global__ void __launch_bounds__(EVAL_BLOCK) k_eval(int32_t n, int32_t nPIs, GraphView g, Lib L,
                                                      const uint32_t *__restrict__ cuts, int useZeros,
                                                      int4 *best) {
    __shared__ int32_t s_match[NMATCH * EVAL_BLOCK];
    __shared__ LV s_mlev[NMATCH * EVAL_BLOCK];
    int64_t v = nPIs + 1 + blockIdx.x * (int64_t)EVAL_BLOCK + threadIdx.x;
    if (v >= n) return;
    Col match{s_match + threadIdx.x};
    ColT<LV> mlev{s_mlev + threadIdx.x};
    const int id = (int)v;
    int reduction = -1, bestLevel = 99999999, bestCut = -1, bestOut = 0;
    const uint32_t *cs = cuts + (size_t)v * NCUT * 5;
    for (int i = 0; i < NCUT; i++) {
        const uint32_t *p = cs + 5 * i;
        uint32_t meta = p[4];
        if (!((meta >> 30) & 1) || ((meta >> 27) & 7) < 3) continue;
        int leaves[4], nL;
        uint32_t tt;
        cut_leaves4(p, leaves, nL, tt);
        Mffc t;
        int saved = calc_mffc(g, id, leaves, nL, t, match);
        uint32_t isC;
        uint64_t matched;
        int cls = match_cut(g, L, id, tt, leaves, t, match, mlev, isC, matched);
        const int nr = lib_nr(L, cls);
        for (int out = 0; out < nr; out++) {
            int rt = lib_root(L, cls, out);
            if (match[rt] == id) continue;
            int added = __popcll(added_set(L, cls, rt, matched));
            int gain = saved - added;
            if (gain < 0 || (gain == 0 && !useZeros)) continue;
            int rl = (int)mlev[rt];
            if (gain < reduction || (gain == reduction && rl >= bestLevel)) continue;
            reduction = gain;
            bestLevel = rl;
            bestCut = i;
            bestOut = out;
        }
    }
    best[v] = bestCut >= 0 ? make_int4(bestCut, bestOut, reduction, bestLevel) : make_int4(-1, 0, 0, 0);
}

// ------------------------------------------------------------------------
// plans and selection
// ------------------------------------------------------------------------
struct Plan {
    int32_t id, gain;
    int32_t nW, nL, nM, nNew, valid, pad;
    int32_t W[MFFC_TAB + 1], L[4], M[PLAN_MAXM];
};

__device__ uint32_t lib_truth(const Lib &L, int cls, int rt, const uint8_t *pp, uint32_t isC) {
    const uint32_t ev[4] = {0xAAAA, 0xCCCC, 0xF0F0, 0xFF00};
    uint32_t tt[64];
    for (int j = 0; j < 4; j++) tt[j] = (ev[pp[j]] ^ (((isC >> j) & 1) ? 0xFFFFu : 0u)) & 0xFFFF;
    for (int j = 4; j <= rt; j++) {
        uint32_t wd = lib_word(L, cls, j - 4);
        uint32_t a = tt[wd & 63] ^ (((wd >> 12) & 1) ? 0xFFFFu : 0u);
        uint32_t b = tt[(wd >> 6) & 63] ^ (((wd >> 13) & 1) ? 0xFFFFu : 0u);
        tt[j] = a & b & 0xFFFF;
    }
    return tt[rt];
}

struct PlanCore {
    int cls, rt, compl_, valid;
    uint32_t isC;
    uint64_t addMask;
};

// Recomputes the chosen replacement of node `id` exactly as evaluated.
__device__ PlanCore plan_core(const GraphView &g, const Lib &L, int id, const uint32_t *p, int out, Mffc &t,
                              Col match, Col mlev) {
    int leaves[4], nL;
    uint32_t tt;
    cut_leaves4(p, leaves, nL, tt);
    calc_mffc(g, id, leaves, nL, t, match);
    PlanCore pc;
    uint64_t matched;
    pc.cls = match_cut(g, L, id, tt, leaves, t, match, mlev, pc.isC, matched);
    pc.rt = lib_root(L, pc.cls, out);
    pc.addMask = added_set(L, pc.cls, pc.rt, matched);
    pc.valid = 1;
    uint32_t w = L.npn[tt];
    uint32_t rtt = lib_truth(L, pc.cls, pc.rt, L.perms4 + 4 * ((w >> 5) & 31), pc.isC);
    if (rtt == tt) pc.compl_ = 0;
    else if (rtt == ((~tt) & 0xFFFF)) pc.compl_ = 1;
    else { pc.compl_ = 0; pc.valid = 0; }
    return pc;
}

__global__ void __launch_bounds__(EVAL_BLOCK) k_plan(int32_t nCand, const uint32_t *__restrict__ candIds,
                                                      GraphView g, Lib L, const uint32_t *__restrict__ cuts,
                                                      const int4 *__restrict__ best, Plan *plans) {
    __shared__ int32_t s_match[NMATCH * EVAL_BLOCK], s_mlev[NMATCH * EVAL_BLOCK];
    int64_t c = blockIdx.x * (int64_t)EVAL_BLOCK + threadIdx.x;
    if (c >= nCand) return;
    Col match{s_match + threadIdx.x}, mlev{s_mlev + threadIdx.x};
    const int id = (int)candIds[c];
    const int4 b = best[id];
    const uint32_t *p = cuts + ((size_t)id * NCUT + b.x) * 5;
    Mffc t;
    PlanCore pc = plan_core(g, L, id, p, b.y, t, match, mlev);
    Plan pl;
    pl.id = id;
    pl.gain = b.z;
    pl.valid = pc.valid;
    pl.nNew = __popcll(pc.addMask);
    pl.pad = 0;
    pl.nW = 0;
    pl.W[pl.nW++] = id;
    for (int i = 0; i < MFFC_TAB; i++)
        if (i < t.size && t.num[i] == 0) pl.W[pl.nW++] = t.id[i];
    uint32_t meta = p[4];
    pl.nL = (meta >> 27) & 7;
    for (int i = 0; i < 4; i++) pl.L[i] = i < pl.nL ? (int32_t)p[i] : 0;
    uint64_t used = 1ull << pc.rt;
    for (int j = pc.rt; j >= 4; j--) {
        if (!((used >> j) & 1)) continue;
        uint32_t wd = lib_word(L, pc.cls, j - 4);
        used |= 1ull << (wd & 63);
        used |= 1ull << ((wd >> 6) & 63);
    }
    pl.nM = 0;
    for (int j = 4; j <= pc.rt; j++) {
        if (!((used >> j) & 1) || match[j] <= -1) continue;
        int m = match[j];
        int dup = 0;
        for (int q = 0; q < PLAN_MAXM; q++)
            if (q < pl.nM && pl.M[q] == m) dup = 1;
        if (m == id) pl.valid = 0;
        if (!dup) {
            if (pl.nM < PLAN_MAXM) pl.M[pl.nM++] = m;
            else pl.valid = 0;
        }
    }
    plans[c] = pl;
}

enum : uint32_t { MK_W = 1, MK_WNR = 2, MK_L = 4, MK_M = 8 };
enum : uint8_t { ST_ACTIVE = 0, ST_SELECTED = 1, ST_REJECTED = 2 };

__device__ __forceinline__ uint64_t stamp(uint32_t round, const Plan &p) {
    return (static_cast<uint64_t>(round) << 48) | (static_cast<uint64_t>(p.gain & 0xFFFF) << 32) |
           (0xFFFFFFFFu - static_cast<uint32_t>(p.id));
}

__global__ void k_sel_init(int32_t nCand, const Plan *__restrict__ plans, uint8_t *state) {
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c < nCand) state[c] = plans[c].valid ? ST_ACTIVE : ST_REJECTED;
}

struct Claims {
    unsigned long long *W, *Wnr, *L, *M;
};

__global__ void k_sel_claim(uint32_t round, int32_t nCand, const Plan *__restrict__ plans,
                            const uint8_t *__restrict__ state, Claims cl) {
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= nCand || state[c] != ST_ACTIVE) return;
    const Plan &p = plans[c];
    unsigned long long s = stamp(round, p);
    for (int i = 0; i < p.nW; i++) {
        atomicMax(&cl.W[p.W[i]], s);
        if (p.W[i] != p.id) atomicMax(&cl.Wnr[p.W[i]], s);
    }
    for (int i = 0; i < p.nL; i++) atomicMax(&cl.L[p.L[i]], s);
    for (int i = 0; i < p.nM; i++) atomicMax(&cl.M[p.M[i]], s);
}

__global__ void k_sel_decide(uint32_t round, int32_t nCand, const Plan *__restrict__ plans, uint8_t *state,
                             Claims cl, uint32_t *mk) {
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= nCand || state[c] != ST_ACTIVE) return;
    const Plan &p = plans[c];
    unsigned long long s = stamp(round, p);
    bool win = true;
    for (int i = 0; i < p.nW && win; i++) {
        int x = p.W[i];
        if (cl.W[x] != s || cl.M[x] >= s) win = false;
        if (x != p.id && cl.L[x] >= s) win = false;
    }
    for (int i = 0; i < p.nL && win; i++)
        if (cl.Wnr[p.L[i]] >= s) win = false;
    for (int i = 0; i < p.nM && win; i++)
        if (cl.W[p.M[i]] >= s) win = false;
    if (!win) return;
    state[c] = ST_SELECTED;
    for (int i = 0; i < p.nW; i++) atomicOr(&mk[p.W[i]], MK_W | (p.W[i] != p.id ? MK_WNR : 0u));
    for (int i = 0; i < p.nL; i++) atomicOr(&mk[p.L[i]], MK_L);
    for (int i = 0; i < p.nM; i++) atomicOr(&mk[p.M[i]], MK_M);
}

__global__ void k_sel_elim(int32_t nCand, const Plan *__restrict__ plans, uint8_t *state,
                           const uint32_t *__restrict__ mk, uint32_t *activeLeft) {
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= nCand || state[c] != ST_ACTIVE) return;
    const Plan &p = plans[c];
    bool conf = false;
    for (int i = 0; i < p.nW && !conf; i++) {
        uint32_t m = mk[p.W[i]];
        if (m & (MK_W | MK_M)) conf = true;
        if (p.W[i] != p.id && (m & MK_L)) conf = true;
    }
    for (int i = 0; i < p.nL && !conf; i++)
        if (mk[p.L[i]] & MK_WNR) conf = true;
    for (int i = 0; i < p.nM && !conf; i++)
        if (mk[p.M[i]] & MK_W) conf = true;
    if (conf) state[c] = ST_REJECTED;
    else atomicAdd(activeLeft, 1u);
}

__global__ void k_new_counts(int32_t nCand, const Plan *__restrict__ plans, const uint8_t *__restrict__ state,
                             uint32_t *cnt, unsigned long long *gainSum) {
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= nCand) return;
    bool s = state[c] == ST_SELECTED;
    cnt[c] = s ? (uint32_t)plans[c].nNew : 0u;
    if (s) atomicAdd(gainSum, (unsigned long long)plans[c].gain);
}

__global__ void k_iota2(int32_t n, uint32_t *repl) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < n; v += stride) repl[v] = 2u * (uint32_t)v;
}

// Creates the new nodes of each selected plan at ids n + off[c] + k, in
// ascending library index order, and records the replacement literal.
__global__ void __launch_bounds__(EVAL_BLOCK) k_build(int32_t n, int32_t nCand, const uint32_t *__restrict__ candIds,
                                                       GraphView g, Lib L, const uint32_t *__restrict__ cuts,
                                                       const int4 *__restrict__ best, const uint8_t *__restrict__ state,
                                                       const uint32_t *__restrict__ off, uint32_t *cf0, uint32_t *cf1,
                                                       uint32_t *repl, uint32_t *err, uint32_t *ctag) {
    __shared__ int32_t s_match[NMATCH * EVAL_BLOCK], s_mlev[NMATCH * EVAL_BLOCK];
    int64_t c = blockIdx.x * (int64_t)EVAL_BLOCK + threadIdx.x;
    if (c >= nCand || state[c] != ST_SELECTED) return;
    Col match{s_match + threadIdx.x}, mlev{s_mlev + threadIdx.x};
    const int id = (int)candIds[c];
    const int4 b = best[id];
    const uint32_t *p = cuts + ((size_t)id * NCUT + b.x) * 5;
    Mffc t;
    PlanCore pc = plan_core(g, L, id, p, b.y, t, match, mlev);
    uint32_t next = (uint32_t)n + off[c];
    for (int j = 4; j <= pc.rt; j++) {
        if (!((pc.addMask >> j) & 1)) continue;
        if (ctag) ctag[next] = ctag[id];   // a new node carries its root's tag
        uint32_t wd = lib_word(L, pc.cls, j - 4);
        int a = wd & 63, bb = (wd >> 6) & 63;
        if (match[a] < 0 || match[bb] < 0) {  // cannot happen for a consistent plan
            atomicOr(err, 2u);
            cf0[next] = 0;
            cf1[next] = 0;
            match[j] = (int32_t)next;
            next++;
            continue;
        }
        uint32_t la = 2u * (uint32_t)match[a] + (isc_bit(pc.isC, a) ^ ((wd >> 12) & 1u));
        uint32_t lb = 2u * (uint32_t)match[bb] + (isc_bit(pc.isC, bb) ^ ((wd >> 13) & 1u));
        cf0[next] = la;
        cf1[next] = lb;
        match[j] = (int32_t)next;
        next++;
    }
    repl[id] = 2u * (uint32_t)match[pc.rt] + (uint32_t)pc.compl_;
}

__global__ void k_compose(int32_t n, int32_t N, int32_t nPIs, const uint32_t *__restrict__ f0,
                          const uint32_t *__restrict__ f1, const uint32_t *__restrict__ repl, uint32_t *cf0,
                          uint32_t *cf1) {
    int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = tid; v < N; v += stride) {
        if (v <= nPIs) { cf0[v] = 0; cf1[v] = 0; continue; }
        uint32_t a = v < n ? f0[v] : cf0[v], b = v < n ? f1[v] : cf1[v];
        uint32_t ia = LID(a), ib = LID(b);
        cf0[v] = (ia < (uint32_t)n ? repl[ia] : 2u * ia) ^ LC(a);
        cf1[v] = (ib < (uint32_t)n ? repl[ib] : 2u * ib) ^ LC(b);
    }
}

__global__ void k_compose_pos(int32_t nPOs, const uint32_t *__restrict__ pos, const uint32_t *__restrict__ repl,
                              uint32_t *cpos) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < nPOs) cpos[i] = repl[LID(pos[i])] ^ LC(pos[i]);
}

__global__ void k_flag_cand(int32_t n, const int4 *__restrict__ best, uint8_t *flag) {
    int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v < n) flag[v] = best[v].x >= 0;
}

// reorder: new position of every node, fanins remapped and sorted
__global__ void k_reorder_ids(int32_t m, const uint32_t *__restrict__ sortedIds, uint32_t *newId) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p == 0) newId[0] = 0;
    if (p < m) newId[sortedIds[p]] = (uint32_t)(p + 1);
}

__global__ void k_reorder_emit(int32_t n, int32_t nPIs, const uint32_t *__restrict__ f0,
                               const uint32_t *__restrict__ f1, const uint32_t *__restrict__ level,
                               const uint32_t *__restrict__ newId, uint32_t *of0, uint32_t *of1, uint32_t *olevel,
                               const uint32_t *__restrict__ tagIn, uint32_t *tagOut) {
    int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t nv = newId[v];
    olevel[nv] = level[v];
    if (tagOut) tagOut[nv] = tagIn[v];
    if (v <= nPIs) { of0[nv] = 0; of1[nv] = 0; return; }
    uint32_t a = 2u * newId[LID(f0[v])] | LC(f0[v]);
    uint32_t b = 2u * newId[LID(f1[v])] | LC(f1[v]);
    if (a > b) { uint32_t x = a; a = b; b = x; }
    of0[nv] = a;
    of1[nv] = b;
}

__global__ void k_reorder_pos(int32_t nPOs, const uint32_t *__restrict__ pos, const uint32_t *__restrict__ newId,
                              uint32_t *opos) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < nPOs) opos[i] = 2u * newId[LID(pos[i])] | LC(pos[i]);
}

__global__ void k_level_starts(int32_t m, const uint32_t *__restrict__ keys, int32_t *starts) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= m) return;
    if (p == 0 || keys[p] != keys[p - 1]) starts[keys[p]] = (int32_t)(p + 1);
}

__global__ void k_iota1(int32_t m, uint32_t *x) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) x[i] = (uint32_t)(i + 1);
}

__global__ void k_gather_level(int32_t m, const uint32_t *__restrict__ level, uint32_t *keys) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) keys[i] = level[i + 1];
}

struct Tmp {
    gls_ctx *ctx;
    std::vector<void *> ptrs;
    bool ok = true;
    template <class T> T *get(size_t n) {
        T *p = pool_get<T>(ctx, n);
        if (p) ptrs.push_back(p);
        else ok = false;
        return p;
    }
    ~Tmp() {
        for (void *p : ptrs) pool_put(ctx, p);
    }
};

// Owns a pool-backed graph until it is freed explicitly or the scope ends.
struct GraphGuard {
    gls_ctx *ctx;
    DevGraph *g;
    void free_now() {
        if (g) dev_graph_free(ctx, *g);
        g = nullptr;
    }
    ~GraphGuard() { free_now(); }
};

int blocks(int64_t n, int b) { return (int)((n + b - 1) / b > 0 ? (n + b - 1) / b : 1); }

uint32_t bits_for(uint32_t x) {
    uint32_t b = 1;
    while (b < 32 && (x >> b) != 0) ++b;
    return b;
}

Lib lib_of(gls_ctx *ctx) {
    return Lib{ctx->d_npn, ctx->d_perms4, ctx->d_lib_nodes, ctx->d_lib_node_off, ctx->d_lib_root_off, ctx->d_lib_roots};
}

// Stage timer of a pass (diagnostics only): an event call that fails makes
// every stage time read 0 instead of failing the pass.
struct Timer {
    gls_ctx *ctx;
    int next = 0;
    bool ok = true;
    void mark() {
        if (next < 16) {
            if (hipEventRecord(ctx->ev[next], ctx->stream) != hipSuccess) ok = false;
            ++next;
        }
    }
    double ms(int a, int b) {
        float x = 0;
        if (!ok || b >= next || hipEventElapsedTime(&x, ctx->ev[a], ctx->ev[b]) != hipSuccess) return 0.0;
        return x;
    }
};

} // namespace

// Signature bits of a batched pass: every node's id in its own AIG, & 31.
// Nodes 1..n-1 sorted stably by tag keep their id order inside each tag; in
// the reordered numbering an AIG's PIs come first, then its ANDs level-major
// (ties by id), exactly as its own reordered numbering, so the rank inside the
// tag group + 1 is that id.
__global__ void __launch_bounds__(kBlock) k_sig_keys(int32_t n, const uint32_t *__restrict__ tag, uint32_t *keys,
                                                     uint32_t *vals, uint32_t *hist) {
    __shared__ uint32_t s_bins[TAG_BINS];
    TagBins tb{s_bins};
    tb.init();
    int64_t b, e;
    tag_chunk(1, n, b, e);
    for (int64_t u = b + threadIdx.x; u < e; u += blockDim.x) {
        const uint32_t k = tag[u];
        keys[u - 1] = k;
        vals[u - 1] = (uint32_t)u;
        tb.add(hist, k, 1u);
    }
    tb.flush_add(hist, 0xFFFFFFFFu);
}

__global__ void k_sig_bits(int32_t m, const uint32_t *__restrict__ skeys, const uint32_t *__restrict__ svals,
                           const uint32_t *__restrict__ start, uint8_t *sigBits) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < m; q += stride)
        sigBits[svals[q]] = (uint8_t)((1u + (uint32_t)(q - start[skeys[q]])) & 31u);
    if (blockIdx.x == 0 && threadIdx.x == 0) sigBits[0] = 0;
}

static int sig_bits_for_tags(gls_ctx *ctx, int32_t n, const uint32_t *tag, uint8_t *sigBits) {
    hipStream_t st = ctx->stream;
    const int32_t m = n - 1;
    if (m <= 0) return GLS_OK;
    Tmp t{ctx};
    uint32_t *mx = t.get<uint32_t>(1);
    size_t tb = 0;
    GLS_HIP(ctx, rocprim::reduce(nullptr, tb, tag, mx, 0u, (size_t)n, rocprim::maximum<uint32_t>(), st));
    void *tmp = t.get<uint8_t>(tb);
    if (!mx || !tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "signature bits: allocation");
    GLS_HIP(ctx, rocprim::reduce(tmp, tb, tag, mx, 0u, (size_t)n, rocprim::maximum<uint32_t>(), st));
    uint32_t maxTag = 0;
    GLS_TRY(stream_copy(ctx, &maxTag, mx, sizeof(uint32_t), hipMemcpyDeviceToHost));
    uint32_t *keys = t.get<uint32_t>(m), *vals = t.get<uint32_t>(m), *skeys = t.get<uint32_t>(m);
    uint32_t *svals = t.get<uint32_t>(m), *hist = t.get<uint32_t>(maxTag + 2), *start = t.get<uint32_t>(maxTag + 2);
    if (!t.ok) return gls_fail(ctx, GLS_ERR_NOMEM, "signature bits: allocation");
    GLS_HIP(ctx, hipMemsetAsync(hist, 0, sizeof(uint32_t) * (maxTag + 2), st));
    k_sig_keys<<<tag_grid(m), kBlock, 0, st>>>(n, tag, keys, vals, hist);
    size_t tb2 = 0;
    const uint32_t eb = bits_for(maxTag + 1);
    GLS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb2, keys, skeys, vals, svals, (size_t)m, 0u, eb, st));
    size_t tb3 = 0;
    GLS_HIP(ctx, rocprim::exclusive_scan(nullptr, tb3, hist, start, 0u, (size_t)maxTag + 1, rocprim::plus<uint32_t>(), st));
    void *tmp2 = t.get<uint8_t>(std::max(tb2, tb3));
    if (!tmp2) return gls_fail(ctx, GLS_ERR_NOMEM, "signature bits: sort storage");
    GLS_HIP(ctx, rocprim::radix_sort_pairs(tmp2, tb2, keys, skeys, vals, svals, (size_t)m, 0u, eb, st));
    GLS_HIP(ctx, rocprim::exclusive_scan(tmp2, tb3, hist, start, 0u, (size_t)maxTag + 1, rocprim::plus<uint32_t>(), st));
    k_sig_bits<<<grid_for(m), kBlock, 0, st>>>(m, skeys, svals, start, sigBits);
    GLS_HIP(ctx, hipGetLastError());
    GLS_HIP(ctx, hipStreamSynchronize(st));
    return GLS_OK;
}

int dev_reorder(gls_ctx *ctx, const DevGraph &in, DevGraph &out, std::vector<int32_t> &levelStart,
                const uint32_t *tagIn, uint32_t **tagOut) {
    const int32_t n = in.nObjs, m = n - 1;
    hipStream_t st = ctx->stream;
    Tmp t{ctx};
    uint32_t *keys = t.get<uint32_t>(m + 1), *keys2 = t.get<uint32_t>(m + 1), *ids = t.get<uint32_t>(m + 1);
    uint32_t *ids2 = t.get<uint32_t>(m + 1), *newId = t.get<uint32_t>(n);
    if (!t.ok) return gls_fail(ctx, GLS_ERR_NOMEM, "reorder: allocation");
    uint32_t maxLevel = 0;
    GLS_TRY(dev_graph_alloc(ctx, out, n, in.nPIs, in.nPOs));
    GraphGuard outGuard{ctx, &out};   // released to the caller on success
    uint32_t *tags = nullptr;
    if (tagIn && tagOut) {
        tags = pool_get<uint32_t>(ctx, (size_t)n);
        if (!tags) return gls_fail(ctx, GLS_ERR_NOMEM, "reorder: tags");
        *tagOut = tags;
    }
    if (m > 0) {
        k_gather_level<<<blocks(m, kBlock), kBlock, 0, st>>>(m, in.level, keys);
        k_iota1<<<blocks(m, kBlock), kBlock, 0, st>>>(m, ids);
        size_t tb = 0;
        GLS_HIP(ctx, rocprim::reduce(nullptr, tb, keys, keys2, 0u, (size_t)m, rocprim::maximum<uint32_t>(), st));
        void *tmp = t.get<uint8_t>(tb);
        if (!tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "reorder: reduce storage");
        GLS_HIP(ctx, rocprim::reduce(tmp, tb, keys, keys2, 0u, (size_t)m, rocprim::maximum<uint32_t>(), st));
        GLS_HIP(ctx, hipMemcpyAsync(&maxLevel, keys2, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        GLS_HIP(ctx, hipStreamSynchronize(st));
        size_t tb2 = 0;
        uint32_t eb = bits_for(maxLevel);
        GLS_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb2, keys, keys2, ids, ids2, (size_t)m, 0u, eb, st));
        void *tmp2 = t.get<uint8_t>(tb2);
        if (!tmp2) return gls_fail(ctx, GLS_ERR_NOMEM, "reorder: sort storage");
        GLS_HIP(ctx, rocprim::radix_sort_pairs(tmp2, tb2, keys, keys2, ids, ids2, (size_t)m, 0u, eb, st));
        k_reorder_ids<<<blocks(m, kBlock), kBlock, 0, st>>>(m, ids2, newId);
        int32_t *d_starts = t.get<int32_t>(maxLevel + 2);
        if (!d_starts) return gls_fail(ctx, GLS_ERR_NOMEM, "reorder: starts");
        GLS_HIP(ctx, hipMemsetAsync(d_starts, 0, sizeof(int32_t) * (maxLevel + 2), st));
        k_level_starts<<<blocks(m, kBlock), kBlock, 0, st>>>(m, keys2, d_starts);
        levelStart.assign(maxLevel + 2, 0);
        GLS_HIP(ctx, hipMemcpyAsync(levelStart.data(), d_starts, sizeof(int32_t) * (maxLevel + 1), hipMemcpyDeviceToHost, st));
        k_reorder_emit<<<blocks(n, kBlock), kBlock, 0, st>>>(n, in.nPIs, in.f0, in.f1, in.level, newId, out.f0,
                                                              out.f1, out.level, tagIn, tags);
        if (in.nPOs > 0)
            k_reorder_pos<<<blocks(in.nPOs, kBlock), kBlock, 0, st>>>(in.nPOs, in.pos, newId, out.pos);
        GLS_HIP(ctx, hipGetLastError());
        GLS_HIP(ctx, hipStreamSynchronize(st));
        levelStart[maxLevel + 1] = n;  // one past the last id
    } else {
        GLS_HIP(ctx, hipMemsetAsync(out.f0, 0, sizeof(uint32_t) * n, st));
        GLS_HIP(ctx, hipMemsetAsync(out.f1, 0, sizeof(uint32_t) * n, st));
        GLS_HIP(ctx, hipMemsetAsync(out.level, 0, sizeof(uint32_t) * n, st));
        if (in.nPOs > 0)
            GLS_HIP(ctx, hipMemcpyAsync(out.pos, in.pos, sizeof(uint32_t) * in.nPOs, hipMemcpyDeviceToDevice, st));
        if (tags) GLS_HIP(ctx, hipMemcpyAsync(tags, tagIn, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
        GLS_HIP(ctx, hipStreamSynchronize(st));
        levelStart.assign(2, n);
    }
    outGuard.g = nullptr;
    return GLS_OK;
}

// per-level cut launches are timed in windows (see dev_enumerate_cuts)
// (GLS_KERNEL_SAMPLE=0 in the environment turns this off, e.g. under a
// profiler that collects counters per dispatch)
static constexpr int KSAMPLE_EVERY = 64;
static constexpr int KSAMPLE_WIN = 8;
static const bool kSampleLaunches = [] {
    const char *e = getenv("GLS_KERNEL_SAMPLE");
    return !(e && e[0] == '0');
}();

// Task lists of the planned schedule for nL launches of K levels each (lv:
// per launch the first ids of its K levels, then its end; K + 1 entries):
// heads / ents on the device, toff[j] = first task of launch j on the host
// (nL + 1 entries).  lv must stay alive until this returns (its upload is
// complete by then).
static int build_cut_plan(gls_ctx *ctx, Tmp &t, const DevGraph &g, int32_t K, const std::vector<int32_t> &lv,
                          uint4 *&heads, uint4 *&ents, std::vector<uint32_t> &toff) {
    hipStream_t st = ctx->stream;
    const int32_t n = g.nObjs, nL = (int32_t)(lv.size() / (size_t)(K + 1));
    int32_t *dlv = t.get<int32_t>(lv.size());
    int32_t *nodeLaunch = t.get<int32_t>((size_t)n);
    uint32_t *own = t.get<uint32_t>((size_t)n);
    uint8_t *coneSize = t.get<uint8_t>((size_t)n);
    unsigned long long *flag = t.get<unsigned long long>((size_t)n + 1);
    unsigned long long *S = t.get<unsigned long long>((size_t)n + 1);
    uint32_t *cnt = t.get<uint32_t>((size_t)nL + 1), *dOff = t.get<uint32_t>((size_t)nL + 1);
    if (!t.ok) return gls_fail(ctx, GLS_ERR_NOMEM, "cuts: plan storage");
    GLS_HIP(ctx, hipMemcpyAsync(dlv, lv.data(), sizeof(int32_t) * lv.size(), hipMemcpyHostToDevice, st));
    GLS_HIP(ctx, hipMemsetAsync(nodeLaunch, 0xFF, sizeof(int32_t) * (size_t)n, st));
    GLS_HIP(ctx, hipMemsetAsync(own, 0xFF, sizeof(uint32_t) * (size_t)n, st));
    k_plan_fill<<<nL, kBlock, 0, st>>>(K, dlv, nodeLaunch);
    for (int p = K - 1; p >= 1; p--)
        k_plan_roots<<<grid_for(n), kBlock, 0, st>>>(n, g.nPIs, K, p, g.f0, g.f1, dlv, nodeLaunch, own, coneSize);
    k_plan_flags<<<grid_for((int64_t)n + 1), kBlock, 0, st>>>(n, K, dlv, nodeLaunch, own, flag);
    GLS_HIP(ctx, hipGetLastError());
    size_t tb = 0, tb2 = 0;
    GLS_HIP(ctx, rocprim::exclusive_scan(nullptr, tb, flag, S, 0ull, (size_t)n + 1,
                                         rocprim::plus<unsigned long long>(), st));
    GLS_HIP(ctx, rocprim::exclusive_scan(nullptr, tb2, cnt, dOff, 0u, (size_t)nL + 1, rocprim::plus<uint32_t>(), st));
    void *tmp = t.get<uint8_t>(std::max<size_t>(std::max(tb, tb2), 1));
    if (!tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "cuts: plan scan storage");
    GLS_HIP(ctx, rocprim::exclusive_scan(tmp, tb, flag, S, 0ull, (size_t)n + 1, rocprim::plus<unsigned long long>(),
                                         st));
    k_plan_counts<<<blocks(nL + 1, kBlock), kBlock, 0, st>>>(nL, K, dlv, S, cnt);
    GLS_HIP(ctx, rocprim::exclusive_scan(tmp, tb2, cnt, dOff, 0u, (size_t)nL + 1, rocprim::plus<uint32_t>(), st));
    toff.assign((size_t)nL + 1, 0u);
    GLS_TRY(stream_copy(ctx, toff.data(), dOff, sizeof(uint32_t) * ((size_t)nL + 1), hipMemcpyDeviceToHost));
    const size_t nT = toff[nL];
    uint32_t *ecnt = t.get<uint32_t>(nT + 1), *eoff = t.get<uint32_t>(nT + 1);
    heads = t.get<uint4>(std::max<size_t>(nT, 1));
    size_t tb3 = 0;
    GLS_HIP(ctx, rocprim::exclusive_scan(nullptr, tb3, ecnt, eoff, 0u, nT + 1, rocprim::plus<uint32_t>(), st));
    void *tmp3 = t.get<uint8_t>(std::max<size_t>(tb3, 1));
    if (!ecnt || !eoff || !heads || !tmp3) return gls_fail(ctx, GLS_ERR_NOMEM, "cuts: plan tasks (%zu)", nT);
    GLS_HIP(ctx, hipMemsetAsync(ecnt, 0, sizeof(uint32_t) * (nT + 1), st));
    k_plan_ecount<<<grid_for(n), kBlock, 0, st>>>(n, g.nPIs, K, dlv, nodeLaunch, own, S, dOff, coneSize, ecnt);
    GLS_HIP(ctx, rocprim::exclusive_scan(tmp3, tb3, ecnt, eoff, 0u, nT + 1, rocprim::plus<uint32_t>(), st));
    // entries: every node once as a root or packed node, plus the cone
    // members recomputed by other roots; bounded by the cone sizes' sum
    uint32_t nE = 0;
    GLS_TRY(stream_copy(ctx, &nE, eoff + nT, sizeof(uint32_t), hipMemcpyDeviceToHost));
    ents = t.get<uint4>(std::max<size_t>(nE, 1));
    if (!ents) return gls_fail(ctx, GLS_ERR_NOMEM, "cuts: plan entries (%u)", nE);
    k_plan_scatter<<<grid_for(n), kBlock, 0, st>>>(n, g.nPIs, K, g.f0, g.f1, dlv, nodeLaunch, own, S, dOff, eoff,
                                                   heads, ents);
    GLS_HIP(ctx, hipGetLastError());
    return GLS_OK;
}

int dev_enumerate_cuts(gls_ctx *ctx, const DevGraph &g, const std::vector<int32_t> &levelStart, uint32_t *d_cuts,
                       int32_t *d_nref, int mode, const uint8_t *sigBits, int *ranMode) {
    int ranDummy = 0;
    int &ran = ranMode ? *ranMode : ranDummy;
    ran = 0;
    hipStream_t st = ctx->stream;
    const int32_t n = g.nObjs;
    ctx->ksamp_n = 0;
    ctx->ksamp_units.clear();
    ctx->ksamp_launches.clear();
    ctx->cut_graph_ms = 0;
    ctx->cut_graph_nodes = 0;
    // The planned schedule is switched off: its kernel made an illegal memory
    // access on the GPU in the schedule tests, also after its LDS records were
    // written through LDS pointers with a wait before each barrier, and the
    // cause is not found (the task lists pass every bounds check in a host
    // build of the plan kernels).  It is refused before anything is launched.
    if (mode & GLS_RW_PLAN_CUTS)
        return gls_fail(ctx, GLS_ERR_ARG, "GLS_RW_PLAN_CUTS is disabled (its kernel faulted the GPU; cause not found)");
    const bool graphed = (mode & GLS_RW_GRAPH_CUTS) != 0;
    mode &= ~GLS_RW_GRAPH_CUTS;
    GLS_HIP(ctx, hipMemsetAsync(d_nref, 0, sizeof(int32_t) * n, st));
    GLS_HIP(ctx, hipMemsetAsync(d_cuts, 0, sizeof(uint32_t) * (size_t)n * NCUT * 5, st));
    k_nref<<<grid_for(n), kBlock, 0, st>>>(n, g.nPIs, g.f0, g.f1, g.nPOs, g.pos, d_nref);
    if (g.nPIs > 0) k_cuts_pi<<<blocks(g.nPIs, kBlock), kBlock, 0, st>>>(g.nPIs, d_nref, d_cuts);
    // levelStart[l] = first id of level l (ids 1.. are level-major); the last
    // entry is one past the last id.
    const int L = (int)levelStart.size() - 1;
    if (L <= 1) return GLS_OK;
    if (mode == GLS_RW_DATAFLOW_CUTS && L > 16) {
        int perCU = 0, dev = 0, nCU = 0;
        GLS_HIP(ctx, hipGetDevice(&dev));
        GLS_HIP(ctx, hipDeviceGetAttribute(&nCU, hipDeviceAttributeMultiprocessorCount, dev));
        GLS_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_cuts_dataflow, CUTW_BLOCK, 0));
        const int G = std::max(1, std::min(perCU, 2) * nCU);
        const size_t nslots = (size_t)n * NCUT;
        unsigned long long *tcuts = pool_get<unsigned long long>(ctx, nslots * 3);
        uint32_t *ctr = pool_get<uint32_t>(ctx, 4);
        if (!tcuts || !ctr) { pool_put(ctx, tcuts); pool_put(ctx, ctr); return gls_fail(ctx, GLS_ERR_NOMEM, "cuts: tagged records"); }
        GLS_HIP(ctx, hipMemsetAsync(tcuts, 0, sizeof(unsigned long long) * nslots * 3, st));
        GLS_HIP(ctx, hipMemsetAsync(ctr, 0, sizeof(uint32_t) * 4, st));
        if (g.nPIs > 0)
            k_cuts_pi_tagged<<<blocks((int64_t)g.nPIs * NCUT, kBlock), kBlock, 0, st>>>(g.nPIs, d_nref, tcuts);
        k_cuts_dataflow<<<G, CUTW_BLOCK, 0, st>>>(g.nPIs + 1, n, g.f0, g.f1, d_nref, tcuts, &ctr[0], &ctr[1],
                                                   sigBits);
        k_untag_cuts<<<grid_for((int64_t)nslots), kBlock, 0, st>>>((int64_t)nslots, tcuts, d_cuts);
        GLS_HIP(ctx, hipGetLastError());
        uint32_t hctr[2] = {0, 0};
        int rc = stream_copy(ctx, hctr, ctr, sizeof(hctr), hipMemcpyDeviceToHost);
        pool_put(ctx, tcuts);
        pool_put(ctx, ctr);
        if (rc != GLS_OK) return rc;
        if (hctr[1]) return gls_fail(ctx, GLS_ERR_INTERNAL, "cut enumeration: a wait exceeded its bound");
        // every node must have been taken (a wavefront stops once tickets pass the end)
        if (hctr[0] < (uint32_t)(n - g.nPIs - 1))
            return gls_fail(ctx, GLS_ERR_INTERNAL, "cut enumeration: %u of %d nodes processed", hctr[0],
                            n - g.nPIs - 1);
        ran = GLS_RW_DATAFLOW_CUTS;
        return GLS_OK;
    }
    if (mode == GLS_RW_XCD_CUTS && L > 16) {
        int dev = 0, nCU = 0;
        GLS_HIP(ctx, hipGetDevice(&dev));
        GLS_HIP(ctx, hipDeviceGetAttribute(&nCU, hipDeviceAttributeMultiprocessorCount, dev));
        const uint32_t nwg = (uint32_t)(nCU / XCD_COUNT);
        int32_t *d_starts = pool_get<int32_t>(ctx, levelStart.size());
        uint32_t *ctl = pool_get<uint32_t>(ctx, 4);
        if (!d_starts || !ctl) { pool_put(ctx, d_starts); pool_put(ctx, ctl); return gls_fail(ctx, GLS_ERR_NOMEM, "cuts: xcd control"); }
        GLS_HIP(ctx, hipMemcpyAsync(d_starts, levelStart.data(), sizeof(int32_t) * levelStart.size(),
                                    hipMemcpyHostToDevice, st));
        GLS_HIP(ctx, hipMemsetAsync(ctl, 0, sizeof(uint32_t) * 4, st));
        const dim3 grid((unsigned)(nwg * XCD_COUNT));
        if (kSampleLaunches) {
            const size_t k = 2 * (size_t)ctx->ksamp_n;
            while (ctx->kev.size() < k + 2) {
                hipEvent_t ev = nullptr;
                GLS_HIP(ctx, hipEventCreate(&ev));
                ctx->kev.push_back(ev);
            }
            hipExtLaunchKernelGGL(k_cuts_xcd, grid, dim3(XCD_WG), 0, st, ctx->kev[k], ctx->kev[k + 1], 0, 1, L,
                                  (const int32_t *)d_starts, (const uint32_t *)g.f0, (const uint32_t *)g.f1,
                                  (const int32_t *)d_nref, d_cuts, sigBits, nwg, ctl);
            ctx->ksamp_units.push_back(levelStart[L] - levelStart[1]);
            ctx->ksamp_launches.push_back(1);
            ctx->ksamp_n++;
        } else {
            k_cuts_xcd<<<grid, XCD_WG, 0, st>>>(1, L, d_starts, g.f0, g.f1, d_nref, d_cuts, sigBits, nwg, ctl);
        }
        GLS_HIP(ctx, hipGetLastError());
        uint32_t hctl[3] = {0, 0, 0};
        const int rc = stream_copy(ctx, hctl, ctl, sizeof(hctl), hipMemcpyDeviceToHost);
        pool_put(ctx, d_starts);
        pool_put(ctx, ctl);
        if (rc != GLS_OK) return rc;
        if (hctl[2] == 0) {
            ran = GLS_RW_XCD_CUTS;
            return GLS_OK;
        }
        // not all on one XCD, or a wait overran: redo with one launch per level
        ctx->ksamp_n = 0;
        ctx->ksamp_units.clear();
        ctx->ksamp_launches.clear();
        GLS_HIP(ctx, hipMemsetAsync(d_cuts, 0, sizeof(uint32_t) * (size_t)n * NCUT * 5, st));
        if (g.nPIs > 0) k_cuts_pi<<<blocks(g.nPIs, kBlock), kBlock, 0, st>>>(g.nPIs, d_nref, d_cuts);
        mode = 0;
    }
    bool coop = mode == GLS_RW_COOP_CUTS && L > 16;
    int G = 0;
    if (coop) {
        int perCU = 0, dev = 0, nCU = 0, coopOk = 0;
        GLS_HIP(ctx, hipGetDevice(&dev));
        GLS_HIP(ctx, hipDeviceGetAttribute(&nCU, hipDeviceAttributeMultiprocessorCount, dev));
        GLS_HIP(ctx, hipDeviceGetAttribute(&coopOk, hipDeviceAttributeCooperativeLaunch, dev));
        GLS_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_cuts_levels, CUTW_BLOCK, 0));
        G = std::min(perCU * nCU, 64);
        coop = coopOk && G > 0;
    }
    if (coop) {
        int32_t *d_starts = pool_get<int32_t>(ctx, levelStart.size());
        if (!d_starts) return gls_fail(ctx, GLS_ERR_NOMEM, "cuts: level starts");
        GLS_HIP(ctx, hipMemcpyAsync(d_starts, levelStart.data(), sizeof(int32_t) * levelStart.size(),
                                    hipMemcpyHostToDevice, st));
        int32_t l0 = 1, l1 = L;
        const uint32_t *a0 = g.f0, *a1 = g.f1;
        const int32_t *ar = d_nref;
        uint32_t *ac = d_cuts;
        const uint8_t *sb = sigBits;
        void *args[] = {&l0, &l1, &d_starts, &a0, &a1, &ar, &ac, &sb};
        hipError_t e = hipLaunchCooperativeKernel((const void *)k_cuts_levels, dim3(G), dim3(CUTW_BLOCK), args, 0, st);
        // the pool block is reused only by later work on the same stream
        pool_put(ctx, d_starts);
        if (e != hipSuccess) return gls_fail(ctx, GLS_ERR_HIP, "cooperative cut launch: %s", hipGetErrorString(e));
        ran = GLS_RW_COOP_CUTS;
    } else {
        // default: three levels per launch (k_cuts_triple); GLS_RW_PAIR_CUTS:
        // two (k_cuts_pair); GLS_RW_LEVEL_CUTS: one (k_cuts_level).  A group
        // with a level wider than cut_wide nodes fills the GPU by itself: the
        // recomputed cones would only add work there, so its levels get one
        // launch each.
        // entries hold node ids in 30 bits
        const bool planned = mode == GLS_RW_PLAN_CUTS && n < (1 << 30);
        const int step = mode == GLS_RW_LEVEL_CUTS ? 1 : mode == GLS_RW_PAIR_CUTS ? 2 : mode == GLS_RW_QUAD_CUTS ? 4
                         : planned ? std::min(std::max(tuning_for(ctx, n).cut_levels, 1), PLAN_MAX_LEVELS) : 3;
        const int32_t cutWide = tuning_for(ctx, n).cut_wide;
        ran = planned ? GLS_RW_PLAN_CUTS : step == 1 ? GLS_RW_LEVEL_CUTS : step == 2 ? GLS_RW_PAIR_CUTS
              : step == 4 ? GLS_RW_QUAD_CUTS : 0;
        struct CutLaunch { int kind; int32_t s, m, m2, e, l; };
        std::vector<CutLaunch> plan;
        plan.reserve((size_t)L / step + 1);
        for (int l = 1; l < L; l += step) {
            const int32_t s = levelStart[l], m = levelStart[std::min(l + 1, L)];
            const int32_t m2 = levelStart[std::min(l + 2, L)];
            const int32_t e = levelStart[std::min(l + step, L)];
            if (e <= s) continue;
            int32_t widest = 0;
            for (int j = l; j < std::min(l + step, L); j++) widest = std::max(widest, levelStart[j + 1] - levelStart[j]);
            if (step > 1 && widest > cutWide) {
                for (int j = l; j < std::min(l + step, L); j++)
                    if (levelStart[j + 1] > levelStart[j])
                        plan.push_back({1, levelStart[j], levelStart[j + 1], levelStart[j + 1], levelStart[j + 1], j});
            } else {
                plan.push_back({step, s, m, m2, e, l});
            }
        }
        // sampled timing of the schedule's own kernel: a window of up to
        // KSAMPLE_WIN consecutive launches of it starts every KSAMPLE_EVERY
        // launches; its first dispatch carries the start event and its last the
        // stop event, so the elapsed time over the launches is the per-launch
        // kernel time a profiler reports, with the events' own cost spread over
        // the window
        // GLS_RW_GRAPH_CUTS: the same launches captured into a HIP graph (kernel
        // nodes only), instantiated and launched once; the host time of the
        // capture and the instantiation is reported apart
        // planned schedule: the task lists of the three-level launches, built
        // before any launch (and before a capture begins)
        Tmp pt{ctx};
        std::vector<int32_t> lv;
        std::vector<uint32_t> toff;
        uint4 *heads = nullptr, *ents = nullptr;
        if (planned && step > 1) {
            for (const CutLaunch &c : plan)
                if (c.kind == step)
                    for (int p = 0; p <= step; p++) lv.push_back(levelStart[std::min(c.l + p, L)]);
            if (!lv.empty()) GLS_TRY(build_cut_plan(ctx, pt, g, step, lv, heads, ents, toff));
        }
        size_t j3 = 0;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        const auto tg0 = std::chrono::steady_clock::now();
        if (graphed) GLS_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int win = 0;
        for (size_t i = 0; i < plan.size(); i++) {
            const CutLaunch &c = plan[i];
            const bool pl = planned && step > 1 && c.kind == step;
            const uint32_t tbase = pl ? toff[j3] : 0u, tcount = pl ? toff[j3 + 1] - toff[j3] : 0u;
            if (pl) ++j3;
            const dim3 grid(pl ? tcount : c.kind > 1 ? (unsigned)(c.e - c.s) : blocks(c.e - c.s, CUTL_BLOCK / 64));
            hipEvent_t evStart = nullptr, evStop = nullptr;
            if (kSampleLaunches && !graphed && win == 0 && c.kind == step && i % KSAMPLE_EVERY == 0) {
                const size_t k = 2 * (size_t)ctx->ksamp_n;
                while (ctx->kev.size() < k + 2) {
                    hipEvent_t ev = nullptr;
                    GLS_HIP(ctx, hipEventCreate(&ev));
                    ctx->kev.push_back(ev);
                }
                evStart = ctx->kev[k];
                ctx->ksamp_units.push_back(0);
                ctx->ksamp_launches.push_back(0);
                ctx->ksamp_n++;
                win = KSAMPLE_WIN;
            }
            if (win > 0) {
                ctx->ksamp_units.back() += c.e - c.s;
                ctx->ksamp_launches.back() += 1;
                // the window ends at its last launch or before a launch of another kernel
                if (--win == 0 || i + 1 == plan.size() || plan[i + 1].kind != step) {
                    win = 0;
                    evStop = ctx->kev[2 * (size_t)ctx->ksamp_n - 1];
                }
                if (pl)
                    hipExtLaunchKernelGGL(k_cuts_plan, grid, dim3(CUT3_BLOCK), 0, st, evStart, evStop, 0,
                                          (const uint4 *)heads, (const uint4 *)ents, (int32_t)tbase, (int32_t)step,
                                          (const int32_t *)d_nref, d_cuts, sigBits);
                else if (c.kind == 4)
                    hipExtLaunchKernelGGL(k_cuts_quad, grid, dim3(CUT4_BLOCK), 0, st, evStart, evStop, 0, c.s, c.m,
                                          c.m2, levelStart[std::min(c.l + 3, L)], (const uint32_t *)g.f0,
                                          (const uint32_t *)g.f1, (const int32_t *)d_nref, d_cuts, sigBits);
                else if (c.kind == 3)
                    hipExtLaunchKernelGGL(k_cuts_triple, grid, dim3(CUT3_BLOCK), 0, st, evStart, evStop, 0, c.s, c.m,
                                          c.m2, (const uint32_t *)g.f0, (const uint32_t *)g.f1,
                                          (const int32_t *)d_nref, d_cuts, sigBits);
                else if (c.kind == 2)
                    hipExtLaunchKernelGGL(k_cuts_pair, grid, dim3(CUTP_BLOCK), 0, st, evStart, evStop, 0, c.s,
                                          c.m - c.s, c.m, (const uint32_t *)g.f0, (const uint32_t *)g.f1,
                                          (const int32_t *)d_nref, d_cuts, sigBits);
                else
                    hipExtLaunchKernelGGL(k_cuts_level, grid, dim3(CUTL_BLOCK), 0, st, evStart, evStop, 0, c.s,
                                          c.e - c.s, (const uint32_t *)g.f0, (const uint32_t *)g.f1,
                                          (const int32_t *)d_nref, d_cuts, sigBits);
            } else if (pl) {
                k_cuts_plan<<<grid, CUT3_BLOCK, 0, st>>>(heads, ents, (int32_t)tbase, step, d_nref, d_cuts, sigBits);
            } else if (c.kind == 4) {
                k_cuts_quad<<<grid, CUT4_BLOCK, 0, st>>>(c.s, c.m, c.m2, levelStart[std::min(c.l + 3, L)], g.f0, g.f1,
                                                         d_nref, d_cuts, sigBits);
            } else if (c.kind == 3) {
                k_cuts_triple<<<grid, CUT3_BLOCK, 0, st>>>(c.s, c.m, c.m2, g.f0, g.f1, d_nref, d_cuts, sigBits);
            } else if (c.kind == 2) {
                k_cuts_pair<<<grid, CUTP_BLOCK, 0, st>>>(c.s, c.m - c.s, c.m, g.f0, g.f1, d_nref, d_cuts, sigBits);
            } else {
                k_cuts_level<<<grid, CUTL_BLOCK, 0, st>>>(c.s, c.e - c.s, g.f0, g.f1, d_nref, d_cuts, sigBits);
            }
        }
        if (graphed) {
            const hipError_t ec = hipStreamEndCapture(st, &graph);
            if (ec != hipSuccess) return gls_fail(ctx, GLS_ERR_HIP, "cut graph capture: %s", hipGetErrorString(ec));
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                (void)hipGraphDestroy(graph);
                return gls_fail(ctx, GLS_ERR_HIP, "cut graph instantiation: %s", hipGetErrorString(ei));
            }
            ctx->cut_graph_ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tg0).count();
            ctx->cut_graph_nodes = (int64_t)plan.size();
            hipError_t el = hipGraphLaunch(exec, st);
            // the executable graph is released once its launch has run
            if (el == hipSuccess) el = hipStreamSynchronize(st);
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            if (el != hipSuccess) return gls_fail(ctx, GLS_ERR_HIP, "cut graph launch: %s", hipGetErrorString(el));
            ran |= GLS_RW_GRAPH_CUTS;
        }
    }
    GLS_HIP(ctx, hipGetLastError());
    return GLS_OK;
}

static int build_table(gls_ctx *ctx, Tmp &t, const DevGraph &g, GraphView &gv) {
    hipStream_t st = ctx->stream;
    const int32_t n = g.nObjs;
    const size_t nAnd = n > g.nPIs + 1 ? (size_t)(n - g.nPIs - 1) : 0;
    uint32_t *cnt = t.get<uint32_t>((size_t)n + 1), *off = t.get<uint32_t>((size_t)n + 1);
    uint2 *ent = t.get<uint2>(std::max<size_t>(nAnd, 1));
    if (!cnt || !off || !ent) return gls_fail(ctx, GLS_ERR_NOMEM, "table allocation");
    GLS_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(uint32_t) * ((size_t)n + 1), st));
    k_table_count<<<grid_for(n), kBlock, 0, st>>>(n, g.nPIs, g.f1, cnt);
    size_t tb = 0;
    GLS_HIP(ctx, rocprim::exclusive_scan(nullptr, tb, cnt, off, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(), st));
    void *tmp = t.get<uint8_t>(std::max<size_t>(tb, 1));
    if (!tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "table scan storage");
    GLS_HIP(ctx, rocprim::exclusive_scan(tmp, tb, cnt, off, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(), st));
    GLS_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(uint32_t) * ((size_t)n + 1), st));
    k_table_fill<<<grid_for(n), kBlock, 0, st>>>(n, g.nPIs, g.f0, g.f1, off, cnt, ent);
    GLS_HIP(ctx, hipGetLastError());
    gv.tOff = off;
    gv.tEnt = ent;
    return GLS_OK;
}

int dev_evaluate(gls_ctx *ctx, const DevGraph &g, const uint32_t *d_cuts, const int32_t *d_nref, int useZeros,
                 int32_t *d_best) {
    GLS_TRY(ctx_ensure_library(ctx));
    hipStream_t st = ctx->stream;
    Tmp t{ctx};
    GraphView gv{g.f0, g.f1, d_nref, g.level, nullptr, nullptr};
    GLS_TRY(build_table(ctx, t, g, gv));
    GLS_HIP(ctx, hipMemsetAsync(d_best, 0xFF, sizeof(int4) * 1, st));
    int32_t nAnd = g.nObjs - g.nPIs - 1;
    // PIs and the constant have no candidate
    GLS_HIP(ctx, hipMemsetAsync(d_best, 0, sizeof(int4) * (size_t)(g.nPIs + 1), st));
    std::vector<int4> none(g.nPIs + 1, make_int4(-1, 0, 0, 0));
    GLS_HIP(ctx, hipMemcpyAsync(d_best, none.data(), sizeof(int4) * none.size(), hipMemcpyHostToDevice, st));
    if (nAnd > 0)
        k_eval<int32_t><<<blocks(nAnd, EVAL_BLOCK), EVAL_BLOCK, 0, st>>>(g.nObjs, g.nPIs, gv, lib_of(ctx), d_cuts, useZeros,
                                                                reinterpret_cast<int4 *>(d_best));
    GLS_HIP(ctx, hipGetLastError());
    GLS_HIP(ctx, hipStreamSynchronize(st));
    return GLS_OK;
}

int dev_rewrite(gls_ctx *ctx, const DevGraph &in, int useZeros, int flags, DevGraph &out, int32_t *depth,
                gls_op_info *info, const uint32_t *tagIn, uint32_t **tagOut, uint32_t *tagDepth) {
    GLS_TRY(ctx_ensure_library(ctx));
    hipStream_t st = ctx->stream;
    Timer tm{ctx};
    tm.mark(); // 0
    DevGraph g;
    std::vector<int32_t> levelStart;
    const bool tagged = tagIn && tagOut;
    uint32_t *tagR = nullptr;   // tags in the reordered numbering (freed on every return)
    struct TagGuard {
        gls_ctx *c;
        uint32_t **p;
        ~TagGuard() { if (*p) pool_put(c, *p); }
    } tagGuard{ctx, &tagR};
    GLS_TRY(dev_reorder(ctx, in, g, levelStart, tagged ? tagIn : nullptr, tagged ? &tagR : nullptr));
    GraphGuard gGuard{ctx, &g};   // frees the reordered graph on every return path
    tm.mark(); // 1
    const int32_t n = g.nObjs, nPIs = g.nPIs, nPOs = g.nPOs;
    int rc = GLS_OK;
    {
        Tmp t{ctx};
        uint32_t *cuts = t.get<uint32_t>((size_t)n * NCUT * 5);
        int32_t *nref = t.get<int32_t>(n);
        int4 *best = t.get<int4>(n);
        uint8_t *flag = t.get<uint8_t>(n);
        uint32_t *candIds = t.get<uint32_t>(n), *d_cnt = t.get<uint32_t>(4), *d_err = t.get<uint32_t>(4);
        uint32_t *repl = t.get<uint32_t>(n);
        unsigned long long *d_gain = t.get<unsigned long long>(1);
        if (!t.ok) return gls_fail(ctx, GLS_ERR_NOMEM, "rewrite: allocation (n=%d)", n);
        GLS_HIP(ctx, hipMemsetAsync(d_err, 0, 16, st));
        uint8_t *sigBits = nullptr;
        if (tagged) {
            sigBits = t.get<uint8_t>(n);
            if (!sigBits) return gls_fail(ctx, GLS_ERR_NOMEM, "rewrite: signature bits");
            GLS_TRY(sig_bits_for_tags(ctx, n, tagR, sigBits));
        }
        int cutRan = 0;
        GLS_TRY(dev_enumerate_cuts(ctx, g, levelStart, cuts, nref,
                                   flags & (GLS_RW_COOP_CUTS | GLS_RW_DATAFLOW_CUTS | GLS_RW_XCD_CUTS |
                                            GLS_RW_LEVEL_CUTS | GLS_RW_PAIR_CUTS | GLS_RW_GRAPH_CUTS |
                                            GLS_RW_PLAN_CUTS | GLS_RW_QUAD_CUTS),
                                   sigBits, &cutRan));
        tm.mark(); // 2
        GraphView gv{g.f0, g.f1, nref, g.level, nullptr, nullptr};
        GLS_TRY(build_table(ctx, t, g, gv));
        tm.mark(); // 3
        std::vector<int4> none(nPIs + 1, make_int4(-1, 0, 0, 0));
        GLS_HIP(ctx, hipMemcpyAsync(best, none.data(), sizeof(int4) * none.size(), hipMemcpyHostToDevice, st));
        const int32_t nAnd = n - nPIs - 1;
        const Lib L = lib_of(ctx);
        // a structure node's level exceeds the graph's deepest by at most
        // NMATCH, so below 65000 levels every value fits 16 bits under 0xFFFF
        if (nAnd > 0 && levelStart.size() < 65000)
            k_eval<uint16_t><<<blocks(nAnd, EVAL_BLOCK), EVAL_BLOCK, 0, st>>>(n, nPIs, gv, L, cuts, useZeros, best);
        else if (nAnd > 0)
            k_eval<int32_t><<<blocks(nAnd, EVAL_BLOCK), EVAL_BLOCK, 0, st>>>(n, nPIs, gv, L, cuts, useZeros, best);
        GLS_HIP(ctx, hipGetLastError());
        tm.mark(); // 4
        // candidates in id order
        k_flag_cand<<<blocks(n, kBlock), kBlock, 0, st>>>(n, best, flag);
        size_t tb = 0;
        rocprim::counting_iterator<uint32_t> ids(0u);
        GLS_HIP(ctx, rocprim::select(nullptr, tb, ids, flag, candIds, d_cnt, (size_t)n, st));
        void *tmp = t.get<uint8_t>(tb);
        if (!tmp) return gls_fail(ctx, GLS_ERR_NOMEM, "rewrite: select storage");
        GLS_HIP(ctx, rocprim::select(tmp, tb, ids, flag, candIds, d_cnt, (size_t)n, st));
        uint32_t nCand = 0;
        GLS_HIP(ctx, hipMemcpyAsync(&nCand, d_cnt, 4, hipMemcpyDeviceToHost, st));
        GLS_HIP(ctx, hipStreamSynchronize(st));
        uint32_t herr = 0;
        uint32_t nSel = 0, rounds = 0, totalNew = 0;
        unsigned long long estGain = 0;
        k_iota2<<<grid_for(n), kBlock, 0, st>>>(n, repl);
        Plan *plans = nullptr;
        uint8_t *state = nullptr;
        uint32_t *newOff = nullptr;
        if (nCand > 0) {
            plans = t.get<Plan>(nCand);
            state = t.get<uint8_t>(nCand);
            uint32_t *mk = t.get<uint32_t>(n), *newCnt = t.get<uint32_t>(nCand + 1);
            newOff = t.get<uint32_t>(nCand + 1);
            Claims cl{t.get<unsigned long long>(n), t.get<unsigned long long>(n), t.get<unsigned long long>(n),
                      t.get<unsigned long long>(n)};
            if (!t.ok) return gls_fail(ctx, GLS_ERR_NOMEM, "rewrite: plan allocation");
            k_plan<<<blocks(nCand, EVAL_BLOCK), EVAL_BLOCK, 0, st>>>((int32_t)nCand, candIds, gv, L, cuts, best, plans);
            k_sel_init<<<blocks(nCand, kBlock), kBlock, 0, st>>>((int32_t)nCand, plans, state);
            GLS_HIP(ctx, hipMemsetAsync(mk, 0, sizeof(uint32_t) * n, st));
            GLS_HIP(ctx, hipMemsetAsync(cl.W, 0, sizeof(unsigned long long) * n, st));
            GLS_HIP(ctx, hipMemsetAsync(cl.Wnr, 0, sizeof(unsigned long long) * n, st));
            GLS_HIP(ctx, hipMemsetAsync(cl.L, 0, sizeof(unsigned long long) * n, st));
            GLS_HIP(ctx, hipMemsetAsync(cl.M, 0, sizeof(unsigned long long) * n, st));
            for (uint32_t r = 1;; r++) {
                GLS_HIP(ctx, hipMemsetAsync(d_cnt, 0, 4, st));
                k_sel_claim<<<blocks(nCand, kBlock), kBlock, 0, st>>>(r, (int32_t)nCand, plans, state, cl);
                k_sel_decide<<<blocks(nCand, kBlock), kBlock, 0, st>>>(r, (int32_t)nCand, plans, state, cl, mk);
                k_sel_elim<<<blocks(nCand, kBlock), kBlock, 0, st>>>((int32_t)nCand, plans, state, mk, d_cnt);
                GLS_HIP(ctx, hipGetLastError());
                uint32_t left = 0;
                GLS_HIP(ctx, hipMemcpyAsync(&left, d_cnt, 4, hipMemcpyDeviceToHost, st));
                GLS_HIP(ctx, hipStreamSynchronize(st));
                rounds = r;
                if (left == 0) break;
                if (r > 60000) return gls_fail(ctx, GLS_ERR_INTERNAL, "selection did not converge");
            }
            GLS_HIP(ctx, hipMemsetAsync(d_gain, 0, 8, st));
            k_new_counts<<<blocks(nCand, kBlock), kBlock, 0, st>>>((int32_t)nCand, plans, state, newCnt, d_gain);
            GLS_HIP(ctx, hipMemsetAsync(newCnt + nCand, 0, 4, st));
            size_t tb2 = 0;
            GLS_HIP(ctx, rocprim::exclusive_scan(nullptr, tb2, newCnt, newOff, 0u, (size_t)nCand + 1, rocprim::plus<uint32_t>(), st));
            void *tmp2 = t.get<uint8_t>(tb2);
            if (!tmp2) return gls_fail(ctx, GLS_ERR_NOMEM, "rewrite: scan storage");
            GLS_HIP(ctx, rocprim::exclusive_scan(tmp2, tb2, newCnt, newOff, 0u, (size_t)nCand + 1, rocprim::plus<uint32_t>(), st));
            GLS_HIP(ctx, hipMemcpyAsync(&totalNew, newOff + nCand, 4, hipMemcpyDeviceToHost, st));
            GLS_HIP(ctx, hipMemcpyAsync(&estGain, d_gain, 8, hipMemcpyDeviceToHost, st));
            // count selected (state == 1)
            std::vector<uint8_t> hs(nCand);
            GLS_HIP(ctx, hipMemcpyAsync(hs.data(), state, nCand, hipMemcpyDeviceToHost, st));
            GLS_HIP(ctx, hipStreamSynchronize(st));
            for (uint8_t s : hs) nSel += (s == ST_SELECTED);
        }
        tm.mark(); // 5
        const int32_t N = n + (int32_t)totalNew;
        DevGraph comb;
        GLS_TRY(dev_graph_alloc(ctx, comb, N, nPIs, nPOs));
        GraphGuard combGuard{ctx, &comb};
        uint32_t *tagC = nullptr;   // tags of the composed graph
        if (tagged) {
            tagC = t.get<uint32_t>(N);
            if (!tagC) return gls_fail(ctx, GLS_ERR_NOMEM, "rewrite: tags");
            GLS_HIP(ctx, hipMemcpyAsync(tagC, tagR, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
        }
        const bool fullStrash = (flags & GLS_RW_FULL_STRASH) != 0;
        if (nCand > 0 && totalNew + nSel > 0)
            k_build<<<blocks(nCand, EVAL_BLOCK), EVAL_BLOCK, 0, st>>>(n, (int32_t)nCand, candIds, gv, L, cuts, best,
                                                                       state, newOff, comb.f0, comb.f1, repl, d_err,
                                                                       tagC);
        k_compose<<<grid_for(N), kBlock, 0, st>>>(n, N, nPIs, g.f0, g.f1, repl, comb.f0, comb.f1);
        if (nPOs > 0) k_compose_pos<<<blocks(nPOs, kBlock), kBlock, 0, st>>>(nPOs, g.pos, repl, comb.pos);
        GLS_HIP(ctx, hipGetLastError());
        GLS_HIP(ctx, hipMemcpyAsync(&herr, d_err, 4, hipMemcpyDeviceToHost, st));
        GLS_HIP(ctx, hipStreamSynchronize(st));
        if (herr) return gls_fail(ctx, GLS_ERR_INTERNAL, "rewrite: inconsistent plan (%u)", herr);
        tm.mark(); // 6
        gls_op_info stInfo = {};
        if (fullStrash) {
            gGuard.free_now();
            rc = dev_strash(ctx, comb, out, depth, &stInfo, tagC, tagged ? tagOut : nullptr,
                            tagged ? tagDepth : nullptr);
        } else {
            // the incremental strash reads the input graph's fanins: g is
            // released only after it
            const RwDelta delta{n, (int32_t)levelStart.size() - 2, g.f0, g.f1, g.level, gv.tOff, gv.tEnt};
            rc = dev_strash_rw(ctx, comb, delta, out, depth, &stInfo, tagC, tagged ? tagOut : nullptr,
                               tagged ? tagDepth : nullptr);
            gGuard.free_now();
        }
        combGuard.free_now();
        tm.mark(); // 7
        GLS_HIP(ctx, hipStreamSynchronize(st));
        if (info) {
            info->reorder_ms = tm.ms(0, 1);
            info->cuts_ms = tm.ms(1, 2);
            info->table_ms = tm.ms(2, 3);
            info->eval_ms = tm.ms(3, 4);
            info->select_ms = tm.ms(4, 5);
            info->build_ms = tm.ms(5, 6);
            info->strash_ms = tm.ms(6, 7);
            info->total_ms = tm.ms(0, 7);
            info->candidates = nCand;
            info->selected = nSel;
            info->new_nodes = totalNew;
            info->rounds = rounds;
            info->levels_enumerated = (int64_t)levelStart.size() - 2;
            info->est_gain = (int64_t)estGain;
            info->hash_probes = stInfo.hash_probes;
            info->final_strash = stInfo.final_strash;
            info->cut_schedule = cutRan;
            info->strash_rounds = stInfo.strash_rounds;
            info->strash_sweeps = stInfo.strash_sweeps;
            info->cut_graph_ms = ctx->cut_graph_ms;
            info->cut_graph_nodes = ctx->cut_graph_nodes;
            double kms = 0;
            int64_t kunits = 0, klaunches = 0;
            for (int i = 0; i < ctx->ksamp_n; i++) {
                float x = 0;
                if (hipEventElapsedTime(&x, ctx->kev[2 * i], ctx->kev[2 * i + 1]) != hipSuccess) continue;
                kms += x;
                kunits += ctx->ksamp_units[i];
                klaunches += ctx->ksamp_launches[i];
            }
            info->kernel_sample_ms = kms;
            info->kernel_samples = klaunches;
            info->kernel_sample_nodes = kunits;
        }
    }
    return rc;
}
