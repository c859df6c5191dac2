/* CPU oracle for the rw/strash hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, sequentially and in plain C, the algorithms the HIP path
 * computes, so that tests can compare the GPU results against it bit for bit.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product never links against it.
 *
 * Reference files restated (paths relative to the CULS tree):
 *   strash -c         src/aig/strash.cu:322-466  (host strash of `st -c`)
 *   balance           src/algorithms/balance.cu:666-1382 (`b`, sortDecId = 1)
 *   strash            src/aig/strash.cu:175-314  (GPU strash: dangling removal,
 *                     level waves, trivial folding, re-id by tentative id),
 *                     src/hash_table/hash_table.h:422-449 (key, trivial rules)
 *   reorder           src/algorithms/rewrite.cu:1229-1279 (Init, Reorder)
 *   cuts / truth      src/algorithms/rewrite.cu:49-294
 *   evaluation        src/algorithms/rewrite.cu:296-447
 *   NPN tables        closed form of the tables read at rewrite.cu:1281-1307
 *   CULS replacement  src/algorithms/rewrite.cu:473-555 (ReplaceSubgr/BuildSubgr),
 *                     :716-770 (re-id), :811-938 (host choice pass, redundancy),
 *                     src/aig_manager.cu:432-564 (post-rw dedupe by strash)
 *
 * Deliberate differences from the reference, all shared with the HIP path:
 *   - strash keeps the smallest (wave, id) of duplicate nodes (insert-min) where
 *     the reference keeps whichever thread won an atomicCAS (hash_table.h:478-486);
 *   - orc_rw replaces the reference's sequential host choice pass with a greedy
 *     conflict-free selection by priority; orc_rw_culs keeps the reference's pass
 *     and is used only to compare quality of results.
 *
 * Parity pins (tests/): NPN tables vs sha256 of the reference tables; forest vs
 * sha256 of the reference class section; cut truth tables vs cone simulation;
 * cut sets vs brute-force enumeration on tiny graphs; strash vs simulation
 * equivalence + canonical-order invariants; MFFC vs brute-force dereferencing;
 * rw output vs simulation equivalence.
 */
#include "oracle.h"
#include "../culs_amd/csrc/rwlib_forest.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NCUT 8
#define MFFC_TABLE 8
#define ID(l) ((l) >> 1)
#define CMP(l) ((l) & 1)

/* ------------------------------------------------------------------ */
/* NPN tables: closed form of the reference's pPhases/pPerms/pMap.      */
/* ------------------------------------------------------------------ */
static const uint8_t PERMS4[24][4] = {
    {0,1,2,3},{1,0,2,3},{0,2,1,3},{2,0,1,3},{2,1,0,3},{1,2,0,3},
    {0,1,3,2},{1,0,3,2},{0,3,1,2},{3,0,1,2},{3,1,0,2},{1,3,0,2},
    {0,3,2,1},{3,0,2,1},{0,2,3,1},{2,0,3,1},{2,3,0,1},{3,2,0,1},
    {3,1,2,0},{1,3,2,0},{3,2,1,0},{2,3,1,0},{2,1,3,0},{1,2,3,0}};

static uint32_t tt_polarize(uint32_t t, int ph) {
    uint32_t r = 0;
    for (int m = 0; m < 16; m++)
        if ((t >> (m ^ ph)) & 1) r |= 1u << m;
    return r;
}

static uint32_t tt_permute(uint32_t t, const uint8_t *p) {
    uint32_t r = 0;
    for (int m = 0; m < 16; m++) {
        int mo = 0;
        for (int j = 0; j < 4; j++)           /* old variable j becomes new variable p[j] */
            if ((m >> p[j]) & 1) mo |= 1 << j;
        if ((t >> mo) & 1) r |= 1u << m;
    }
    return r;
}

void orc_npn_tables(uint8_t *phase, uint8_t *perm, uint8_t *cls, uint8_t *perms4) {
    uint8_t *vis = calloc(65536, 1);
    memset(phase, 0, 65536); memset(perm, 0, 65536); memset(cls, 0, 65536);
    vis[0] = vis[0xFFFF] = 1;
    phase[0xFFFF] = 16;
    int next = 1;
    for (uint32_t u = 1; u < 0xFFFF; u++) {
        if (vis[u]) continue;
        int c = next++;
        for (int i = 0; i < 16; i++)
            for (int k = 0; k < 24; k++) {
                uint32_t v = tt_permute(tt_polarize(u, i), PERMS4[k]);
                if (!vis[v]) { vis[v] = 1; phase[v] = (uint8_t)i; perm[v] = (uint8_t)k; cls[v] = (uint8_t)c; }
                uint32_t nv = (~v) & 0xFFFF;
                if (!vis[nv]) { vis[nv] = 1; phase[nv] = (uint8_t)(i | 16); perm[nv] = (uint8_t)k; cls[nv] = (uint8_t)c; }
            }
    }
    for (int k = 0; k < 24; k++)
        for (int j = 0; j < 4; j++) perms4[k * 4 + j] = PERMS4[k][j];
    free(vis);
}

static uint8_t g_phase[65536], g_perm[65536], g_cls[65536], g_perms4[96];
static int g_npn_ready = 0;
static void npn_init(void) {
    if (!g_npn_ready) { orc_npn_tables(g_phase, g_perm, g_cls, g_perms4); g_npn_ready = 1; }
}

/* ------------------------------------------------------------------ */
/* A small open-addressing map uint64 -> int32.                        */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t *k; int32_t *v; uint64_t mask; } Map;
#define MAP_EMPTY (~0ull)
static void map_init(Map *m, int64_t n) {
    uint64_t cap = 16;
    while (cap < (uint64_t)(2 * n + 16)) cap <<= 1;
    m->k = malloc(cap * sizeof(uint64_t));
    m->v = malloc(cap * sizeof(int32_t));
    memset(m->k, 0xFF, cap * sizeof(uint64_t));
    m->mask = cap - 1;
}
static void map_free(Map *m) { free(m->k); free(m->v); }
static uint64_t mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
    return x;
}
static int32_t map_get(const Map *m, uint64_t key) {
    uint64_t s = mix64(key) & m->mask;
    while (m->k[s] != MAP_EMPTY) {
        if (m->k[s] == key) return m->v[s];
        s = (s + 1) & m->mask;
    }
    return -1;
}
/* insert if absent; returns the stored value */
static int32_t map_put(Map *m, uint64_t key, int32_t val) {
    uint64_t s = mix64(key) & m->mask;
    while (m->k[s] != MAP_EMPTY) {
        if (m->k[s] == key) return m->v[s];
        s = (s + 1) & m->mask;
    }
    m->k[s] = key; m->v[s] = val;
    return val;
}
static uint64_t and_key(uint32_t l0, uint32_t l1) { return ((uint64_t)l0 << 32) | l1; }

/* ------------------------------------------------------------------ */
/* Levels of an AIG with arbitrary (acyclic) id order.                  */
/* live == NULL: all AND nodes.  Iterative post-order DFS.             */
/* ------------------------------------------------------------------ */
static int compute_levels(int n, int nPIs, const uint32_t *f0, const uint32_t *f1,
                          const uint8_t *live, int32_t *lev) {
    /* -1 unvisited, -2 on the DFS path; returns -1 if a cycle is found */
    for (int v = 0; v < n; v++) lev[v] = v <= nPIs ? 0 : -1;
    int32_t *stk = malloc(sizeof(int32_t) * (size_t)(n + 1));
    for (int r = nPIs + 1; r < n; r++) {
        if (lev[r] >= 0 || (live && !live[r])) continue;
        int sp = 0;
        stk[sp++] = r;
        while (sp) {
            int v = stk[sp - 1];
            if (lev[v] >= 0) { sp--; continue; }
            int a = ID(f0[v]), b = ID(f1[v]);
            if (lev[a] == -2 || lev[b] == -2) { free(stk); return -1; }
            if (lev[a] == -1) { lev[v] = -2; stk[sp++] = a; continue; }
            if (lev[b] == -1) { lev[v] = -2; stk[sp++] = b; continue; }
            lev[v] = 1 + (lev[a] > lev[b] ? lev[a] : lev[b]);
            sp--;
        }
    }
    free(stk);
    return 0;
}

int orc_depth(int n, int nPIs, const uint32_t *f0, const uint32_t *f1) {
    int32_t *lev = malloc(sizeof(int32_t) * (size_t)n);
    if (compute_levels(n, nPIs, f0, f1, NULL, lev)) { free(lev); return -1; }
    int d = 0;
    for (int v = nPIs + 1; v < n; v++) if (lev[v] > d) d = lev[v];
    free(lev);
    return d;
}

/* ------------------------------------------------------------------ */
/* Structural hashing (GPU semantics of strash.cu:190-314, made        */
/* deterministic: the representative of duplicates is the smallest    */
/* (wave, id); the output order is (wave, id)).                        */
/* ------------------------------------------------------------------ */
static uint32_t trivial_and(uint32_t a, uint32_t b, int *isTriv) {
    /* hash_table.h:435-449 in AIGER convention (0 = FALSE, 1 = TRUE) */
    *isTriv = 1;
    if (a == b) return a;
    if (a == (b ^ 1)) return 0;
    if (a == 1) return b;
    if (b == 1) return a;
    if (a == 0 || b == 0) return 0;
    *isTriv = 0;
    return 0;
}

int orc_strash(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
               const uint32_t *pos, uint32_t *of0, uint32_t *of1, uint32_t *opos, int *pDepth) {
    int32_t *ref = calloc((size_t)n, sizeof(int32_t));
    uint8_t *live = malloc((size_t)n);
    for (int v = nPIs + 1; v < n; v++) { ref[ID(f0[v])]++; ref[ID(f1[v])]++; }
    for (int i = 0; i < nPOs; i++) ref[ID(pos[i])]++;
    /* dangling removal: strash.cu:175-188, 221-230 */
    memset(live, 1, (size_t)n);
    int32_t *stk = malloc(sizeof(int32_t) * (size_t)(n + 1));
    int sp = 0;
    for (int v = nPIs + 1; v < n; v++) if (ref[v] == 0) stk[sp++] = v;
    while (sp) {
        int v = stk[--sp];
        live[v] = 0;
        int a = ID(f0[v]), b = ID(f1[v]);
        if (a > nPIs && --ref[a] == 0) stk[sp++] = a;
        if (b > nPIs && --ref[b] == 0) stk[sp++] = b;
    }
    free(stk);
    free(ref);
    /* waves = levels of the dangling-free graph (strash.cu:247-277) */
    int32_t *wave = malloc(sizeof(int32_t) * (size_t)n);
    if (compute_levels(n, nPIs, f0, f1, live, wave)) { free(wave); free(live); return -3; }
    int depth = 0;
    for (int v = nPIs + 1; v < n; v++) if (live[v] && wave[v] > depth) depth = wave[v];
    /* (wave, id) order via counting sort */
    int32_t *cnt = calloc((size_t)depth + 2, sizeof(int32_t));
    for (int v = nPIs + 1; v < n; v++) if (live[v]) cnt[wave[v] + 1]++;
    for (int w = 1; w <= depth + 1; w++) cnt[w] += cnt[w - 1];
    int nLive = cnt[depth + 1];
    int32_t *ord = malloc(sizeof(int32_t) * (size_t)(nLive + 1));
    for (int v = nPIs + 1; v < n; v++) if (live[v]) ord[cnt[wave[v]]++] = v;
    free(cnt);
    /* hashing in (wave, id) order: first occurrence = min (wave, id) */
    uint32_t *res = malloc(sizeof(uint32_t) * (size_t)n);   /* lit over representative ids */
    uint8_t *surv = calloc((size_t)n, 1);
    for (int v = 0; v <= nPIs && v < n; v++) res[v] = 2u * (uint32_t)v;
    Map m; map_init(&m, nLive);
    for (int k = 0; k < nLive; k++) {
        int v = ord[k];
        uint32_t a = res[ID(f0[v])] ^ CMP(f0[v]);
        uint32_t b = res[ID(f1[v])] ^ CMP(f1[v]);
        if (a > b) { uint32_t t = a; a = b; b = t; }
        int triv;
        uint32_t r = trivial_and(a, b, &triv);
        if (triv) { res[v] = r; continue; }
        int32_t rep = map_put(&m, and_key(a, b), v);
        res[v] = 2u * (uint32_t)rep;
        if (rep == v) surv[v] = 1;
    }
    map_free(&m);
    /* consecutive re-id in (wave, id) order (strash.cu:113-160) */
    int32_t *newId = malloc(sizeof(int32_t) * (size_t)n);
    for (int v = 0; v <= nPIs && v < n; v++) newId[v] = v;
    int next = nPIs + 1;
    for (int k = 0; k < nLive; k++) if (surv[ord[k]]) newId[ord[k]] = next++;
    for (int v = 0; v <= nPIs && v < n; v++) { of0[v] = 0; of1[v] = 0; }
    for (int k = 0; k < nLive; k++) {
        int v = ord[k];
        if (!surv[v]) continue;
        uint32_t a = res[ID(f0[v])] ^ CMP(f0[v]);
        uint32_t b = res[ID(f1[v])] ^ CMP(f1[v]);
        a = 2u * (uint32_t)newId[ID(a)] | CMP(a);
        b = 2u * (uint32_t)newId[ID(b)] | CMP(b);
        if (a > b) { uint32_t t = a; a = b; b = t; }
        of0[newId[v]] = a; of1[newId[v]] = b;
    }
    for (int i = 0; i < nPOs; i++) {
        uint32_t l = res[ID(pos[i])] ^ CMP(pos[i]);
        opos[i] = 2u * (uint32_t)newId[ID(l)] | CMP(l);
    }
    free(newId); free(res); free(surv); free(ord); free(wave); free(live);
    if (pDepth) *pDepth = depth;
    return next;
}

/* ------------------------------------------------------------------ */
/* Host strash of `st -c` (AIGMan::strash(fCPU=true), strash.cu:322-466)*/
/* ------------------------------------------------------------------ */
/* The live set is computed as the transitive fanin of the POs (the      */
/* reference removes zero-fanout nodes recursively, :360-383; both give  */
/* the same set).  Nodes are visited in id order; no trivial folding.    */
/* Fanins are ordered by the reference's own literal values, where lit 0 */
/* is TRUE (common.h:285-286), i.e. the AIGER constant literal flipped.  */
static uint32_t culs_val(uint32_t l) { return l < 2 ? 1u - l : l; }

int orc_strash_cpu(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
                   const uint32_t *pos, uint32_t *of0, uint32_t *of1, uint32_t *opos, uint32_t *olevel,
                   int *pDepth) {
    uint8_t *live = calloc((size_t)n, 1);
    for (int i = 0; i < nPOs; i++) live[ID(pos[i])] = 1;
    for (int v = n - 1; v > nPIs; v--)          /* ids are topological: one reverse sweep */
        if (live[v]) { live[ID(f0[v])] = 1; live[ID(f1[v])] = 1; }
    uint32_t *newId = malloc(sizeof(uint32_t) * (size_t)n);
    for (int v = 0; v <= nPIs && v < n; v++) { newId[v] = (uint32_t)v; of0[v] = of1[v] = 0; olevel[v] = 0; }
    Map m; map_init(&m, n);
    int next = nPIs + 1, depth = -1;
    for (int v = nPIs + 1; v < n; v++) {
        if (!live[v]) continue;
        uint32_t a = 2u * newId[ID(f0[v])] + CMP(f0[v]);
        uint32_t b = 2u * newId[ID(f1[v])] + CMP(f1[v]);
        if (culs_val(a) > culs_val(b)) { uint32_t t = a; a = b; b = t; }
        int32_t id = map_put(&m, and_key(a, b), next);
        newId[v] = (uint32_t)id;
        if (id == next) {
            of0[next] = a; of1[next] = b;
            uint32_t la = olevel[ID(a)], lb = olevel[ID(b)];
            olevel[next] = 1u + (la > lb ? la : lb);
            if ((int)olevel[next] > depth) depth = (int)olevel[next];
            next++;
        }
    }
    for (int i = 0; i < nPOs; i++) opos[i] = 2u * newId[ID(pos[i])] + CMP(pos[i]);
    map_free(&m); free(newId); free(live);
    if (pDepth) *pDepth = depth;
    return next;
}

/* ------------------------------------------------------------------ */
/* Balance `b` (balancePerformV2, src/algorithms/balance.cu:912-1382)   */
/* ------------------------------------------------------------------ */
/* Literals are AIGER here (0 = FALSE); where the reference's result     */
/* depends on literal VALUES (input order ties) its own values are used, */
/* in which lit 0 is TRUE (common.h:285-286): culs_val().                */
#define BAL_STACK 128            /* DFS_COVER_STACK_SIZE, balance.cu:18 */
#define BAL_MAX_COVER 1024       /* MAX_LOCAL_COVER_SIZE, balance.cu:17 */
#define BAL_BATCH 32768          /* iterLen, balance.cu:1129 */

typedef struct { int32_t *lit; int len; } BCover;   /* len -1: not computed */

/* AND with the trivial rules of checkTrivialAndCases (hash_table.h:435-449);
 * returns the literal, or -1 if non-trivial */
static int64_t bal_trivial(uint32_t a, uint32_t b) {
    if (a == b) return a;
    if (a == (b ^ 1u)) return 0;                 /* FALSE */
    if (a == 1u) return b;                       /* TRUE & b */
    if (b == 1u) return a;
    if (a == 0u || b == 0u) return 0;
    return -1;
}

/* localCoverTravToTable (balance.cu:666-750): DFS from the root, fanin0 first;
 * a literal is an input if complemented, not an AND, multi-fanout or at a
 * limit; a repeated input is dropped; an input and its complement make the
 * cover constant FALSE (len 0). */
static void bal_cover(int root, int nPIs, const uint32_t *f0, const uint32_t *f1, const int32_t *nFo, BCover *c) {
    uint32_t stack[BAL_STACK];
    int top = -1;
    int32_t *in = malloc(sizeof(int32_t) * (BAL_MAX_COVER + 64));   /* the limit is checked before expanding */
    int len = 0;
    const uint32_t rootLit = 2u * (uint32_t)root;
    stack[++top] = rootLit;
    while (top != -1) {
        uint32_t cur = stack[top--];
        int rep = 0;
        for (int i = 0; i < len; i++) {
            if ((uint32_t)in[i] == cur) { rep = 1; break; }
            if ((uint32_t)in[i] == (cur ^ 1u)) { c->lit = in; c->len = 0; return; }
        }
        if (rep) continue;
        int id = (int)ID(cur);
        if (cur != rootLit &&
            (CMP(cur) || id <= nPIs || nFo[id] > 1 || top + 2 >= BAL_STACK ||
             len + 2 * (top + 1) + 1 >= BAL_MAX_COVER)) {
            in[len++] = (int32_t)cur;
            continue;
        }
        stack[++top] = f1[id];
        stack[++top] = f0[id];
    }
    c->lit = in;
    c->len = len;
}

/* level of a literal in the rebuilt graph (PIs and constants 0) */
static uint32_t bal_level(const Map *lev, uint32_t lit, int nPIs) {
    int id = (int)ID(lit);
    if (id <= nPIs) return 0;
    int32_t l = map_get(lev, (uint64_t)id);
    return l < 0 ? 0u : (uint32_t)l;
}

/* AND lookup of retrieveAndNode (balance.cu:87-111): trivial result, or the
 * existing node's literal, or -1 */
static int64_t bal_lookup(const Map *tab, uint32_t a, uint32_t b) {
    int64_t t = bal_trivial(a, b);
    if (t >= 0) return t;
    if (a > b) { uint32_t x = a; a = b; b = x; }
    int32_t v = map_get(tab, and_key(a, b));
    return v < 0 ? -1 : 2 * (int64_t)v;
}

/* thrust::shuffle(thrust::seq, ..., default_random_engine(seed)) as a
 * permutation: position k of the shuffled sequence holds input element
 * perm[k].  The `b -s` tie-break shuffles each cover's inputs this way
 * (balance.cu:404-411).  The algorithm is thrust's own (a third-party
 * dependency of the reference, restated from rocThrust's
 * thrust/system/detail/generic/shuffle.inl as shipped with ROCm 7.2): a
 * 24-round Feistel bijection over the next power of two (at least 16), keys
 * drawn from minstd_rand (x <- 48271 x mod 2^31-1, seed 0 taken as 1), then
 * the indices below len kept in order. */
void orc_thrust_shuffle_perm(uint32_t len, uint32_t seed, uint32_t *perm) {
    uint32_t x = seed % 2147483647u;
    if (x == 0) x = 1;
    uint64_t bits = 4;
    if (len > 16) {
        bits = 0;
        for (uint64_t m = (uint64_t)len - 1; m; m >>= 1) bits++;
    }
    const uint64_t lb = bits / 2, rb = bits - lb;
    const uint64_t lmask = (1ull << lb) - 1, rmask = (1ull << rb) - 1;
    uint32_t key[24];
    for (int i = 0; i < 24; i++) {
        x = (uint32_t)(((uint64_t)x * 48271u) % 2147483647u);
        key[i] = x;
    }
    uint32_t k = 0;
    for (uint64_t j = 0; j < (1ull << bits); j++) {
        uint32_t s0 = (uint32_t)(j >> rb), s1 = (uint32_t)(j & rmask);
        for (int r = 0; r < 24; r++) {
            const uint64_t prod = 0xD2B74407B1CE6E93ull * (uint64_t)s0;
            const uint32_t hi = (uint32_t)(prod >> 32);
            uint32_t lo = (uint32_t)prod;
            lo = (lo << (rb - lb)) | (s1 >> lb);
            s0 = (uint32_t)(((hi ^ key[r]) ^ s1) & lmask);
            s1 = (uint32_t)(lo & rmask);
        }
        const uint64_t out = ((uint64_t)s0 << rb) | s1;
        if (out < len) perm[k++] = (uint32_t)out;
    }
}

int orc_balance(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs, const uint32_t *pos,
                int cap, uint32_t *of0, uint32_t *of1, uint32_t *opos) {
    return orc_balance_ex(n, nPIs, f0, f1, nPOs, pos, cap, of0, of1, opos, 1);
}

int orc_balance_ex(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs, const uint32_t *pos,
                   int cap, uint32_t *of0, uint32_t *of1, uint32_t *opos, int sortDecId) {
    int32_t *nFo = calloc((size_t)n, sizeof(int32_t));
    for (int v = nPIs + 1; v < n; v++) { nFo[ID(f0[v])]++; nFo[ID(f1[v])]++; }
    for (int i = 0; i < nPOs; i++) nFo[ID(pos[i])]++;
    BCover *cov = malloc(sizeof(BCover) * (size_t)n);
    for (int v = 0; v < n; v++) { cov[v].lit = NULL; cov[v].len = -1; }
    int32_t *lastLevel = malloc(sizeof(int32_t) * (size_t)n);
    for (int v = 0; v < n; v++) lastLevel[v] = -1;
    /* phase 1: cover roots level by level from the POs (balance.cu:958-1037) */
    int32_t *list = malloc(sizeof(int32_t) * (size_t)(n + nPOs + 1)), *next = malloc(sizeof(int32_t) * (size_t)(n + 1));
    uint8_t *mark = calloc((size_t)n, 1);
    int len = 0;
    for (int i = 0; i < nPOs; i++) {
        int id = (int)ID(pos[i]);
        if (id > nPIs && !mark[id]) { mark[id] = 1; list[len++] = id; }
    }
    int levels = 0;
    memset(mark, 0, (size_t)n);
    while (len > 0) {
        int nlen = 0;
        for (int k = 0; k < len; k++) {
            int v = list[k];
            lastLevel[v] = levels;
            if (cov[v].len < 0) bal_cover(v, nPIs, f0, f1, nFo, &cov[v]);
            for (int i = 0; i < cov[v].len; i++) {
                int id = (int)ID((uint32_t)cov[v].lit[i]);
                if (id > nPIs && !mark[id]) { mark[id] = 1; next[nlen++] = id; }
            }
        }
        for (int k = 0; k < nlen; k++) mark[next[k]] = 0;
        memcpy(list, next, sizeof(int32_t) * (size_t)nlen);   /* order within a level does not matter */
        len = nlen;
        levels++;
    }
    /* covers by (last level, id) */
    int nCov = 0;
    for (int v = nPIs + 1; v < n; v++) if (lastLevel[v] >= 0) nCov++;
    int32_t *byLevel = malloc(sizeof(int32_t) * (size_t)(nCov + 1));
    int32_t *lvStart = calloc((size_t)levels + 2, sizeof(int32_t));
    for (int v = nPIs + 1; v < n; v++) if (lastLevel[v] >= 0) lvStart[lastLevel[v] + 1]++;
    for (int l = 1; l <= levels + 1; l++) lvStart[l] += lvStart[l - 1];
    {
        int32_t *fill = malloc(sizeof(int32_t) * (size_t)(levels + 1));
        memcpy(fill, lvStart, sizeof(int32_t) * (size_t)(levels + 1));
        for (int v = nPIs + 1; v < n; v++) if (lastLevel[v] >= 0) byLevel[fill[lastLevel[v]]++] = v;
        free(fill);
    }
    /* phase 2: rebuild from the deepest level (balance.cu:1135-1229) */
    Map tab; map_init(&tab, 2 * (int64_t)n + 16);     /* (lit0, lit1) -> id */
    Map lev; map_init(&lev, 2 * (int64_t)n + 16);     /* id -> level */
    uint32_t *newLit = malloc(sizeof(uint32_t) * (size_t)n);
    int64_t counter = nPIs + 1;
    int32_t *arr = malloc(sizeof(int32_t) * (size_t)BAL_BATCH * (BAL_MAX_COVER + 64));
    uint32_t *alv = malloc(sizeof(uint32_t) * (size_t)BAL_BATCH * (BAL_MAX_COVER + 64));
    int32_t *alen = malloc(sizeof(int32_t) * BAL_BATCH);
    int8_t *mask = malloc(BAL_BATCH);
    uint32_t *sval = malloc(sizeof(uint32_t) * BAL_BATCH), *slev = malloc(sizeof(uint32_t) * BAL_BATCH);
    uint64_t *skey = malloc(sizeof(uint64_t) * BAL_BATCH);
    int64_t nCreated = 0;
    for (int l = levels - 1; l >= 0; l--) {
        for (int b0 = lvStart[l]; b0 < lvStart[l + 1]; b0 += BAL_BATCH) {
            int nb = lvStart[l + 1] - b0 < BAL_BATCH ? lvStart[l + 1] - b0 : BAL_BATCH;
            /* gather and sort (prepareReconstructArrays, balance.cu:272-427) */
            for (int c = 0; c < nb; c++) {
                int v = byLevel[b0 + c];
                int32_t *A = arr + (size_t)c * (BAL_MAX_COVER + 64);
                uint32_t *LV = alv + (size_t)c * (BAL_MAX_COVER + 64);
                alen[c] = cov[v].len;
                /* sortDecId = 0 (`b -s`): the inputs in the order thrust's
                 * shuffle seeded with the cover's index in its batch puts them
                 * (balance.cu:301, 404-411) */
                uint32_t perm[BAL_MAX_COVER + 64];
                if (!sortDecId && cov[v].len > 0) orc_thrust_shuffle_perm((uint32_t)cov[v].len, (uint32_t)c, perm);
                for (int i = 0; i < cov[v].len; i++) {
                    uint32_t lit = (uint32_t)cov[v].lit[sortDecId ? (uint32_t)i : perm[i]];
                    int id = (int)ID(lit);
                    uint32_t nl = id <= nPIs ? lit : (newLit[id] ^ CMP(lit));
                    A[i] = (int32_t)nl;
                    LV[i] = bal_level(&lev, nl, nPIs);
                }
                /* level descending, then (sortDecId = 1) the reference's
                 * literal value descending; insertion sort, so that equal
                 * keys keep their order as thrust's sequential sort does */
                for (int i = 1; i < alen[c]; i++) {
                    int32_t x = A[i]; uint32_t xl = LV[i];
                    int j = i - 1;
                    while (j >= 0 && (LV[j] < xl || (sortDecId && LV[j] == xl &&
                                                     culs_val((uint32_t)A[j]) < culs_val((uint32_t)x)))) {
                        A[j + 1] = A[j]; LV[j + 1] = LV[j]; j--;
                    }
                    A[j + 1] = x; LV[j + 1] = xl;
                }
            }
            for (;;) {
                /* sharedNodeDrivenPermute (balance.cu:187-247), against the table as it stands */
                for (int c = 0; c < nb; c++) {
                    int L = alen[c];
                    if (L <= 1) continue;
                    int32_t *A = arr + (size_t)c * (BAL_MAX_COVER + 64);
                    uint32_t *LV = alv + (size_t)c * (BAL_MAX_COVER + 64);
                    int left;
                    if (L < 3) left = 0;
                    else {
                        left = L - 2;
                        uint32_t border = LV[left];
                        for (left--; left >= 0; left--) if (LV[left] != border) break;
                        left++;
                    }
                    int right = L - 2;
                    if (left < right) {
                        uint32_t l1 = (uint32_t)A[right + 1], l2 = (uint32_t)A[right];
                        for (int i = right; i >= left; i--) {
                            uint32_t l3 = (uint32_t)A[i];
                            if (bal_lookup(&tab, l1, l3) >= 0) {
                                if (l3 != l2) {
                                    A[i] = (int32_t)l2; A[right] = (int32_t)l3;
                                    uint32_t t = LV[i]; LV[i] = LV[right]; LV[right] = t;
                                }
                                break;
                            }
                        }
                    }
                }
                /* prepareDataToInsert (balance.cu:113-185) */
                int any = 0;
                for (int c = 0; c < nb; c++) {
                    int L = alen[c];
                    int32_t *A = arr + (size_t)c * (BAL_MAX_COVER + 64);
                    uint32_t *LV = alv + (size_t)c * (BAL_MAX_COVER + 64);
                    if (L <= 1) { mask[c] = 0; continue; }
                    uint32_t l1 = (uint32_t)A[L - 1], l2 = (uint32_t)A[L - 2];
                    uint32_t v1 = LV[L - 1], v2 = LV[L - 2];
                    alen[c] = L - 2;
                    int64_t r = bal_lookup(&tab, l1, l2);
                    any = 1;
                    if (r >= 0) {
                        mask[c] = 2;
                        sval[c] = (uint32_t)r;
                        slev[c] = bal_level(&lev, (uint32_t)r, nPIs);
                        continue;
                    }
                    if (l1 > l2) { uint32_t t = l1; l1 = l2; l2 = t; }
                    mask[c] = 1;
                    skey[c] = and_key(l1, l2);
                    sval[c] = (uint32_t)(counter + c);
                    slev[c] = (v1 > v2 ? v1 : v2) + 1u;
                }
                if (!any) break;
                /* insert with the smallest id per key, then read back (balance.cu:1188-1210) */
                int64_t maxId = 0;
                for (int c = 0; c < nb; c++) {
                    if (mask[c] != 1) continue;
                    int32_t got = map_put(&tab, skey[c], (int32_t)sval[c]);   /* smaller c inserted first */
                    if (got == (int32_t)sval[c]) nCreated++;
                    sval[c] = (uint32_t)got;
                }
                for (int c = 0; c < nb; c++) if (mask[c] == 1 && (int64_t)sval[c] > maxId) maxId = sval[c];
                if (maxId != 0) counter = maxId + 1;
                for (int c = 0; c < nb; c++) if (mask[c] == 1) map_put(&lev, (uint64_t)sval[c], (int32_t)slev[c]);
                /* addBackLocalArrays (balance.cu:429-467) */
                for (int c = 0; c < nb; c++) {
                    if (mask[c] == 0) continue;
                    uint32_t lit = mask[c] == 2 ? sval[c] : 2u * sval[c];
                    int32_t *A = arr + (size_t)c * (BAL_MAX_COVER + 64);
                    uint32_t *LV = alv + (size_t)c * (BAL_MAX_COVER + 64);
                    int L = alen[c], dup = 0;
                    for (int i = 0; i < L; i++) if ((uint32_t)A[i] == lit) { dup = 1; break; }
                    if (dup) continue;
                    A[L] = (int32_t)lit; LV[L] = slev[c];
                    alen[c] = L + 1;
                    for (int i = L; i > 0; i--) {
                        if (LV[i] <= LV[i - 1]) break;
                        int32_t t = A[i]; A[i] = A[i - 1]; A[i - 1] = t;
                        uint32_t u = LV[i]; LV[i] = LV[i - 1]; LV[i - 1] = u;
                    }
                }
            }
            /* recordReconstructedCovers (balance.cu:469-491) */
            for (int c = 0; c < nb; c++) {
                int v = byLevel[b0 + c];
                newLit[v] = alen[c] == 0 ? 0u : (uint32_t)arr[(size_t)c * (BAL_MAX_COVER + 64)];
            }
        }
    }
    int ret;
    if (nCreated + nPIs + 1 > cap) ret = -2;
    else {
        /* re-id: created nodes in id order (balance.cu:1237-1262) */
        int64_t m = 0;
        uint64_t *keys = malloc(sizeof(uint64_t) * (size_t)(nCreated + 1));
        uint32_t *ids = malloc(sizeof(uint32_t) * (size_t)(nCreated + 1));
        for (uint64_t s = 0; s <= tab.mask; s++)
            if (tab.k[s] != MAP_EMPTY) { keys[m] = tab.k[s]; ids[m] = (uint32_t)tab.v[s]; m++; }
        /* sort by id (insertion into an id-indexed order via counting on a copy) */
        int64_t *ord = malloc(sizeof(int64_t) * (size_t)(m + 1));
        for (int64_t i = 0; i < m; i++) ord[i] = i;
        /* shell sort on ids: m is the number of created nodes */
        for (int64_t gap = m / 2; gap > 0; gap /= 2)
            for (int64_t i = gap; i < m; i++) {
                int64_t t = ord[i], j = i;
                while (j >= gap && ids[ord[j - gap]] > ids[t]) { ord[j] = ord[j - gap]; j -= gap; }
                ord[j] = t;
            }
        Map outId; map_init(&outId, m + 16);
        for (int64_t i = 0; i < m; i++) map_put(&outId, (uint64_t)ids[ord[i]], (int32_t)(nPIs + 1 + i));
        for (int v = 0; v <= nPIs; v++) { of0[v] = 0; of1[v] = 0; }
        for (int64_t i = 0; i < m; i++) {
            uint64_t k = keys[ord[i]];
            uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
            uint32_t ia = ID(a), ib = ID(b);
            uint32_t na = ia <= (uint32_t)nPIs ? ia : (uint32_t)map_get(&outId, ia);
            uint32_t nb2 = ib <= (uint32_t)nPIs ? ib : (uint32_t)map_get(&outId, ib);
            of0[nPIs + 1 + i] = 2u * na | CMP(a);
            of1[nPIs + 1 + i] = 2u * nb2 | CMP(b);
        }
        /* POs (processPO, balance.cu:882-910) */
        for (int i = 0; i < nPOs; i++) {
            uint32_t l = pos[i];
            int id = (int)ID(l);
            if (id <= nPIs) { opos[i] = l; continue; }
            uint32_t nl = newLit[id];
            uint32_t nid = ID(nl);
            uint32_t oid = nid <= (uint32_t)nPIs ? nid : (uint32_t)map_get(&outId, nid);
            opos[i] = 2u * oid | (CMP(l) ^ CMP(nl));
        }
        ret = (int)(nPIs + 1 + m);
        map_free(&outId); free(ord); free(keys); free(ids);
    }
    for (int v = 0; v < n; v++) free(cov[v].lit);
    free(cov); free(nFo); free(lastLevel); free(list); free(next); free(mark); free(byLevel); free(lvStart);
    map_free(&tab); map_free(&lev); free(newLit); free(arr); free(alv); free(alen); free(mask);
    free(sval); free(slev); free(skey);
    return ret;
}

/* ------------------------------------------------------------------ */
/* Reorder (rewrite.cu:1245-1279) + fanin sort of Init (:1229-1233).   */
/* ------------------------------------------------------------------ */
int orc_reorder(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
                const uint32_t *pos, uint32_t *of0, uint32_t *of1, uint32_t *opos, uint32_t *olevel) {
    int32_t *lev = malloc(sizeof(int32_t) * (size_t)n);
    if (compute_levels(n, nPIs, f0, f1, NULL, lev)) { free(lev); return -3; }
    int maxL = 0;
    for (int v = 1; v < n; v++) if (lev[v] > maxL) maxL = lev[v];
    int32_t *cnt = calloc((size_t)maxL + 1, sizeof(int32_t));
    for (int v = 1; v < n; v++) cnt[lev[v]]++;
    for (int l = 1; l <= maxL; l++) cnt[l] += cnt[l - 1];
    int32_t *order = malloc(sizeof(int32_t) * (size_t)n);
    order[0] = 0;
#ifndef ORC_MUT_REORDER
    for (int v = n - 1; v >= 1; v--) order[v] = cnt[lev[v]]--;
#else
    for (int v = 1; v < n; v++) order[v] = cnt[lev[v]]--;
#endif
    free(cnt);
    for (int v = 0; v <= nPIs; v++) { of0[v] = 0; of1[v] = 0; }
    for (int v = nPIs + 1; v < n; v++) {
        uint32_t a = 2u * (uint32_t)order[ID(f0[v])] | CMP(f0[v]);
        uint32_t b = 2u * (uint32_t)order[ID(f1[v])] | CMP(f1[v]);
        if (a > b) { uint32_t t = a; a = b; b = t; }
        of0[order[v]] = a; of1[order[v]] = b;
    }
    for (int i = 0; i < nPOs; i++) opos[i] = 2u * (uint32_t)order[ID(pos[i])] | CMP(pos[i]);
    if (olevel) for (int v = 0; v < n; v++) olevel[order[v]] = (uint32_t)lev[v];
    free(order); free(lev);
    return n;
}

/* ------------------------------------------------------------------ */
/* Cut enumeration (rewrite.cu:49-294)                                 */
/* (ORC_MUT_* macros plant one mistake each for tests/test_oracle_     */
/* mutants.py, which checks that the pins catch it; never defined in   */
/* the normal build.)                                                  */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t leaves[4]; uint32_t tt, value; int nLeaves, used; } Cut;

/* Signature bit of each node (rewrite.cu:68, 262: leaf & 31).  A batched pass
 * over a union of AIGs uses every node's bit in its own AIG instead; tests set
 * that table here (indexed by the ids of the reordered graph). */
static const uint8_t *g_sig_bits = NULL;
void orc_set_sig_bits(const uint8_t *bits) { g_sig_bits = bits; }

static uint32_t cut_sign(const Cut *c) {
    uint32_t s = 0;
    for (int i = 0; i < c->nLeaves; i++) s |= 1u << (g_sig_bits ? g_sig_bits[c->leaves[i]] : (c->leaves[i] & 31));
    return s;
}

static uint32_t cut_value(const Cut *c, const int32_t *nRef) {     /* rewrite.cu:49-61 */
    int value = 0, nOnes = 0;
    for (int i = 0; i < c->nLeaves; i++) {
        value += nRef[c->leaves[i]];
        nOnes += (nRef[c->leaves[i]] == 1);
    }
    if (c->nLeaves < 2) return 1001;
    if (value > 1000) value = 1000;
#ifndef ORC_MUT_VALUE
    if (nOnes > 3) value = 5 - nOnes;
#endif
    return (uint32_t)value;
}

static int find_cut(Cut *cs) {                                    /* rewrite.cu:84-109 */
    for (int i = 0; i < NCUT; i++) if (!cs[i].used) return i;
    int ans = -1;
    for (int i = 0; i < NCUT; i++)
#ifndef ORC_MUT_FINDCUT
        if (cs[i].nLeaves > 2 && (ans == -1 || cs[i].value < cs[ans].value)) ans = i;
#else
        if (cs[i].nLeaves > 2 && (ans == -1 || cs[i].value <= cs[ans].value)) ans = i;
#endif
    if (ans == -1)
        for (int i = 0; i < NCUT; i++)
            if (cs[i].nLeaves == 2 && (ans == -1 || cs[i].value < cs[ans].value)) ans = i;
    if (ans == -1)
        for (int i = 0; i < NCUT; i++)
            if (cs[i].nLeaves < 2 && (ans == -1 || cs[i].value < cs[ans].value)) ans = i;
    cs[ans].used = 0;
    return ans;
}

/* a.nLeaves >= b.nLeaves; rewrite.cu:112-152 */
static int merge_ordered(const Cut *a, const Cut *b, Cut *c) {
    int n = 0, i = 0, k = 0;
    for (n = 0; n < 4; n++) {
        if (k == b->nLeaves) {
            if (i == a->nLeaves) { c->nLeaves = n; return 1; }
            c->leaves[n] = a->leaves[i++];
            continue;
        }
        if (i == a->nLeaves) {
            c->leaves[n] = b->leaves[k++];
            continue;
        }
        if (a->leaves[i] < b->leaves[k]) { c->leaves[n] = a->leaves[i++]; continue; }
        if (a->leaves[i] > b->leaves[k]) { c->leaves[n] = b->leaves[k++]; continue; }
        c->leaves[n] = a->leaves[i++];
        k++;
    }
    if (i < a->nLeaves || k < b->nLeaves) return 0;
    c->nLeaves = n;
    return 1;
}

static int merge_cut(const Cut *a, const Cut *b, Cut *c) {        /* rewrite.cu:155-164 */
    if (a->nLeaves >= b->nLeaves) { if (!merge_ordered(a, b, c)) return 0; }
    else if (!merge_ordered(b, a, c)) return 0;
    c->used = 1;
    return 1;
}

static int dominates(const Cut *a, const Cut *b) {                /* a subset of b; :167-176 */
    for (int i = 0; i < a->nLeaves; i++) {
        int ok = 0;
        for (int j = 0; j < b->nLeaves; j++) if (b->leaves[j] == a->leaves[i]) ok = 1;
        if (!ok) return 0;
    }
    return 1;
}

static int cut_filter(Cut *cs, int id) {                          /* rewrite.cu:178-193 */
    uint32_t sid = cut_sign(&cs[id]);
    for (int i = 0; i < NCUT; i++) {
        if (i == id || !cs[i].used) continue;
        uint32_t si = cut_sign(&cs[i]);
        if (cs[i].nLeaves > cs[id].nLeaves) {
            if ((si & sid) != sid) continue;
            if (dominates(&cs[id], &cs[i])) cs[i].used = 0;
        } else {
            if ((si & sid) != si) continue;
#ifndef ORC_MUT_DOMINANCE
            if (dominates(&cs[i], &cs[id])) { cs[id].used = 0; return 1; }
#else
            if (dominates(&cs[id], &cs[i])) { cs[id].used = 0; return 1; }
#endif
        }
    }
    return 0;
}

static uint32_t swap_adj(uint32_t t, int v) {                     /* rewrite.cu:203-211 */
    if (v == 0) return (t & 0x99999999u) | ((t & 0x22222222u) << 1) | ((t & 0x44444444u) >> 1);
    if (v == 1) return (t & 0xC3C3C3C3u) | ((t & 0x0C0C0C0Cu) << 2) | ((t & 0x30303030u) >> 2);
    if (v == 2) return (t & 0xF00FF00Fu) | ((t & 0x00F000F0u) << 4) | ((t & 0x0F000F00u) >> 4);
    return 0;
}

static uint32_t tt_stretch(uint32_t t, int nVar, int phase) {     /* rewrite.cu:213-221 */
    int var = nVar - 1;
    for (int i = 3; i >= 0; i--)
        if ((phase >> i) & 1) {
            for (int k = var; k < i; k++) t = swap_adj(t, k);
            var--;
        }
    return t;
}

static int truth_phase(const Cut *x, const Cut *c) {              /* rewrite.cu:195-201 */
    int ph = 0;
    for (int i = 0; i < c->nLeaves; i++)
        for (int j = 0; j < x->nLeaves; j++)
            if (x->leaves[j] == c->leaves[i]) ph |= 1 << i;
    return ph;
}

static uint32_t tt_shrink(uint32_t t, int phase) {                /* rewrite.cu:231-241 */
    int var = 0;
    for (int i = 0; i < 4; i++)
        if (phase & (1 << i)) {
            for (int k = i - 1; k >= var; k--) t = swap_adj(t, k);
            var++;
        }
    return t;
}

/* The two variable placements above over arrays, for the tests of the
 * library's delta-swap form (tt_place.h): out[i] = tt_stretch(t[i],
 * popcount(phase[i]), phase[i]) when shrink == 0, tt_shrink(t[i], phase[i])
 * otherwise. */
void orc_tt_place(int n, const uint32_t *t, const uint8_t *phase, int shrink, uint32_t *out) {
    for (int i = 0; i < n; i++)
        out[i] = shrink ? tt_shrink(t[i], phase[i]) : tt_stretch(t[i], __builtin_popcount(phase[i]), phase[i]);
}

static int minimize_support(Cut *c) {                             /* rewrite.cu:243-266 */
    static const uint32_t masks[4][2] = {{0x5555, 0xAAAA}, {0x3333, 0xCCCC}, {0x0F0F, 0xF0F0}, {0x00FF, 0xFF00}};
    int phase = 0, nL = c->nLeaves;
    uint32_t t = c->tt & 0xFFFF;
    for (int i = 0; i < c->nLeaves; i++) {
        if ((t & masks[i][0]) == ((t & masks[i][1]) >> (1 << i))) nL--;
        else phase |= 1 << i;
    }
    if (nL == c->nLeaves) return 0;
    t = tt_shrink(t, phase);
    c->tt = t & 0xFFFF;
    int k = 0;
    for (int i = 0; i < c->nLeaves; i++) if ((phase >> i) & 1) c->leaves[k++] = c->leaves[i];
    c->nLeaves = nL;
    return 1;
}

static void enumerate_node(int idx, uint32_t l0, uint32_t l1, Cut *cuts, const int32_t *nRef) {
    Cut *cs = cuts + (size_t)idx * NCUT;
    cs[0].used = 1; cs[0].tt = 0xAAAA; cs[0].nLeaves = 1; cs[0].leaves[0] = (uint32_t)idx;
    cs[0].value = cut_value(&cs[0], nRef);
    int in0 = ID(l0), in1 = ID(l1), c0 = CMP(l0), c1 = CMP(l1);
    for (int i = 0; i < NCUT; i++) {
        if (!cuts[(size_t)in0 * NCUT + i].used) continue;
        for (int j = 0; j < NCUT; j++) {
            if (!cuts[(size_t)in1 * NCUT + j].used) continue;
            Cut a = cuts[(size_t)in0 * NCUT + i], b = cuts[(size_t)in1 * NCUT + j];
            if (__builtin_popcount(cut_sign(&a) | cut_sign(&b)) > 4) continue;
            int id = find_cut(cs);
            if (!merge_cut(&a, &b, &cs[id])) continue;
            if (cut_filter(cs, id)) continue;
            uint32_t t0 = c0 ? ~a.tt : a.tt, t1 = c1 ? ~b.tt : b.tt;
            t0 = tt_stretch(t0, a.nLeaves, truth_phase(&a, &cs[id]));
            t1 = tt_stretch(t1, b.nLeaves, truth_phase(&b, &cs[id]));
            cs[id].tt = (t0 & t1) & 0xFFFF;
            if (minimize_support(&cs[id])) cut_filter(cs, id);
            cs[id].value = cut_value(&cs[id], nRef);
            if (cs[id].nLeaves < 2) return;
        }
    }
}

static void fanout_counts(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
                          const uint32_t *pos, int32_t *nRef) {
    memset(nRef, 0, sizeof(int32_t) * (size_t)n);
    for (int v = nPIs + 1; v < n; v++) { nRef[ID(f0[v])]++; nRef[ID(f1[v])]++; }
    for (int i = 0; i < nPOs; i++) nRef[ID(pos[i])]++;
}

static Cut *enumerate_all(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, const int32_t *nRef) {
    Cut *cuts = calloc((size_t)n * NCUT, sizeof(Cut));
    for (int v = 1; v <= nPIs; v++) {                               /* Inputs, rewrite.cu:63-74 */
        Cut *c = &cuts[(size_t)v * NCUT];
        c->used = 1; c->tt = 0xAAAA; c->nLeaves = 1; c->leaves[0] = (uint32_t)v;
        c->value = cut_value(c, nRef);
    }
    for (int v = nPIs + 1; v < n; v++) enumerate_node(v, f0[v], f1[v], cuts, nRef);
    return cuts;
}

static void pack_cuts(int n, const Cut *cuts, uint32_t *out) {
    for (size_t k = 0; k < (size_t)n * NCUT; k++) {
        const Cut *c = &cuts[k];
        uint32_t *o = out + 5 * k;
        for (int i = 0; i < 4; i++) o[i] = (c->used && i < c->nLeaves) ? c->leaves[i] : 0;
        o[4] = c->used ? ((c->tt & 0xFFFF) | (c->value & 0x7FF) << 16 | (uint32_t)(c->nLeaves & 7) << 27 | 1u << 30) : 0;
    }
}

static void unpack_cuts(int n, const uint32_t *in, Cut *cuts) {
    for (size_t k = 0; k < (size_t)n * NCUT; k++) {
        const uint32_t *o = in + 5 * k;
        Cut *c = &cuts[k];
        for (int i = 0; i < 4; i++) c->leaves[i] = o[i];
        c->tt = o[4] & 0xFFFF; c->value = (o[4] >> 16) & 0x7FF;
        c->nLeaves = (o[4] >> 27) & 7; c->used = (o[4] >> 30) & 1;
    }
}

int orc_cuts(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
             const uint32_t *pos, uint32_t *out) {
    int32_t *nRef = malloc(sizeof(int32_t) * (size_t)n);
    fanout_counts(n, nPIs, f0, f1, nPOs, pos, nRef);
    Cut *cuts = enumerate_all(n, nPIs, f0, f1, nRef);
    pack_cuts(n, cuts, out);
    free(cuts); free(nRef);
    return n;
}

/* ------------------------------------------------------------------ */
/* Evaluation (rewrite.cu:296-447)                                     */
/* ------------------------------------------------------------------ */
typedef struct {
    int n, nPIs;
    const uint32_t *f0, *f1;
    const int32_t *nRef, *level;
    Map tab;                  /* (lit0, lit1) -> AND id, fanins sorted by literal */
} Graph;

typedef struct { int size; int32_t id[MFFC_TABLE]; int32_t num[MFFC_TABLE]; } MffcTab;

static int mffc_decrease(const Graph *g, int x, const int *leaves, int nL, MffcTab *t) {
    for (int i = 0; i < nL; i++) if (x == leaves[i]) return 1;
    for (int i = 0; i < t->size; i++) if (t->id[i] == x) return --t->num[i];
    if (t->size == MFFC_TABLE) return 1;
    t->id[t->size] = x; t->num[t->size] = g->nRef[x] - 1; t->size++;
    return g->nRef[x] - 1;
}

/* CalcMFFC (rewrite.cu:319-326) as an explicit-stack DFS with the same visit
 * order: fanin0's subtree completes before fanin1 is decreased. */
static int calc_mffc(const Graph *g, int root, const int *leaves, int nL, MffcTab *t) {
    int stkNode[MFFC_TABLE + 2], stkStage[MFFC_TABLE + 2], sp = 0, ans = 0;
    stkNode[0] = root; stkStage[0] = 0; sp = 1; ans = 1;
    while (sp) {
        int cur = stkNode[sp - 1];
        if (stkStage[sp - 1] == 0) {
            stkStage[sp - 1] = 1;
            int x = ID(g->f0[cur]);
            if (mffc_decrease(g, x, leaves, nL, t) == 0) { stkNode[sp] = x; stkStage[sp] = 0; sp++; ans++; }
        } else if (stkStage[sp - 1] == 1) {
            stkStage[sp - 1] = 2;
            int x = ID(g->f1[cur]);
            if (mffc_decrease(g, x, leaves, nL, t) == 0) { stkNode[sp] = x; stkStage[sp] = 0; sp++; ans++; }
        } else sp--;
    }
    return ans;
}

static int mffc_deleted(const MffcTab *t, int x) {
    for (int i = 0; i < t->size; i++) if (t->id[i] == x) return t->num[i] == 0;
    return 0;
}

static int table_lookup(const Graph *g, int in0, int in1, int c0, int c1) {  /* rewrite.cu:345-367 */
    if (in0 > in1) { int t = in0; in0 = in1; in1 = t; t = c0; c0 = c1; c1 = t; }
    return map_get(&g->tab, and_key(2u * (uint32_t)in0 + (uint32_t)c0, 2u * (uint32_t)in1 + (uint32_t)c1));
}

/* complement of library input a (only the 4 cut inputs carry one; the
 * reference keeps these bits in a 64-bit word, rewrite.cu:388-395) */
static inline uint32_t isc_bit(uint32_t isC, int a) { return a < 4 ? (isC >> a) & 1u : 0u; }
static inline int lib_in0(int c, int j) { return gls_forest_nodes[gls_forest_node_off[c] + j] & 63; }
static inline int lib_in1(int c, int j) { return (gls_forest_nodes[gls_forest_node_off[c] + j] >> 6) & 63; }
static inline int lib_c0(int c, int j) { return (gls_forest_nodes[gls_forest_node_off[c] + j] >> 12) & 1; }
static inline int lib_c1(int c, int j) { return (gls_forest_nodes[gls_forest_node_off[c] + j] >> 13) & 1; }
static inline int lib_nnodes(int c) { return gls_forest_node_off[c + 1] - gls_forest_node_off[c]; }
static inline int lib_nroots(int c) { return gls_forest_root_off[c + 1] - gls_forest_root_off[c]; }
static inline int lib_root(int c, int k) { return gls_forest_roots[gls_forest_root_off[c] + k]; }

/* Library matching for one cut (rewrite.cu:391-418).  match[] = existing node id
 * or -1; mlev[] = level estimate.  Returns class. */
static int match_cut(const Graph *g, int id, const Cut *cut, const int *leaves, const MffcTab *t,
                     int32_t *match, int32_t *mlev, uint32_t *isC) {
    int tt = (int)cut->tt;
    int uPhase = g_phase[tt], cls = g_cls[tt];
    const uint8_t *perm = &g_perms4[4 * g_perm[tt]];
    for (int j = 0; j < 54; j++) match[j] = mlev[j] = -1;
    *isC = 0;
    for (int j = 0; j < 4; j++) {
        match[j] = leaves[perm[j]];
        mlev[j] = g->level[match[j]];
        if ((uPhase >> j) & 1) *isC |= 1u << j;
    }
    for (int j = 0; j < lib_nnodes(cls); j++) {
        int num = j + 4, a = lib_in0(cls, j), b = lib_in1(cls, j);
        mlev[num] = 1 + (mlev[a] > mlev[b] ? mlev[a] : mlev[b]);
        if (match[a] == -1 || match[b] == -1 || match[a] == id || match[b] == id) continue;
        int nid = table_lookup(g, match[a], match[b], (int)isc_bit(*isC, a) ^ lib_c0(cls, j),
                               (int)isc_bit(*isC, b) ^ lib_c1(cls, j));
        if (nid != -1 && !mffc_deleted(t, nid)) { match[num] = nid; mlev[num] = g->level[nid]; }
    }
    return cls;
}

/* Nodes to add for root rt: unmatched library nodes reachable from rt through
 * unmatched nodes (Eval, rewrite.cu:369-374).  Returns the bit set. */
static uint64_t added_set(int cls, int rt, const int32_t *match) {
    uint64_t used = 1ull << rt, add = 0;
    for (int j = rt; j >= 4; j--) {
#ifndef ORC_MUT_ADDED
        if (!((used >> j) & 1) || match[j] > -1) continue;
#else
        if (!((used >> j) & 1)) continue;
#endif
        add |= 1ull << j;
        used |= 1ull << lib_in0(cls, j - 4);
        used |= 1ull << lib_in1(cls, j - 4);
    }
    return add;
}

static void cut_leaves4(const Cut *c, int *leaves, int *nL) {
    for (int i = 0; i < c->nLeaves; i++) leaves[i] = (int)c->leaves[i];
    *nL = c->nLeaves;
    if (*nL == 3) { leaves[3] = 0; *nL = 4; }
}

static void eval_node(const Graph *g, int id, const Cut *cs, int useZeros, int32_t *best) {
    int reduction = -1, bestLevel = 99999999, bestCut = -1, bestOut = 0;
    int32_t match[54], mlev[54];
    for (int i = 0; i < NCUT; i++) {
        const Cut *cut = &cs[i];
        if (!cut->used || cut->nLeaves < 3) continue;
        int leaves[4], nL;
        cut_leaves4(cut, leaves, &nL);
        MffcTab t; t.size = 0;
        int saved = calc_mffc(g, id, leaves, nL, &t);
        uint32_t isC;
        int cls = match_cut(g, id, cut, leaves, &t, match, mlev, &isC);
        for (int out = 0; out < lib_nroots(cls); out++) {
            int rt = lib_root(cls, out);
            if (match[rt] == id) continue;
            int added = __builtin_popcountll(added_set(cls, rt, match));
            int gain = saved - added;
#ifndef ORC_MUT_ZERO
            if (gain < 0 || (gain == 0 && !useZeros)) continue;
#else
            if (gain < 0) continue;
#endif
            if (gain < reduction || (gain == reduction && mlev[rt] >= bestLevel)) continue;
            reduction = gain; bestLevel = mlev[rt]; bestCut = i; bestOut = out;
        }
    }
    best[0] = bestCut; best[1] = bestCut >= 0 ? bestOut : 0;
    best[2] = bestCut >= 0 ? reduction : 0; best[3] = bestCut >= 0 ? bestLevel : 0;
}

static void graph_init(Graph *g, int n, int nPIs, const uint32_t *f0, const uint32_t *f1,
                       const int32_t *nRef, const int32_t *level) {
    g->n = n; g->nPIs = nPIs; g->f0 = f0; g->f1 = f1; g->nRef = nRef; g->level = level;
    map_init(&g->tab, n);
    for (int v = nPIs + 1; v < n; v++) map_put(&g->tab, and_key(f0[v], f1[v]), v);
}

int orc_eval(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
             const uint32_t *pos, const uint32_t *packed, int useZeros, int32_t *best) {
    npn_init();
    int32_t *nRef = malloc(sizeof(int32_t) * (size_t)n), *lev = malloc(sizeof(int32_t) * (size_t)n);
    fanout_counts(n, nPIs, f0, f1, nPOs, pos, nRef);
    compute_levels(n, nPIs, f0, f1, NULL, lev);
    Cut *cuts = malloc(sizeof(Cut) * (size_t)n * NCUT);
    unpack_cuts(n, packed, cuts);
    Graph g; graph_init(&g, n, nPIs, f0, f1, nRef, lev);
    for (int v = 0; v < n; v++) {
        if (v == 0) { best[0] = -1; best[1] = best[2] = best[3] = 0; continue; }
        eval_node(&g, v, &cuts[(size_t)v * NCUT], useZeros, &best[4 * v]);
    }
    map_free(&g.tab); free(cuts); free(nRef); free(lev);
    return n;
}

/* MFFC of `root` bounded by `leaves` as counted by the evaluation (8-entry
 * table); *tabSize receives the table fill so callers know when it is exact. */
int orc_mffc(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs, const uint32_t *pos,
             int root, const int *leavesIn, int nL, int *tabSize) {
    int32_t *nRef = malloc(sizeof(int32_t) * (size_t)n);
    fanout_counts(n, nPIs, f0, f1, nPOs, pos, nRef);
    Graph g;
    g.n = n; g.nPIs = nPIs; g.f0 = f0; g.f1 = f1; g.nRef = nRef; g.level = NULL;
    int leaves[4] = {0, 0, 0, 0};
    for (int i = 0; i < nL && i < 4; i++) leaves[i] = leavesIn[i];
    MffcTab t; t.size = 0;
    int saved = calc_mffc(&g, root, leaves, nL, &t);
    if (tabSize) *tabSize = t.size;
    free(nRef);
    return saved;
}

/* ------------------------------------------------------------------ */
/* Replacement plan for one candidate and this package's conflict-free  */
/* selection (sequential restatement of the GPU's priority rounds).     */
/* ------------------------------------------------------------------ */
typedef struct {
    int id, gain, cls, rt, valid, compl;
    int nW, nL, nM, nNew;
    int32_t W[MFFC_TABLE + 1], L[4], M[64];
    uint32_t isC;
    uint64_t addMask;
    int32_t match[54];
} Plan;

static uint32_t lib_truth(int cls, int rt, const uint8_t *perm, uint32_t isC) {
    static const uint32_t ev[4] = {0xAAAA, 0xCCCC, 0xF0F0, 0xFF00};
    uint32_t tt[64];
    for (int j = 0; j < 4; j++) tt[j] = (ev[perm[j]] ^ (((isC >> j) & 1) ? 0xFFFF : 0)) & 0xFFFF;
    for (int j = 4; j <= rt; j++) {
        int k = j - 4;
        uint32_t a = tt[lib_in0(cls, k)] ^ (lib_c0(cls, k) ? 0xFFFF : 0);
        uint32_t b = tt[lib_in1(cls, k)] ^ (lib_c1(cls, k) ? 0xFFFF : 0);
        tt[j] = a & b & 0xFFFF;
    }
    return tt[rt];
}

static void make_plan(const Graph *g, int id, const Cut *cut, int out, int gain, Plan *p) {
    int leaves[4], nL;
    cut_leaves4(cut, leaves, &nL);
    MffcTab t; t.size = 0;
    calc_mffc(g, id, leaves, nL, &t);
    int32_t mlev[54];
    uint32_t isC;
    int cls = match_cut(g, id, cut, leaves, &t, p->match, mlev, &isC);
    int rt = lib_root(cls, out);
    p->id = id; p->gain = gain; p->cls = cls; p->rt = rt; p->isC = isC; p->valid = 1;
    p->nW = 0;
    p->W[p->nW++] = id;
    for (int i = 0; i < t.size; i++) if (t.num[i] == 0) p->W[p->nW++] = t.id[i];
    p->nL = 0;
    for (int i = 0; i < cut->nLeaves; i++) p->L[p->nL++] = (int32_t)cut->leaves[i];
    p->addMask = added_set(cls, rt, p->match);
    p->nNew = __builtin_popcountll(p->addMask);
    /* matched library nodes (index >= 4) used by the root's cone */
    uint64_t used = 1ull << rt;
    for (int j = rt; j >= 4; j--) {
        if (!((used >> j) & 1)) continue;
        used |= 1ull << lib_in0(cls, j - 4);
        used |= 1ull << lib_in1(cls, j - 4);
    }
    p->nM = 0;
    for (int j = 4; j <= rt; j++)
        if (((used >> j) & 1) && p->match[j] > -1) {
            int dup = 0;
            for (int q = 0; q < p->nM; q++) if (p->M[q] == p->match[j]) dup = 1;
            if (!dup) p->M[p->nM++] = p->match[j];
            if (p->match[j] == id) p->valid = 0;
        }
    /* output polarity: library root function over the cut leaves vs the cut */
    uint32_t rtt = lib_truth(cls, rt, &g_perms4[4 * g_perm[cut->tt]], isC);
    if (rtt == cut->tt) p->compl = 0;
    else if (rtt == ((~cut->tt) & 0xFFFF)) p->compl = 1;
    else p->valid = 0;
}

/* priority: larger gain first, then smaller id */
static int plan_cmp(const void *x, const void *y) {
    const Plan *a = *(const Plan *const *)x, *b = *(const Plan *const *)y;
    if (a->gain != b->gain) return a->gain > b->gain ? -1 : 1;
    return a->id < b->id ? -1 : (a->id > b->id);
}

enum { MK_W = 1, MK_WNR = 2, MK_L = 4, MK_M = 8 };

static int plan_conflicts(const Plan *p, const uint8_t *mk) {
    for (int i = 0; i < p->nW; i++) {
        int x = p->W[i];
        if (mk[x] & (MK_W | MK_M)) return 1;
        if (x != p->id && (mk[x] & MK_L)) return 1;
    }
    for (int i = 0; i < p->nL; i++) if (mk[p->L[i]] & MK_WNR) return 1;
    for (int i = 0; i < p->nM; i++) if (mk[p->M[i]] & MK_W) return 1;
    return 0;
}

static void plan_mark(const Plan *p, uint8_t *mk) {
    for (int i = 0; i < p->nW; i++) mk[p->W[i]] |= MK_W | (p->W[i] != p->id ? MK_WNR : 0);
    for (int i = 0; i < p->nL; i++) mk[p->L[i]] |= MK_L;
    for (int i = 0; i < p->nM; i++) mk[p->M[i]] |= MK_M;
}

typedef struct {
    int n, nPIs, nPOs;
    uint32_t *f0, *f1, *pos;
    int32_t *nRef, *level;
} RwIn;

static void rw_prepare(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
                       const uint32_t *pos, RwIn *r) {
    r->n = n; r->nPIs = nPIs; r->nPOs = nPOs;
    r->f0 = malloc(sizeof(uint32_t) * (size_t)n); r->f1 = malloc(sizeof(uint32_t) * (size_t)n);
    r->pos = malloc(sizeof(uint32_t) * (size_t)(nPOs + 1));
    uint32_t *lev = malloc(sizeof(uint32_t) * (size_t)n);
    orc_reorder(n, nPIs, f0, f1, nPOs, pos, r->f0, r->f1, r->pos, lev);
    r->level = (int32_t *)lev;
    r->nRef = malloc(sizeof(int32_t) * (size_t)n);
    fanout_counts(n, nPIs, r->f0, r->f1, nPOs, r->pos, r->nRef);
}

static void rw_free(RwIn *r) { free(r->f0); free(r->f1); free(r->pos); free(r->nRef); free(r->level); }

static int g_dbg_max_sel = -1;   /* debugging aid: apply at most this many selections */
void orc_debug_max_sel(int k) { g_dbg_max_sel = k; }

static int rw_impl(int n0, int nPIs, const uint32_t *f0in, const uint32_t *f1in, int nPOs,
                   const uint32_t *posin, int useZeros, int cap, uint32_t *of0, uint32_t *of1,
                   uint32_t *opos, int64_t *stats, int composeOnly);

int orc_rw(int n0, int nPIs, const uint32_t *f0in, const uint32_t *f1in, int nPOs,
           const uint32_t *posin, int useZeros, int cap, uint32_t *of0, uint32_t *of1,
           uint32_t *opos, int64_t *stats) {
    return rw_impl(n0, nPIs, f0in, f1in, nPOs, posin, useZeros, cap, of0, of1, opos, stats, 0);
}

int orc_rw_compose(int n0, int nPIs, const uint32_t *f0in, const uint32_t *f1in, int nPOs,
                   const uint32_t *posin, int useZeros, int cap, uint32_t *of0, uint32_t *of1,
                   uint32_t *opos, int64_t *stats) {
    return rw_impl(n0, nPIs, f0in, f1in, nPOs, posin, useZeros, cap, of0, of1, opos, stats, 1);
}

static int rw_impl(int n0, int nPIs, const uint32_t *f0in, const uint32_t *f1in, int nPOs,
                   const uint32_t *posin, int useZeros, int cap, uint32_t *of0, uint32_t *of1,
                   uint32_t *opos, int64_t *stats, int composeOnly) {
    npn_init();
    RwIn r;
    rw_prepare(n0, nPIs, f0in, f1in, nPOs, posin, &r);
    int n = r.n;
    Cut *cuts = enumerate_all(n, nPIs, r.f0, r.f1, r.nRef);
    Graph g; graph_init(&g, n, nPIs, r.f0, r.f1, r.nRef, r.level);
    int32_t best[4];
    int nCand = 0;
    Plan *plans = malloc(sizeof(Plan) * 1024);
    int capPlans = 1024;
    for (int v = nPIs + 1; v < n; v++) {
        eval_node(&g, v, &cuts[(size_t)v * NCUT], useZeros, best);
        if (best[0] < 0) continue;
        if (nCand == capPlans) { capPlans *= 2; plans = realloc(plans, sizeof(Plan) * (size_t)capPlans); }
        make_plan(&g, v, &cuts[(size_t)v * NCUT + best[0]], best[1], best[2], &plans[nCand]);
        nCand++;
    }
    Plan **ordp = malloc(sizeof(Plan *) * (size_t)(nCand + 1));
    for (int i = 0; i < nCand; i++) ordp[i] = &plans[i];
    qsort(ordp, (size_t)nCand, sizeof(Plan *), plan_cmp);
    uint8_t *mk = calloc((size_t)n, 1), *sel = calloc((size_t)nCand + 1, 1);
    int nSel = 0;
    int64_t estGain = 0;
    for (int i = 0; i < nCand; i++) {
        Plan *p = ordp[i];
        if (!p->valid || plan_conflicts(p, mk)) continue;
        if (g_dbg_max_sel >= 0 && nSel >= g_dbg_max_sel) break;
        plan_mark(p, mk);
        sel[p - plans] = 1; nSel++; estGain += p->gain;
    }
    /* build new nodes: selected candidates in id order, library index order */
    int64_t nNew = 0;
    for (int i = 0; i < nCand; i++) if (sel[i]) nNew += plans[i].nNew;
    int N = n + (int)nNew;
    uint32_t *c0 = malloc(sizeof(uint32_t) * (size_t)N), *c1 = malloc(sizeof(uint32_t) * (size_t)N);
    uint32_t *cpos = malloc(sizeof(uint32_t) * (size_t)(nPOs + 1));
    uint32_t *repl = malloc(sizeof(uint32_t) * (size_t)n);
    for (int v = 0; v < n; v++) repl[v] = 2u * (uint32_t)v;
    int next = n;
    for (int i = 0; i < nCand; i++) {
        if (!sel[i]) continue;
        Plan *p = &plans[i];
        int32_t nodeOf[54];
        for (int j = 0; j < 54; j++) nodeOf[j] = p->match[j];
        for (int j = 4; j <= p->rt; j++) {
            if (!((p->addMask >> j) & 1)) continue;
            int k = j - 4, a = lib_in0(p->cls, k), b = lib_in1(p->cls, k);
            uint32_t la = 2u * (uint32_t)nodeOf[a] + ((isc_bit(p->isC, a) ^ (uint32_t)lib_c0(p->cls, k)));
            uint32_t lb = 2u * (uint32_t)nodeOf[b] + ((isc_bit(p->isC, b) ^ (uint32_t)lib_c1(p->cls, k)));
            nodeOf[j] = next;
            c0[next] = la; c1[next] = lb;
            next++;
        }
        repl[p->id] = 2u * (uint32_t)nodeOf[p->rt] + (uint32_t)p->compl;
    }
    /* compose: substitute selected roots everywhere (one level suffices) */
    for (int v = 0; v <= nPIs; v++) { c0[v] = 0; c1[v] = 0; }
    for (int v = nPIs + 1; v < n; v++) {
        c0[v] = repl[ID(r.f0[v])] ^ CMP(r.f0[v]);
        c1[v] = repl[ID(r.f1[v])] ^ CMP(r.f1[v]);
    }
    for (int v = n; v < N; v++) {
        uint32_t a = c0[v], b = c1[v];
        c0[v] = (ID(a) < (uint32_t)n ? repl[ID(a)] : 2u * ID(a)) ^ CMP(a);
        c1[v] = (ID(b) < (uint32_t)n ? repl[ID(b)] : 2u * ID(b)) ^ CMP(b);
    }
    for (int i = 0; i < nPOs; i++) cpos[i] = repl[ID(r.pos[i])] ^ CMP(r.pos[i]);
    int ret;
    if (N > cap) ret = -2;
    else if (composeOnly) {
        memcpy(of0, c0, sizeof(uint32_t) * (size_t)N);
        memcpy(of1, c1, sizeof(uint32_t) * (size_t)N);
        memcpy(opos, cpos, sizeof(uint32_t) * (size_t)nPOs);
        ret = N;
        if (stats) {
            stats[0] = nCand; stats[1] = nSel; stats[2] = nNew; stats[3] = N;
            stats[4] = 0; stats[5] = 0; stats[6] = estGain; stats[7] = 0;
        }
    } else {
        int depth;
        ret = orc_strash(N, nPIs, c0, c1, nPOs, cpos, of0, of1, opos, &depth);
        if (stats) {
            stats[0] = nCand; stats[1] = nSel; stats[2] = nNew; stats[3] = N;
            stats[4] = depth; stats[5] = 0; stats[6] = estGain; stats[7] = 0;
        }
    }
    free(c0); free(c1); free(cpos); free(repl); free(mk); free(sel); free(ordp); free(plans);
    map_free(&g.tab); free(cuts); rw_free(&r);
    return ret;
}

/* ------------------------------------------------------------------ */
/* The reference's replacement method (QoR comparison only).           */
/* ------------------------------------------------------------------ */
typedef struct {
    int nIn, n, cap;
    int32_t *fanin0, *fanin1;   /* literals, AIGER convention */
    int32_t *ref, *isDel, *phase;
} Cpu;

static int is_redundant(const Cpu *s, int v) {                    /* rewrite.h:118-120 */
    return v > s->nIn && (s->fanin0[v] == s->fanin1[v] || s->fanin0[v] == 1);
}

/* FastDelete exactly as recursive in the reference: when a fanin's count reaches
 * zero inside the fanin0 subtree and is tested again for fanin1, it is deleted
 * twice (a reference quirk kept so the comparison follows its real behaviour). */
static int fast_delete(Cpu *s, int node) {                        /* rewrite.h:208-220 */
    if (node <= s->nIn) return 0;
    int a = ID(s->fanin0[node]), b = ID(s->fanin1[node]);
    s->ref[a]--; s->ref[b]--;
    int cnt = !is_redundant(s, node);
    if (s->ref[a] == 0) cnt += fast_delete(s, a);
    if (s->ref[b] == 0) cnt += fast_delete(s, b);
    s->isDel[node] = 1;
    return cnt;
}

static int install(Cpu *s, int idx, int32_t *stk) {               /* rewrite.cu:837-849 */
    int cnt = 0, sp = 0;
    stk[sp++] = idx;
    while (sp) {
        int v = stk[--sp];
        if (!s->isDel[v]) continue;
        s->isDel[v] = 0;
        cnt += !is_redundant(s, v);
        int a = ID(s->fanin0[v]), b = ID(s->fanin1[v]);
        s->ref[a]++; s->ref[b]++;
        stk[sp++] = a; stk[sp++] = b;
    }
    return cnt;
}

static int32_t fanin_resolve(const Cpu *s, int32_t lit) {         /* rewrite.h:121-124 */
    int32_t c = 0;
    while (is_redundant(s, ID(lit))) { c ^= CMP(lit); lit = s->fanin1[ID(lit)]; }
    return lit ^ c;
}

int orc_rw_culs(int n0, int nPIs, const uint32_t *f0in, const uint32_t *f1in, int nPOs,
                const uint32_t *posin, int useZeros, int cap, uint32_t *of0, uint32_t *of1,
                uint32_t *opos, int64_t *stats) {
    npn_init();
    RwIn r;
    rw_prepare(n0, nPIs, f0in, f1in, nPOs, posin, &r);
    int n = r.n;
    Cut *cuts = enumerate_all(n, nPIs, r.f0, r.f1, r.nRef);
    Graph g; graph_init(&g, n, nPIs, r.f0, r.f1, r.nRef, r.level);
    int32_t *bestCut = malloc(sizeof(int32_t) * (size_t)n), *bestOut = malloc(sizeof(int32_t) * (size_t)n);
    for (int v = 0; v < n; v++) bestCut[v] = -1;
    int32_t best[4];
    for (int v = nPIs + 1; v < n; v++) {
        eval_node(&g, v, &cuts[(size_t)v * NCUT], useZeros, best);
        bestCut[v] = best[0]; bestOut[v] = best[1];
    }
    /* phase: value under the all-zero input (rewrite.cu:1239) */
    int32_t *phase0 = calloc((size_t)n, sizeof(int32_t));
    for (int v = nPIs + 1; v < n; v++)
        phase0[v] = (phase0[ID(r.f0[v])] ^ (int32_t)CMP(r.f0[v])) & (phase0[ID(r.f1[v])] ^ (int32_t)CMP(r.f1[v]));
    /* ReplaceSubgr/BuildSubgr (rewrite.cu:473-555), in id order */
    int capN = n + 16;
    int32_t *t0 = malloc(sizeof(int32_t) * (size_t)capN), *t1 = malloc(sizeof(int32_t) * (size_t)capN);
    int32_t *ph = malloc(sizeof(int32_t) * (size_t)capN);
    int32_t *replace = malloc(sizeof(int32_t) * (size_t)n);
    for (int v = 0; v < n; v++) { t0[v] = (int32_t)r.f0[v]; t1[v] = (int32_t)r.f1[v]; ph[v] = phase0[v]; replace[v] = -1; }
    int N = n;
    Map newTab; map_init(&newTab, n);
    for (int id = 1; id < n; id++) {
        if (bestCut[id] < 0) continue;
        const Cut *cut = &cuts[(size_t)id * NCUT + bestCut[id]];
        if (cut->nLeaves < 3) continue;
        int leaves[4], nL;
        cut_leaves4(cut, leaves, &nL);
        int tt = (int)cut->tt, cls = g_cls[tt], uPhase = g_phase[tt];
        const uint8_t *perm = &g_perms4[4 * g_perm[tt]];
        int32_t match[54];
        for (int j = 0; j < 54; j++) match[j] = -1;
        uint32_t isC = 0;
        for (int j = 0; j < 4; j++) { match[j] = leaves[perm[j]]; if ((uPhase >> j) & 1) isC |= 1u << j; }
        int rt = lib_root(cls, bestOut[id]);
        uint64_t used = 1ull << rt;
        for (int j = rt; j >= 4; j--) if ((used >> j) & 1) {
            used |= 1ull << lib_in0(cls, j - 4);
            used |= 1ull << lib_in1(cls, j - 4);
        }
        for (int j = 0; j < lib_nnodes(cls); j++) {
            if (!((used >> (j + 4)) & 1)) continue;
            int a = lib_in0(cls, j), b = lib_in1(cls, j);
            if (match[a] == -1 || match[b] == -1) continue;
            int nid = table_lookup(&g, match[a], match[b], (int)isc_bit(isC, a) ^ lib_c0(cls, j),
                                   (int)isc_bit(isC, b) ^ lib_c1(cls, j));
            if (nid != -1) match[j + 4] = nid;
        }
        if (match[rt] == id) continue;
        /* DFS order of creation matters only for the temporary ids, which the
         * rebuild below renumbers; create in a DFS post-order anyway */
        int dstk[64], dst[64], dsp = 0;
        dstk[dsp] = rt; dst[dsp] = 0; dsp++;
        while (dsp) {
            int cur = dstk[dsp - 1];
            if (match[cur] != -1) { dsp--; continue; }
            int k = cur - 4, a = lib_in0(cls, k), b = lib_in1(cls, k);
            if (dst[dsp - 1] == 0) { dst[dsp - 1] = 1; dstk[dsp] = a; dst[dsp] = 0; dsp++; continue; }
            if (dst[dsp - 1] == 1) { dst[dsp - 1] = 2; dstk[dsp] = b; dst[dsp] = 0; dsp++; continue; }
            dsp--;
            if (N + 1 >= capN) {
                capN *= 2;
                t0 = realloc(t0, sizeof(int32_t) * (size_t)capN); t1 = realloc(t1, sizeof(int32_t) * (size_t)capN);
                ph = realloc(ph, sizeof(int32_t) * (size_t)capN);
            }
            int node = N;
            int32_t in0 = match[a], in1 = match[b];
            int32_t cc0 = (int32_t)(isc_bit(isC, a) ^ (uint32_t)lib_c0(cls, k));
            int32_t cc1 = (int32_t)(isc_bit(isC, b) ^ (uint32_t)lib_c1(cls, k));
            if (in0 > in1) { int32_t tmp = in0; in0 = in1; in1 = tmp; tmp = cc0; cc0 = cc1; cc1 = tmp; }
            int32_t l0 = 2 * in0 + cc0, l1 = 2 * in1 + cc1;
            int32_t actual = map_put(&newTab, and_key((uint32_t)l0, (uint32_t)l1), node);
            if (actual == node) {
                t0[node] = l0; t1[node] = l1;
                ph[node] = (ph[in0] ^ cc0) & (ph[in1] ^ cc1);
                N++;
            }
            match[cur] = actual;
        }
        replace[id] = 2 * match[rt] + (ph[match[rt]] ^ ph[id]);
    }
    map_free(&newTab);
    /* re-id new nodes consecutively (rewrite.cu:743-762) */
    Cpu s;
    s.nIn = nPIs; s.cap = n + (N - n) + 16;
    s.fanin0 = malloc(sizeof(int32_t) * (size_t)s.cap); s.fanin1 = malloc(sizeof(int32_t) * (size_t)s.cap);
    s.ref = calloc((size_t)s.cap, sizeof(int32_t)); s.isDel = calloc((size_t)s.cap, sizeof(int32_t));
    for (int v = 0; v < n; v++) { s.fanin0[v] = (int32_t)r.f0[v]; s.fanin1[v] = (int32_t)r.f1[v]; }
    int32_t *mapping = malloc(sizeof(int32_t) * (size_t)(N + 1));
    for (int v = 0; v <= N; v++) mapping[v] = -1;
    int nn = n;
    s.n = n;
    int32_t *rstk = malloc(sizeof(int32_t) * (size_t)(2 * (N + 2)));
    int32_t *rst = malloc(sizeof(int32_t) * (size_t)(2 * (N + 2)));
    for (int id = 1; id < nn; id++) {
        if (replace[id] == -1) continue;
        int root = replace[id] / 2;
        if (root < nn) continue;
        int sp = 0;
        rstk[sp] = root; rst[sp] = 0; sp++;
        while (sp) {
            int cur = rstk[sp - 1];
            if (cur < nn || mapping[cur] != -1) { sp--; continue; }
            if (rst[sp - 1] == 0) { rst[sp - 1] = 1; rstk[sp] = t0[cur] / 2; rst[sp] = 0; sp++; continue; }
            if (rst[sp - 1] == 1) { rst[sp - 1] = 2; rstk[sp] = t1[cur] / 2; rst[sp] = 0; sp++; continue; }
            sp--;
            int a = t0[cur] / 2, b = t1[cur] / 2;
            int ma = a < nn ? a : mapping[a], mb = b < nn ? b : mapping[b];
            int nid = s.n;
            s.fanin0[nid] = 2 * ma + (t0[cur] & 1);
            s.fanin1[nid] = 2 * mb + (t1[cur] & 1);
            mapping[cur] = nid;
            s.n++;
        }
    }
    for (int id = 1; id < nn; id++)
        if (replace[id] != -1) {
            int root = replace[id] / 2;
            replace[id] = 2 * (root < nn ? root : mapping[root]) + (replace[id] & 1);
        }
    free(rstk); free(rst); free(mapping); free(t0); free(t1); free(ph);
    int total = s.n;       /* ids 0 .. total-1 */
    /* ref counts of Init (rewrite.cu:1234-1242) */
    for (int v = nPIs + 1; v < nn; v++) { s.ref[ID(s.fanin0[v])]++; s.ref[ID(s.fanin1[v])]++; }
    int32_t *outp = malloc(sizeof(int32_t) * (size_t)(nPOs + 1));
    for (int i = 0; i < nPOs; i++) { outp[i] = (int32_t)r.pos[i]; s.ref[ID(outp[i])]++; }
    for (int v = nn; v < total; v++) { s.isDel[v] = 1; s.ref[v] = 0; }
    /* choice pass (rewrite.cu:851-893) */
    int32_t *stk = malloc(sizeof(int32_t) * (size_t)(4 * total + 16));
    int64_t replaced = 0, expected = 0;
    for (int i = nPIs + 1; i < nn; i++) {
        if (bestCut[i] < 0) continue;
        Cut cut = cuts[(size_t)i * NCUT + bestCut[i]];
        if (cut.nLeaves < 3) continue;
        if (replace[i] == -1) continue;
        int newIdx = replace[i] / 2;
        if (newIdx == i || (int)ID(s.fanin1[newIdx]) == i) continue;
        int leaves[4], nL;
        cut_leaves4(&cut, leaves, &nL);
        int comp = 0;
        for (int j = 0; j < nL; j++) if (s.isDel[leaves[j]]) comp = 1;
        if (comp) continue;
        int added = install(&s, newIdx, stk);
        s.ref[newIdx]++;
        int saved = fast_delete(&s, i);
        s.ref[newIdx]--;
        if (saved - added > 0 || (saved == added && useZeros)) {
            replaced++; expected += saved - added;
            s.isDel[i] = 0;
            s.fanin0[i] = 1; s.fanin1[i] = replace[i];
            s.ref[0]++; s.ref[newIdx]++;
        } else {
            install(&s, i, stk);
            s.ref[i]++;
            if (s.ref[newIdx] == 0) fast_delete(&s, newIdx);
            s.ref[i]--;
        }
    }
    /* redundancy removal (rewrite.cu:911-921) */
    for (int i = nPIs + 1; i < total; i++) {
        if (s.isDel[i]) continue;
        if (!is_redundant(&s, i)) {
            s.fanin0[i] = fanin_resolve(&s, s.fanin0[i]);
            s.fanin1[i] = fanin_resolve(&s, s.fanin1[i]);
        } else s.isDel[i] = 1;
    }
    for (int i = 0; i < nPOs; i++) outp[i] = fanin_resolve(&s, outp[i]);
    /* Reorder keeps every node (rewrite.cu:1252-1255); the dedupe strash then
     * drops what is unreachable.  Reorder is a renumbering, so strash directly. */
    uint32_t *u0 = malloc(sizeof(uint32_t) * (size_t)total), *u1 = malloc(sizeof(uint32_t) * (size_t)total);
    uint32_t *upos = malloc(sizeof(uint32_t) * (size_t)(nPOs + 1));
    for (int v = 0; v < total; v++) {
        u0[v] = v <= nPIs ? 0 : (uint32_t)s.fanin0[v];
        u1[v] = v <= nPIs ? 0 : (uint32_t)s.fanin1[v];
    }
    for (int i = 0; i < nPOs; i++) upos[i] = (uint32_t)outp[i];
    /* reorder first so that the strash order follows the reference */
    uint32_t *v0 = malloc(sizeof(uint32_t) * (size_t)total), *v1 = malloc(sizeof(uint32_t) * (size_t)total);
    uint32_t *vpos = malloc(sizeof(uint32_t) * (size_t)(nPOs + 1));
    int ret;
    if (orc_reorder(total, nPIs, u0, u1, nPOs, upos, v0, v1, vpos, NULL) < 0) ret = -3;
    else if (total > cap) ret = -2;
    else {
        int depth;
        ret = orc_strash(total, nPIs, v0, v1, nPOs, vpos, of0, of1, opos, &depth);
        if (stats) {
            stats[0] = replaced; stats[1] = expected; stats[2] = total - nn; stats[3] = total;
            stats[4] = depth; stats[5] = stats[6] = stats[7] = 0;
        }
    }
    free(u0); free(u1); free(upos); free(v0); free(v1); free(vpos);
    free(stk); free(outp); free(s.fanin0); free(s.fanin1); free(s.ref); free(s.isDel);
    free(replace); free(phase0); free(bestCut); free(bestOut);
    map_free(&g.tab); free(cuts); rw_free(&r);
    return ret;
}

/* ------------------------------------------------------------------ */
/* Simulation-based equivalence                                        */
/* ------------------------------------------------------------------ */
static uint64_t splitmix(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void sim_words(int n, int nPIs, const uint32_t *f0, const uint32_t *f1, const uint64_t *pi, uint64_t *val) {
    val[0] = 0;
    for (int v = 1; v <= nPIs; v++) val[v] = pi[v - 1];
    for (int v = nPIs + 1; v < n; v++) {
        uint64_t a = val[ID(f0[v])] ^ (CMP(f0[v]) ? ~0ull : 0);
        uint64_t b = val[ID(f1[v])] ^ (CMP(f1[v]) ? ~0ull : 0);
        val[v] = a & b;
    }
}

/* Pattern words of the library's simulation (gls_simulate / gls_cec): the
 * exhaustive pattern p = 64w + b gives PI j bit j of p; the random word w of
 * PI j is splitmix64's output function at state seed + (j << 32 | w) + 1
 * golden-gamma steps.  An own definition (the reference's simulation-based
 * CEC is closed-source, README.md:102-111). */
static uint64_t sim_pi_word(int j, uint64_t w, uint64_t seed, int exhaustive) {
    if (!exhaustive) {
        uint64_t st = seed + ((((uint64_t)j << 32) | w)) * 0x9E3779B97F4A7C15ULL;
        return splitmix(&st);
    }
    uint64_t x = 0;
    for (int b = 0; b < 64; b++) {
        uint64_t p = w * 64 + (uint64_t)b;
        if (j < 64 && ((p >> j) & 1)) x |= 1ull << b;
    }
    return x;
}

void orc_sim_po_words(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs, const uint32_t *pos,
                      int nWords, uint64_t seed, int exhaustive, uint64_t *out) {
    uint64_t *val = malloc(sizeof(uint64_t) * (size_t)(nObjs > 0 ? nObjs : 1));
    uint64_t *pi = malloc(sizeof(uint64_t) * (size_t)(nPIs + 1));
    for (int w = 0; w < nWords; w++) {
        for (int j = 0; j < nPIs; j++) pi[j] = sim_pi_word(j, (uint64_t)w, seed, exhaustive);
        sim_words(nObjs, nPIs, f0, f1, pi, val);
        for (int o = 0; o < nPOs; o++)
            out[(size_t)o * nWords + w] = val[ID(pos[o])] ^ (CMP(pos[o]) ? ~0ull : 0);
    }
    free(val);
    free(pi);
}

int orc_sim_cmp(int nA, const uint32_t *a0, const uint32_t *a1, const uint32_t *aPos,
                int nB, const uint32_t *b0, const uint32_t *b1, const uint32_t *bPos,
                int nPIs, int nPOs, int nWords, uint64_t seed) {
    uint64_t *va = malloc(sizeof(uint64_t) * (size_t)nA), *vb = malloc(sizeof(uint64_t) * (size_t)nB);
    uint64_t *pi = malloc(sizeof(uint64_t) * (size_t)(nPIs + 1));
    int exhaustive = nPIs <= 16;
    int rounds = exhaustive ? ((1 << nPIs) + 63) / 64 : nWords;
    uint64_t st = seed;
    int bad = -1;
    for (int w = 0; w < rounds && bad < 0; w++) {
        for (int i = 0; i < nPIs; i++) {
            if (exhaustive) {
                uint64_t x = 0;
                for (int b = 0; b < 64; b++) {
                    uint64_t m = (uint64_t)w * 64 + (uint64_t)b;
                    if ((m >> i) & 1) x |= 1ull << b;
                }
                pi[i] = x;
            } else pi[i] = splitmix(&st);
        }
        sim_words(nA, nPIs, a0, a1, pi, va);
        sim_words(nB, nPIs, b0, b1, pi, vb);
        uint64_t mask = ~0ull;
        if (exhaustive && (uint64_t)(w + 1) * 64 > (1ull << nPIs)) mask = (1ull << ((1ull << nPIs) - (uint64_t)w * 64)) - 1;
        for (int o = 0; o < nPOs; o++) {
            uint64_t x = va[ID(aPos[o])] ^ (CMP(aPos[o]) ? ~0ull : 0);
            uint64_t y = vb[ID(bPos[o])] ^ (CMP(bPos[o]) ? ~0ull : 0);
            if ((x ^ y) & mask) { bad = o; break; }
        }
    }
    free(va); free(vb); free(pi);
    return bad;
}
