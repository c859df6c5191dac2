/* CPU oracle for the rw/strash hot path — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker.  The product (culs_amd / libgls.so)
 * never links, loads or calls it.
 *
 * Graph convention used throughout: AIGER literals (lit = 2*id + compl,
 * lit 0 = constant FALSE, lit 1 = constant TRUE), node 0 = constant,
 * ids 1..nPIs = PIs, ids > nPIs = AND nodes.  f0/f1 are indexed by id; entries
 * for the constant and the PIs are 0.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* NPN tables (closed form).  phase: bits 0..3 input complements, bit 4 output
 * complement; perm: index into perms4; cls: class 0..221; perms4: 24x4. */
void orc_npn_tables(uint8_t *phase, uint8_t *perm, uint8_t *cls, uint8_t *perms4);

/* Structural hashing.  Returns nObjs of the result (or <0 on error).  Output
 * arrays must hold nObjs entries (fanins) / nPOs (pos).  *pDepth = number of
 * level waves of the dangling-free input. */
int orc_strash(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
               const uint32_t *pos, uint32_t *of0, uint32_t *of1, uint32_t *opos, int *pDepth);

/* Host strash of `st -c`: dangling removal, id-order hashing, no trivial
 * folding.  olevel receives node delays; *pDepth = largest delay, -1 if no
 * AND remains.  Returns nObjs of the result. */
int orc_strash_cpu(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
                   const uint32_t *pos, uint32_t *of0, uint32_t *of1, uint32_t *opos, uint32_t *olevel,
                   int *pDepth);

/* Balance `b` (sortDecId = 1) without the strash that the `b` command runs
 * afterwards.  cap: capacity of of0/of1.  Returns nObjs of the result, or
 * -2 if it would exceed cap. */
int orc_balance(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs, const uint32_t *pos,
                int cap, uint32_t *of0, uint32_t *of1, uint32_t *opos);
/* The same with the reference's sortDecId: 1 = `b` (level, then literal),
 * 0 = `b -s` (level, ties in thrust's shuffle order, seeded by the cover's
 * index in its batch). */
int orc_balance_ex(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs, const uint32_t *pos,
                   int cap, uint32_t *of0, uint32_t *of1, uint32_t *opos, int sortDecId);
/* thrust::shuffle(thrust::seq, first, first + len, default_random_engine(seed))
 * as a permutation: output position k holds input element perm[k]. */
void orc_thrust_shuffle_perm(uint32_t len, uint32_t seed, uint32_t *perm);
void orc_tt_place(int n, const uint32_t *t, const uint8_t *phase, int shrink, uint32_t *out);

/* Level-major renumbering (ties by id) of an AIG whose ids are topological. */
int orc_reorder(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
                const uint32_t *pos, uint32_t *of0, uint32_t *of1, uint32_t *opos,
                uint32_t *olevel);

/* 4-feasible cut enumeration on a reordered AIG.  cuts: nObjs*8 records of
 * 5 uint32: leaves[4], meta = tt | value<<16 | nLeaves<<27 | used<<30. */
int orc_cuts(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
             const uint32_t *pos, uint32_t *cuts);

/* Per-node evaluation.  best: nObjs*4 int32 {slot or -1, subgraph index, gain, root level}. */
int orc_eval(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
             const uint32_t *pos, const uint32_t *cuts, int useZeros, int32_t *best);

/* One rw pass, this package's pipeline (strashed input -> strashed output).
 * cap: capacity of of0/of1.  stats[8] (may be NULL): candidates, selected,
 * new nodes, combined objs, depth, rounds(unused), estimated gain, 0. */
int orc_rw(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
           const uint32_t *pos, int useZeros, int cap, uint32_t *of0, uint32_t *of1,
           uint32_t *opos, int64_t *stats);

/* Signature bits of the cut prefilter, indexed by node id of the reordered
 * graph (NULL: id & 31, the reference's rule).  Models a batched pass, where
 * every node keeps the bit of its id in its own AIG. */
void orc_set_sig_bits(const uint8_t *bits);

/* The graph orc_rw hands to its final strash: the reordered input with the
 * selected roots substituted and the new nodes appended (ids n..N-1, fanins
 * possibly above their own id's level order but acyclic).  Returns N. */
int orc_rw_compose(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
                   const uint32_t *pos, int useZeros, int cap, uint32_t *of0, uint32_t *of1,
                   uint32_t *opos, int64_t *stats);

/* One rw pass following the reference's own replacement method (sequential
 * host choice pass) — used for QoR comparison only. */
int orc_rw_culs(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs,
                const uint32_t *pos, int useZeros, int cap, uint32_t *of0, uint32_t *of1,
                uint32_t *opos, int64_t *stats);

/* Equivalence by simulation: exhaustive when nPIs <= 16, else nWords*64 random
 * patterns from seed.  Returns -1 if all POs agree, else the first differing PO. */
int orc_sim_cmp(int nObjsA, const uint32_t *a0, const uint32_t *a1, const uint32_t *aPos,
                int nObjsB, const uint32_t *b0, const uint32_t *b1, const uint32_t *bPos,
                int nPIs, int nPOs, int nWords, uint64_t seed);

/* PO words over nWords*64 patterns with the pattern definition of
 * gls_simulate (exhaustive: pattern p gives PI j bit j of p; else splitmix64
 * words keyed by (j << 32 | w) under seed): out[po * nWords + w]. */
void orc_sim_po_words(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs, const uint32_t *pos,
                      int nWords, uint64_t seed, int exhaustive, uint64_t *out);

/* MFFC size of root within the given leaves, as counted by the evaluation. */
int orc_mffc(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1, int nPOs, const uint32_t *pos,
             int root, const int *leaves, int nL, int *tabSize);

/* Depth (max level) of an AIG with topological ids. */
int orc_depth(int nObjs, int nPIs, const uint32_t *f0, const uint32_t *f1);

#ifdef __cplusplus
}
#endif
