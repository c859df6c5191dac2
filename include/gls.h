/* gls.h — C ABI of the MI355X-native rw / strash library (libgls.so).
 *
 * Plain C types only.  Every entry point returns GLS_OK (0) or a negative
 * error code; gls_last_error() gives the message.  A context owns one HIP
 * device and one HIP stream; it is not thread-safe, use one context per host
 * thread.  AIG handles belong to the context that created them.
 *
 * Literal convention (lit_convention argument):
 *   GLS_LIT_AIGER: lit = 2*id + complement, lit 0 = constant FALSE (AIGER);
 *   GLS_LIT_CULS : same, but lit 0 = constant TRUE, lit 1 = FALSE — the
 *                  convention of the reference's AIGMan arrays
 *                  (src/common.h:285-286, AigConst1 = 0).
 * Arrays are indexed by object id: id 0 = constant, ids 1..nPIs = PIs,
 * ids nPIs+1 .. nObjs-1 = AND nodes with fanin ids smaller than their own
 * (topological), nObjs = 1 + nPIs + nAnds.  Entries of fanin arrays for the
 * constant and the PIs are ignored on input and written as 0 on output.
 */
#ifndef GLS_H
#define GLS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GLS_OK 0
#define GLS_ERR_ARG (-1)      /* invalid argument or malformed graph */
#define GLS_ERR_HIP (-2)      /* HIP runtime error */
#define GLS_ERR_NOMEM (-3)    /* device or host allocation failed */
#define GLS_ERR_FORMAT (-4)   /* AIGER parse error */
#define GLS_ERR_NODEV (-5)    /* no GPU: context is host-only */
#define GLS_ERR_IO (-6)       /* file open/read/write failed */
#define GLS_ERR_INTERNAL (-7) /* an internal consistency check failed */
#define GLS_ERR_TIMEOUT (-8)  /* gls_server_wait: the job is not done yet */

#define GLS_LIT_AIGER 0
#define GLS_LIT_CULS 1

typedef struct gls_ctx gls_ctx;
typedef struct gls_aig gls_aig;

typedef struct {
    int32_t nPIs, nPOs, nAnds, nObjs;
    int32_t levels;       /* depth, or -1 if not known */
    int32_t on_device;    /* 1 if the device copy is authoritative */
} gls_stats;

/* Timings of the last operation on a handle, in milliseconds (HIP events on
 * the context stream).  Fields not used by an operation are 0. */
typedef struct {
    double total_ms;
    double reorder_ms, cuts_ms, table_ms, eval_ms, select_ms, build_ms, strash_ms;
    int64_t candidates, selected, new_nodes, rounds, levels_enumerated;
    int64_t est_gain;
    /* Sampled launch timing of the cut kernel: every 64 levels a window of 8
     * consecutive launches, the first dispatch carrying a HIP start event and
     * the last a stop event; the windows' summed elapsed time, the launches
     * they cover and the nodes those launches processed. */
    double kernel_sample_ms;
    int64_t kernel_samples, kernel_sample_nodes;
    /* strash table slot inspections of the operation's strash sweeps: key
     * probes (compare-and-swap attempts on a slot) plus value updates through
     * a pre-inserted key's stored location */
    int64_t hash_probes;
    /* gls_rewrite: the cut schedule that ran (its GLS_RW_*_CUTS flag, 0 for
     * the default three-level launches; a requested persistent schedule that
     * could not run reports the schedule that replaced it), and the final
     * strash: 0 = the wave sweep over the composed graph, 1 = incremental
     * (only what the replacements changed is rehashed), 2 = incremental that
     * fell back to the wave sweep; strash_rounds = rounds of the incremental
     * strash */
    int32_t cut_schedule, final_strash;
    int64_t strash_rounds;
    /* gls_rewrite with GLS_RW_GRAPH_CUTS: host time of capturing the cut
     * launches into a HIP graph and instantiating it, and the kernel nodes
     * of that graph (cuts_ms covers capture, instantiation and execution) */
    double cut_graph_ms;
    int64_t cut_graph_nodes;
    /* gls_rewrite's incremental final strash: the one-workgroup level sweeps
     * it ran (0-2).  The composed graph's levels are swept only where its
     * relaxed topological keys are not already its levels, and the output's
     * only where a merge or fold put a node's representative on a lower
     * level than the node. */
    int32_t strash_sweeps, pad0;
} gls_op_info;

/* ---- context ---------------------------------------------------------- */
/* device >= 0 opens that HIP device; device < 0 gives a host-only context
 * (AIGER I/O, generators, array conversion; device operations then return
 * GLS_ERR_NODEV).  Replaces the process-wide state of AIGMan's constructor
 * (src/aig_manager.cu:21-33) and rw's __managed__ globals (rewrite.cu:24-27). */
int gls_ctx_create(int device, gls_ctx **out);
void gls_ctx_destroy(gls_ctx *ctx);
const char *gls_last_error(const gls_ctx *ctx);
/* 1 if built for and running on a GPU, else 0 */
int gls_ctx_has_device(const gls_ctx *ctx);

/* ---- AIG handles ------------------------------------------------------ */
/* Replaces AIGMan's host arrays + toDevice (src/aig_manager.cu:276-312). */
int gls_aig_from_arrays(gls_ctx *ctx, int32_t nPIs, int32_t nPOs, int32_t nAnds,
                        const uint32_t *fanin0, const uint32_t *fanin1, const uint32_t *pos,
                        int32_t lit_convention, gls_aig **out);
/* Replaces AIGMan::toHost (src/aig_manager.cu:314-336).  fanin0/fanin1 hold
 * nObjs entries, pos nPOs entries. */
int gls_aig_to_arrays(gls_aig *aig, uint32_t *fanin0, uint32_t *fanin1, uint32_t *pos,
                      int32_t lit_convention);
/* Replaces AIGMan::printStats counters (src/aig_manager.cu:604-612). */
int gls_aig_stats(gls_aig *aig, gls_stats *out);
int gls_aig_copy(gls_aig *src, gls_aig **out);
void gls_aig_free(gls_aig *aig);
/* Replaces AIGMan::readFile / readFromMemory (src/aig_manager.cu:80-229):
 * binary AIGER "aig M I L O A", L must be 0. */
int gls_aig_read_aiger(gls_ctx *ctx, const char *path, gls_aig **out);
int gls_aig_read_aiger_mem(gls_ctx *ctx, const char *buf, int64_t len, gls_aig **out);
/* Replaces AIGMan::saveFile (src/aig_manager.cu:231-267). */
int gls_aig_write_aiger(gls_aig *aig, const char *path);

/* ---- in-process networks (the ABC patch's gget) ------------------------ */
/* An AIG network as a caller holds it in memory (ABC's strashed Abc_Ntk_t in
 * the reference's ABC patch): objects 0 .. nObjs-1, each a constant-1 node, a
 * combinational input, an AND or a combinational output; ANDs and outputs
 * name their fanin objects with complement bits; ci / co list the inputs and
 * outputs in the network's own order.  Replaces AbcNtkToGpuMan
 * (src/abc_patch/abc_patch_transform.cpp:5-87): PI k is ci[k]; the ANDs are
 * numbered in depth-first post-order from the outputs (output order, fanin 0
 * before fanin 1, each node once; nodes no output reaches are left out), the
 * order Abc_NtkDfs gives them; the constant-1 node is the constant.  A fanin
 * that is not a constant, a listed input or an AND, or a cycle, gives
 * GLS_ERR_ARG; an object of kind GLS_NET_CI that is not in ci and that
 * nothing references is allowed (unused object ids).  The other direction (GpuManToAbcNtk,
 * abc_patch_transform.cpp:89-149) is gls_aig_to_arrays followed by one
 * Abc_AigAnd per AND in id order (INTEGRATION.md shows both). */
#define GLS_NET_CONST1 0
#define GLS_NET_CI 1
#define GLS_NET_AND 2
#define GLS_NET_CO 3
typedef struct {
    int32_t nObjs, nCis, nCos, pad;
    const uint8_t *kind;               /* nObjs entries, GLS_NET_* */
    const int32_t *fanin0, *fanin1;    /* nObjs entries: fanin objects of an AND; a CO uses fanin0 */
    const uint8_t *compl0, *compl1;    /* nObjs entries: complement bits of those edges */
    const int32_t *ci, *co;            /* the CI objects and the CO objects, in order */
} gls_net;
int gls_aig_from_net(gls_ctx *ctx, const gls_net *net, gls_aig **out);

/* ---- the hot path ----------------------------------------------------- */
/* Structural hashing + dangling-node removal on the device.  Replaces
 * Aig::strash (src/aig/strash.cuh:12-14, strash.cu:190-314) behind
 * AIGMan::strash(fCPU=false) (strash.cu:317).
 * flags = GLS_ST_CPU selects the host strash of `st -c`
 * (AIGMan::strash(fCPU=true), strash.cu:322-466): dangling removal and
 * id-order hashing without trivial-case folding, on the host (also in a
 * host-only context); levels = the largest node delay, -1 if no AND remains.
 * The device strash has two schedules with the same result: the wave sweep
 * (frontier by frontier) and, for graphs of at least 2^18 objects whose ids
 * are topological, one chunked level sweep over id order followed by hashing
 * rounds over all nodes at once (it gives way to the wave sweep when the
 * graph turns out wide or its ids are not topological).  GLS_ST_WAVES and
 * GLS_ST_LEVELS force either at any size (for measurement and testing). */
#define GLS_ST_CPU 1
#define GLS_ST_WAVES 2
#define GLS_ST_LEVELS 4
int gls_strash(gls_aig *aig, int32_t flags);

/* One DAG-aware rewriting pass with 4-input cuts (rw; use_zeros=1 is rw -z /
 * rwz).  Replaces AIGMan::rewrite (src/aig_manager.cu:632-672) and
 * rw::CPUSolver::Rewrite (src/algorithms/rewrite.cu:811-938).  The AIG stays on
 * the device; the result is strashed.  flags select how cut enumeration is
 * scheduled (the results are identical): 0 = one kernel launch per three
 * levels (a node recomputes the cut sets of its fanin cone inside the launch
 * in its own workgroup), GLS_RW_PAIR_CUTS = the same with two levels per
 * launch, GLS_RW_LEVEL_CUTS = one kernel launch per level,
 * GLS_RW_COOP_CUTS = one cooperative launch with a grid barrier per level,
 * GLS_RW_DATAFLOW_CUTS = persistent wavefronts that wait only for their fanins,
 * GLS_RW_XCD_CUTS = one persistent kernel on the CUs of one XCD with a barrier
 * per level (falls back to one launch per level where that cannot run). */
#define GLS_RW_COOP_CUTS 2
#define GLS_RW_DATAFLOW_CUTS 4
#define GLS_RW_XCD_CUTS 8
#define GLS_RW_LEVEL_CUTS 16
#define GLS_RW_PAIR_CUTS 32
/* GLS_RW_FULL_STRASH: finish with the wave-sweep strash of the composed graph
 * instead of the incremental one (the results are identical). */
#define GLS_RW_FULL_STRASH 64
/* GLS_RW_GRAPH_CUTS: with the default, pair or per-level schedule, capture the
 * pass's cut launches into a HIP graph and launch that once (for measuring
 * graph launch against stream launches; the results are identical). */
#define GLS_RW_GRAPH_CUTS 128
/* GLS_RW_PLAN_CUTS: cut_levels levels per launch from cone lists built once
 * per pass on the device (a cone's workgroup also stores the nodes it
 * recomputes, one owner each).  Currently DISABLED: refused with GLS_ERR_ARG,
 * because its kernel faulted the GPU in testing and the cause is not found. */
#define GLS_RW_PLAN_CUTS 256
/* GLS_RW_QUAD_CUTS: four levels per launch (k_cuts_quad: a node recomputes its
 * in-launch cone of up to 8 + 4 + 2 nodes in its workgroup), for measurement
 * against the default three. */
#define GLS_RW_QUAD_CUTS 512
int gls_rewrite(gls_aig *aig, int32_t use_zeros, int32_t flags);

/* Balance (`b`): rebuild every AND cover as a tree of least depth, sharing
 * existing nodes.  Replaces AIGMan::balance (aig_manager.cu:682-737) and
 * balancePerformV2 (src/algorithms/balance.cu:912-1382).  sort_dec_id 1
 * (`b`): a cover's inputs ordered by level, then literal; 0 (`b -s`): by
 * level, equal levels in the order of thrust::shuffle (thrust::seq,
 * default_random_engine seeded with the cover's index in its batch,
 * balance.cu:301, 404-411), thrust's algorithm as in rocThrust.
 * update_level = 1 strashes afterwards as the `b` command does
 * (command_manager.cpp:77); with 0 the result keeps its balance numbering
 * and its depth is unknown (levels = -1), as after `gb`. */
int gls_balance(gls_aig *aig, int32_t sort_dec_id, int32_t update_level);

/* Info about the last gls_strash / gls_rewrite on this handle. */
int gls_aig_last_op(gls_aig *aig, gls_op_info *out);

/* ---- batched passes ---------------------------------------------------- */
/* No counterpart in the reference, which runs one AIG per process through
 * AIGMan::strash / AIGMan::rewrite (aig_manager.cu:632-672).  The `count`
 * AIGs (distinct handles of ctx) are packed into one device arena — a
 * disjoint union with CSR offsets — rewritten by one pass and split back in
 * place.  Each AIG's result is bit-identical to gls_strash / gls_rewrite on it
 * alone.  gls_batch_rewrite strashes members that are not strashed first (as
 * gls_rewrite does); flags as in gls_rewrite.  gls_aig_last_op of every
 * member then reports the batch as a whole. */
int gls_batch_strash(gls_ctx *ctx, gls_aig **aigs, int32_t count, int32_t flags);
int gls_batch_rewrite(gls_ctx *ctx, gls_aig **aigs, int32_t count, int32_t use_zeros, int32_t flags);
/* Batched gls_balance: one device pass over the packed members; a member with
 * more than 32768 covers on one level is balanced alone (same result). */
int gls_batch_balance(gls_ctx *ctx, gls_aig **aigs, int32_t count, int32_t sort_dec_id, int32_t update_level);

/* ---- tuning ----------------------------------------------------------- */
/* No counterpart in the reference (its launch shapes and table sizes are
 * constants).  Settings that change only how a pass is scheduled or how much
 * room its tables have; every setting gives the same results.  A context
 * keeps one set per size bucket (floor(log2(objects)) of the graph or batch
 * arena a pass runs on), initially the defaults below.
 *   cut_wide: a group of levels (three, or cut_levels when planned) whose
 *     widest level has more nodes runs one cut launch per level (default 2048);
 *   rw_sweep_width: rw finishes with the incremental strash while the input
 *     has at most this many objects per level on average, else with the
 *     wave sweep (default 1024);
 *   strash_table_factor: the strash table has at least this many slots per
 *     object, a power of two, 4 or 8 (default 4);
 *   strash_wave_batch: wide strash waves launched per host check, 1-64
 *     (default 4);
 *   balance_narrow_max: balance levels of at most this many covers are
 *     rebuilt in one persistent workgroup, wider ones by grid-wide phase
 *     kernels (default 4096);
 *   cut_levels: levels per cut launch of the planned schedule
 *     (GLS_RW_PLAN_CUTS, currently disabled), 1-6 (default 3). */
typedef struct {
    int32_t cut_wide, rw_sweep_width, strash_table_factor, strash_wave_batch, balance_narrow_max, cut_levels;
} gls_tuning;
/* The settings a pass over `objects` objects uses. */
int gls_ctx_get_tuning(gls_ctx *ctx, int64_t objects, gls_tuning *out);
/* Sets them for that size bucket (objects < 0: every bucket); GLS_ERR_ARG
 * for a value out of range. */
int gls_ctx_set_tuning(gls_ctx *ctx, int64_t objects, const gls_tuning *t);
/* Times `command` (GLS_TUNE_*) on device copies of the AIGs (one AIG: the
 * single pass; several: the batched pass) under candidate settings, one
 * setting at a time from the current ones (coordinate descent, `reps` runs
 * each, the fastest kept), stores the best for the bucket of the inputs'
 * total objects and returns it with its time and the time of the settings it
 * started from.  The inputs are not changed. */
#define GLS_TUNE_RW 0
#define GLS_TUNE_RWZ 1
#define GLS_TUNE_STRASH 2
#define GLS_TUNE_BALANCE 3
int gls_autotune(gls_ctx *ctx, gls_aig **aigs, int32_t count, int32_t command, int32_t reps, gls_tuning *out,
                 double *best_ms, double *start_ms);

/* ---- serving ---------------------------------------------------------- */
/* No counterpart in the reference (one script, one AIG, one process).  A
 * server owns a worker thread and its own context on `device`.  Jobs (an AIG
 * as arrays + a script of st / strash, rw, rwz / rw -z, b / balance separated
 * by ';') are admitted in arrival order while their estimated device
 * footprint fits hbm_budget_bytes (and at most max_active run at once); an
 * admitted AIG stays on the device for its whole script.  Each round runs
 * every command kind once as a batched pass over the active jobs whose next
 * command it is, so jobs that arrive while others run join the next round.
 * A job's result equals running its script alone.  All functions are
 * thread-safe; gls_server_last_error's text is valid until the same thread's
 * next call to it. */
typedef struct gls_server gls_server;
typedef struct {
    int32_t status;                 /* 0 queued, 1 on the device, 2 done, < 0 error code */
    int32_t nPIs, nPOs, nAnds, levels;  /* the result */
    int32_t nAnds_in;
    int64_t footprint_bytes;        /* device bytes reserved while active */
    double submit_ms, admit_ms, done_ms;  /* milliseconds since the server started */
} gls_job_info;
int gls_server_create(int device, int64_t hbm_budget_bytes, int32_t max_active, gls_server **out);
int gls_server_submit(gls_server *s, int32_t nPIs, int32_t nPOs, int32_t nAnds, const uint32_t *fanin0,
                      const uint32_t *fanin1, const uint32_t *pos, int32_t lit_convention, const char *script,
                      int64_t *job);
/* Blocks until the job is done (timeout_ms < 0: no limit; GLS_ERR_TIMEOUT
 * when it runs out).  info may be NULL. */
int gls_server_wait(gls_server *s, int64_t job, double timeout_ms, gls_job_info *info);
/* The result of a done job: fanin arrays of info.nPIs + 1 + info.nAnds
 * entries, pos of info.nPOs. */
int gls_server_result(gls_server *s, int64_t job, uint32_t *fanin0, uint32_t *fanin1, uint32_t *pos,
                      int32_t lit_convention);
int gls_server_release(gls_server *s, int64_t job);
const char *gls_server_last_error(gls_server *s);
/* Finishes every submitted job, then stops the worker. */
void gls_server_destroy(gls_server *s);

/* ---- stage-level entry points (parity tests and tools) ----------------- */
/* Level-major renumbering of a strashed AIG as done before rw
 * (rw::CPUSolver::Reorder, rewrite.cu:1245-1279), applied in place.
 * level_out (nObjs entries, may be NULL) receives node levels. */
int gls_reorder(gls_aig *aig, uint32_t *level_out);
/* Cut sets of every node of a reordered AIG (GPUSolver::EnumerateAndPreEvaluate
 * cut stage, rewrite.cu:689-705).  cuts: nObjs*8 records of 5 uint32:
 * leaves[4], meta = truth | value<<16 | nLeaves<<27 | used<<30. */
int gls_enumerate_cuts(gls_aig *aig, uint32_t *cuts);
/* The same under a cut schedule of gls_rewrite's flags (GLS_RW_*_CUTS). */
int gls_enumerate_cuts_sched(gls_aig *aig, int32_t flags, uint32_t *cuts);
/* Per-node best replacement of a reordered AIG (EvaluateNode, rewrite.cu:378-447).
 * best: nObjs*4 int32 {cut slot or -1, subgraph index, gain, root level}. */
int gls_evaluate(gls_aig *aig, int32_t use_zeros, int32_t *best);

/* ---- equivalence checking by simulation -------------------------------- */
/* No counterpart in the reference's source: CULS ships its GPU
 * simulation-based CEC only as a binary (README.md:102-111).  Device
 * bit-parallel simulation of both AIGs over the same input patterns,
 * comparing PO i of a with PO i of b (same nPIs and nPOs).  With
 * nPIs <= GLS_SIM_EXHAUSTIVE_PIS every one of the 2^nPIs patterns is
 * simulated and `equivalent` is a proof; otherwise nWords * 64 seeded
 * patterns are, and only a difference is conclusive.
 * Patterns: exhaustive pattern p assigns PI j (0-based) bit j of p; random
 * word w of PI j is splitmix64's output function of
 * seed + ((j << 32 | w) + 1) * 0x9E3779B97F4A7C15.
 * po / pattern: the smallest differing PO index, and the smallest pattern on
 * which it differs (-1 when equivalent).  Both handles must belong to one
 * context. */
#define GLS_SIM_EXHAUSTIVE_PIS 16
typedef struct {
    int32_t equivalent;  /* 1: all POs agree on every simulated pattern */
    int32_t exhaustive;  /* 1: all 2^nPIs patterns were simulated */
    int32_t po;          /* smallest differing PO, -1 if none */
    int32_t pad;
    int64_t pattern;     /* smallest pattern on which `po` differs, -1 if none */
    int64_t patterns;    /* patterns simulated */
} gls_cec_info;
int gls_cec(gls_aig *a, gls_aig *b, int32_t nWords, uint64_t seed, gls_cec_info *out);
/* gls_cec without any allocation.  It has no counterpart in the reference;
 * it is for callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc.  The simulation
 * rows live in a static device buffer of GLS_CEC_NOALLOC_BYTES (one per GPU,
 * held until the process ends); both AIGs must already be on the device (any
 * device operation on a handle puts it there) and together hold at most
 * (GLS_CEC_NOALLOC_BYTES - 256) / 8 objects (nObjs a + nObjs b), else
 * GLS_ERR_ARG.  Results equal gls_cec's.  Calls on one GPU share the buffer:
 * calls that overlap on one GPU race. */
#define GLS_CEC_NOALLOC_BYTES 134217728
int gls_cec_noalloc(gls_aig *a, gls_aig *b, int32_t nWords, uint64_t seed, gls_cec_info *out);
/* PO values of one AIG over nWords * 64 patterns (exhaustive = 1: pattern p
 * assigns PI j bit j of p, for any nPIs; else the seeded random words above):
 * po_words[po * nWords + w], bit b of word w = pattern 64 w + b. */
int gls_simulate(gls_aig *aig, int32_t nWords, uint64_t seed, int32_t exhaustive, uint64_t *po_words);

/* ---- NPN tables and library (host) ------------------------------------ */
/* Closed-form NPN tables (the content read by rw::CPUSolver::ReadLibrary,
 * rewrite.cu:1281-1307): phase (bit 4 = output complement), perm index into
 * perms4 (24x4), class 0..221.  Each array 65536 bytes; perms4 96 bytes. */
int gls_npn_tables(uint8_t *phase, uint8_t *perm, uint8_t *cls, uint8_t *perms4);

/* ---- synthetic inputs (host) ------------------------------------------ */
/* Seeded random AIG: each AND takes two distinct fanins drawn uniformly from
 * the PIs and the previous `window` ANDs, complemented with p=0.5; POs are the
 * last nPOs ANDs. */
int gls_gen_random(gls_ctx *ctx, int32_t nPIs, int32_t nAnds, int32_t nPOs, int32_t window,
                   uint64_t seed, gls_aig **out);
/* Seeded deep random AIG: fanin distance ~ Geometric(p) over the previous
 * `window` ANDs (re-drawn uniformly over PIs and window when out of range). */
int gls_gen_geometric(gls_ctx *ctx, int32_t nPIs, int32_t nAnds, int32_t nPOs, int32_t window,
                      double p, uint64_t seed, gls_aig **out);
/* n x n unsigned array multiplier (XOR = 3 ANDs, MAJ = 3 ANDs) with a fraction
 * `redundant` of AND nodes re-expressed redundantly (seeded). */
int gls_gen_multiplier(gls_ctx *ctx, int32_t nbits, double redundant, uint64_t seed,
                       gls_aig **out);

#ifdef __cplusplus
}
#endif
#endif /* GLS_H */
