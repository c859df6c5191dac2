"""GPU parity: every stage of the HIP path against the CPU oracle, bit for bit.

Inputs are the seeded generators at sizes the oracle finishes in seconds, plus
size-independent properties (equivalence by simulation, determinism, node-count
monotonicity) at the larger sizes.
"""
import sys

import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    ("rand_small", dict(kind="random", nPIs=16, nAnds=400, nPOs=8, window=100, seed=11)),
    ("cfg0_random_20k", dict(kind="random", nPIs=64, nAnds=20000, nPOs=32, window=1000, seed=1)),
    ("mult8_red", dict(kind="mult", nbits=8, redundant=0.1, seed=2)),
    ("cfg1_mult32_red", dict(kind="mult", nbits=32, redundant=0.1, seed=2)),
    ("geom_50k", dict(kind="geom", nPIs=128, nAnds=50000, nPOs=256, window=5000, p=0.01, seed=3)),
    # the 10M-AND workload's generator and parameters (1,024 PIs, 4,096 POs,
    # window 50,000, p = 0.001, seed 3) at 200k ANDs
    ("cfg2_shape_200k", dict(kind="geom", nPIs=1024, nAnds=200000, nPOs=4096, window=50000, p=0.001, seed=3)),
]


def make(ctx, spec):
    if spec["kind"] == "random":
        return ctx.gen_random(spec["nPIs"], spec["nAnds"], spec["nPOs"], spec["window"], spec["seed"])
    if spec["kind"] == "mult":
        return ctx.gen_multiplier(spec["nbits"], spec["redundant"], spec["seed"])
    return ctx.gen_geometric(spec["nPIs"], spec["nAnds"], spec["nPOs"], spec["window"], spec["p"], spec["seed"])


def same(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


@pytest.mark.parametrize("name,spec", CASES, ids=[c[0] for c in CASES])
def test_strash_parity(gpu_ctx, name, spec):
    """Both device schedules (the wave sweep, and the level sweep with hashing
    rounds over all nodes) and the size-based default give the oracle's AIG
    and depth."""
    import culs_amd
    a = make(gpu_ctx, spec)
    g = a.arrays()
    want, depth = O.strash(*g)
    for sched in (0, culs_amd.GLS_ST_WAVES, culs_amd.GLS_ST_LEVELS):
        log("strash parity: %s, schedule %d" % (name, sched))
        b = a.copy()
        b.strash(schedule=sched)
        got = b.arrays()
        assert same(got, want), "GPU strash differs from the oracle (schedule %d)" % sched
        assert b.stats().levels == depth, "depth under schedule %d" % sched
    assert O.sim_cmp(g, got) == -1


@pytest.mark.parametrize("name,spec", CASES, ids=[c[0] for c in CASES])
def test_reorder_cuts_eval_parity(gpu_ctx, name, spec):
    log("reorder/cuts/eval parity: %s" % name)
    a = make(gpu_ctx, spec)
    a.strash()
    s = a.arrays()
    lev = a.reorder()
    r = a.arrays()
    want, want_lev = O.reorder(*s)
    assert same(r, want), "GPU reorder differs"
    assert np.array_equal(lev, want_lev)
    cuts = a.enumerate_cuts()
    want_cuts = O.cuts(*r)
    assert np.array_equal(cuts, want_cuts), "GPU cut sets differ (first node %d)" % (
        int(np.argmax((cuts != want_cuts).reshape(-1, 40).any(axis=1))))
    for z in (False, True):
        best = a.evaluate(use_zeros=z)
        want_best = O.evaluate(*r, want_cuts, use_zeros=z)
        assert np.array_equal(best, want_best), "GPU evaluation differs (zeros=%s)" % z


@pytest.mark.parametrize("name,spec", CASES, ids=[c[0] for c in CASES])
def test_rewrite_parity(gpu_ctx, name, spec):
    log("rw parity: %s" % name)
    a = make(gpu_ctx, spec)
    g = a.arrays()
    a.strash()
    s = a.arrays()
    for z in (False, True):
        a.rewrite(use_zeros=z)
        got = a.arrays()
        want, st = O.rw(*s, use_zeros=z)
        info = a.last_op()
        assert info["candidates"] == st[0] and info["selected"] == st[1] and info["new_nodes"] == st[2]
        assert same(got, want), "GPU rw (zeros=%s) differs from the oracle" % z
        # the selected replacements are disjoint, so the pass removes at least
        # the sum of their gains (the final strash can only remove more)
        assert info["est_gain"] == st[6]
        assert O.n_ands(s) - O.n_ands(got) >= info["est_gain"]
        assert O.sim_cmp(g, got) == -1, "rw output not equivalent to the input"
        s = got


@pytest.mark.parametrize("name,spec", CASES[1:], ids=[c[0] for c in CASES[1:]])
def test_cut_launch_modes_agree(gpu_ctx, name, spec):
    """Three and two levels per launch, one launch per level, one
    cooperative launch, the dataflow and the XCD-local schedules agree."""
    import culs_amd
    log("cut launch modes: %s" % name)
    a = make(gpu_ctx, spec)
    a.strash()
    b, c, d, e, f = a.copy(), a.copy(), a.copy(), a.copy(), a.copy()
    a.rewrite(cut_mode=0)
    b.rewrite(cut_mode=culs_amd.GLS_RW_DATAFLOW_CUTS)
    c.rewrite(cut_mode=culs_amd.GLS_RW_COOP_CUTS)
    d.rewrite(cut_mode=culs_amd.GLS_RW_XCD_CUTS)
    e.rewrite(cut_mode=culs_amd.GLS_RW_LEVEL_CUTS)
    f.rewrite(cut_mode=culs_amd.GLS_RW_PAIR_CUTS)
    for x in (b, c, d, e, f):
        assert same(a.arrays(), x.arrays())
    # the schedule that ran (gls_op_info.cut_schedule): every one as requested;
    # the XCD-local kernel recomputes with the default schedule when its
    # workgroups were not all placed on one XCD, and says so
    ran = [x.last_op()["cut_schedule"] for x in (a, b, c, d, e, f)]
    log("cut schedules that ran: %s" % ran)
    assert ran[0] == 0 and ran[1] == culs_amd.GLS_RW_DATAFLOW_CUTS and ran[2] == culs_amd.GLS_RW_COOP_CUTS
    assert ran[3] in (culs_amd.GLS_RW_XCD_CUTS, 0)
    assert ran[4] == culs_amd.GLS_RW_LEVEL_CUTS and ran[5] == culs_amd.GLS_RW_PAIR_CUTS


@pytest.mark.parametrize("name,spec", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_rewrite_after_host_strash(gpu_ctx, name, spec):
    """`st -c; rw`: rw on the host strash's output (trivial nodes kept)."""
    log("st -c; rw: %s" % name)
    a = make(gpu_ctx, spec)
    g = a.arrays()
    a.strash(cpu=True)
    s = a.arrays()
    want_s, _, depth = O.strash_cpu(*g)
    assert same(s, want_s) and a.stats().levels == depth
    a.rewrite()
    want, st = O.rw(*s)
    assert same(a.arrays(), want)
    assert O.sim_cmp(g, a.arrays()) == -1


def test_rewrite_deterministic(gpu_ctx):
    log("rw determinism")
    outs = []
    for _ in range(3):
        a = gpu_ctx.gen_random(64, 20000, 32, 1000, 7)
        a.strash().rewrite().rewrite(use_zeros=True)
        outs.append(a.arrays())
    assert same(outs[0], outs[1]) and same(outs[0], outs[2])


def test_empty_and_trivial_graphs(gpu_ctx):
    log("edge cases")
    # no AND nodes
    a = gpu_ctx.from_arrays(3, np.zeros(4, np.uint32), np.zeros(4, np.uint32), np.array([2, 5, 0, 1], np.uint32))
    a.strash().rewrite()
    nPIs, f0, f1, pos = a.arrays()
    assert len(f0) == 4 and list(pos) == [2, 5, 0, 1]
    # ANDs that all fold to constants / literals
    f0 = np.array([0, 0, 0, 2, 2, 6, 0], np.uint32)
    f1 = np.array([0, 0, 0, 3, 4, 1, 7], np.uint32)
    # node3 = a & !a = 0, node4 = a & b, node5 = node3 & 1 = 0, node6 = 0 & ... = 0
    pos = np.array([6, 8, 10, 13], np.uint32)
    g = (2, f0, f1, pos)
    b = gpu_ctx.from_arrays(*g)
    b.strash()
    got = b.arrays()
    want, _ = O.strash(*g)
    assert same(got, want)
    b.rewrite()
    assert O.sim_cmp(g, b.arrays()) == -1


@pytest.mark.slow
def test_large_geometric_equivalence(gpu_ctx):
    """Config-2 generator at 1M ANDs: size-independent checks only."""
    log("large geometric 1M")
    a = gpu_ctx.gen_geometric(1024, 1_000_000, 4096, 50000, 0.001, 3)
    g = a.arrays()
    a.strash()
    s = a.arrays()
    n0 = a.n_ands
    a.rewrite()
    r = a.arrays()
    assert a.n_ands <= n0
    assert O.sim_cmp(g, r, words=8) == -1
    b = gpu_ctx.from_arrays(*s)
    b.strash().rewrite()
    assert same(b.arrays(), r), "rw not deterministic at 1M nodes"
