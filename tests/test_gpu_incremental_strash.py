"""The incremental final strash of a rewrite (dev_strash_rw) against the wave
sweep over the composed graph (GLS_RW_FULL_STRASH) and the oracle: every
output array, the depth, and the levels the next pass reorders by (checked
through a second pass) must be identical; gls_aig_last_op reports which path
ran (final_strash 1 = incremental, no fallback) and how many level sweeps it
needed (strash_sweeps: the composed graph's levels are swept only where the
relaxed keys are not already the levels, the output's only where a merge or
fold lowered a representative).

Cases: the generators of the parity tests, a fanout-heavy AIG (a window of
20 nodes), the batched pass (tags carried to survivors, per-member depth),
and chains of passes (rw; rwz; rw) where the composed graphs merge and fold
nodes through several rounds.
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
    ("rand_small", lambda c: c.gen_random(16, 400, 8, 100, 11)),
    ("cfg0_random_20k", lambda c: c.gen_random(64, 20000, 32, 1000, 1)),
    ("mult8_red", lambda c: c.gen_multiplier(8, 0.1, 2)),
    ("cfg1_mult32_red", lambda c: c.gen_multiplier(32, 0.1, 2)),
    ("geom_50k", lambda c: c.gen_geometric(128, 50000, 256, 5000, 0.01, 3)),
    ("cfg2_shape_200k", lambda c: c.gen_geometric(1024, 200000, 4096, 50000, 0.001, 3)),
    ("narrow_window", lambda c: c.gen_random(8, 30000, 16, 20, 5)),
]


def same(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_incremental_equals_full_sweep(gpu_ctx, name, make):
    a = make(gpu_ctx)
    a.strash()
    b = a.copy()
    x = a.arrays()
    paths = set()
    for step, z in enumerate((False, True, False)):
        log("%s pass %d zeros=%s" % (name, step, z))
        a.rewrite(use_zeros=z)
        b.rewrite(use_zeros=z, full_strash=True)
        ia, ib = a.last_op(), b.last_op()
        log("incremental: final_strash %d rounds %d levels %d ands %d; sweep: final_strash %d levels %d ands %d"
            % (ia["final_strash"], ia["strash_rounds"], a.stats().levels, a.stats().nAnds, ib["final_strash"],
               b.stats().levels, b.stats().nAnds))
        ga, gb = a.arrays(), b.arrays()
        for k, (u, w) in enumerate(zip(ga[1:], gb[1:])):
            if len(u) != len(w) or not np.array_equal(u, w):
                d = np.nonzero(u[:min(len(u), len(w))] != w[:min(len(u), len(w))])[0]
                log("array %d differs: lengths %d / %d, first index %s" % (k, len(u), len(w), d[:5]))
        assert ia["final_strash"] == 1, ia
        paths.add(ia["strash_sweeps"])
        assert ib["final_strash"] == 0, ib
        assert same(ga, gb), "incremental strash differs from the wave sweep"
        assert a.stats().levels == b.stats().levels
        want, st = O.rw(*x, use_zeros=z)
        assert same(ga, want), "incremental strash differs from the oracle"
        assert a.stats().levels == st[4]
        x = ga
    log("%s: %d rounds in the last pass; level sweeps per pass %s" % (name, ia["strash_rounds"], sorted(paths)))


@pytest.mark.parametrize("count,window,path", [(12, 30, 1), (48, 1000, 0)], ids=["deep_incremental", "wide_sweep"])
def test_final_strash_in_batches(gpu_ctx, count, window, path):
    """Batched rw with tags: a deep narrow arena takes the incremental strash,
    a wide one (more than 1,024 objects per level) the wave sweep; either way
    every member equals its pass alone."""
    aigs = [gpu_ctx.gen_random(16 + i % 12, 1500 + 700 * (i % 12), 8, window, 100 + i) for i in range(count)]
    gpu_ctx.batch_strash(aigs)
    alone = [x.copy() for x in aigs]
    gpu_ctx.batch_rewrite(aigs)
    info = aigs[0].last_op()
    log("batch window %d: final_strash %d" % (window, info["final_strash"]))
    assert info["final_strash"] == path, info
    for x, y in zip(aigs, alone):
        y.rewrite(full_strash=True)
        assert same(x.arrays(), y.arrays())
        assert x.stats().levels == y.stats().levels
