"""`st -c`: the host strash (AIGMan::strash(fCPU=true), src/aig/strash.cu:322-466).

The product's host strash runs without a GPU, so these tests run in the CPU
suite: product vs oracle bit for bit, plus properties of the reference's method
that the oracle's own formula does not restate (no trivial folding, dangling
removal, id order, delays) on hand-built graphs.
"""
import numpy as np
import pytest

import oracle_ffi as O


def same(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


def run_cpu(ctx, g):
    a = ctx.from_arrays(*g)
    a.strash(cpu=True)
    return a


GENS = [
    ("random_small", lambda c: c.gen_random(8, 300, 6, 40, 5)),
    ("random_20k", lambda c: c.gen_random(64, 20000, 32, 1000, 1)),
    ("mult8_red", lambda c: c.gen_multiplier(8, 0.1, 2)),
    ("mult16_red", lambda c: c.gen_multiplier(16, 0.2, 4)),
    ("geom_50k", lambda c: c.gen_geometric(128, 50000, 256, 5000, 0.01, 3)),
]


@pytest.mark.parametrize("name,gen", GENS, ids=[g[0] for g in GENS])
def test_cpu_strash_matches_oracle(host_ctx, name, gen):
    g = gen(host_ctx).arrays()
    a = run_cpu(host_ctx, g)
    want, want_lev, want_depth = O.strash_cpu(*g)
    assert same(a.arrays(), want)
    assert a.stats().levels == want_depth
    assert O.sim_cmp(g, a.arrays()) == -1


def test_cpu_strash_keeps_trivial_nodes_and_removes_duplicates(host_ctx):
    # ids: 0 const, 1..2 PIs a, b
    # 3 = a & a      (trivial: stays, the host strash does not fold)
    # 4 = a & b
    # 5 = b & a      (duplicate of 4 after ordering)
    # 6 = 3 & 5      -> (lit 6, lit 8) after mapping 5 -> 4
    # 7 = a & !a     (trivial: stays)
    # 8 = 4 & 4      dangling (no fanout), and with it 4 (its only fanout is 8)
    f0 = np.array([0, 0, 0, 2, 2, 4, 6, 2, 8], np.uint32)
    f1 = np.array([0, 0, 0, 2, 4, 2, 10, 3, 8], np.uint32)
    pos = np.array([12, 14, 5], np.uint32)
    a = run_cpu(host_ctx, (2, f0, f1, pos))
    nPIs, g0, g1, gp = a.arrays()
    # survivors in id order: 3 -> 3, 5 -> 4, 6 -> 5, 7 -> 6
    assert list(g0) == [0, 0, 0, 2, 2, 6, 2]
    assert list(g1) == [0, 0, 0, 2, 4, 8, 3]
    assert list(gp) == [10, 12, 5]
    assert a.stats().levels == 2          # delays: 3:1 4:1 5:2 6:1
    want, lev, depth = O.strash_cpu(2, f0, f1, pos)
    assert same((nPIs, g0, g1, gp), want) and depth == 2
    assert list(lev) == [0, 0, 0, 1, 1, 2, 1]


def test_cpu_strash_no_and_left_reports_depth_minus_one(host_ctx):
    f0 = np.array([0, 0, 0, 2], np.uint32)
    f1 = np.array([0, 0, 0, 4], np.uint32)
    a = run_cpu(host_ctx, (2, f0, f1, np.array([2, 5], np.uint32)))
    assert a.stats().nAnds == 0 and a.stats().levels == -1
    assert O.strash_cpu(2, f0, f1, np.array([2, 5], np.uint32))[2] == -1


def test_cpu_strash_is_idempotent_and_dangling_free(host_ctx):
    g = host_ctx.gen_random(32, 5000, 8, 300, 9).arrays()
    a = run_cpu(host_ctx, g)
    once = a.arrays()
    a.strash(cpu=True)
    assert same(a.arrays(), once)
    nPIs, f0, f1, pos = once
    ref = np.zeros(len(f0), np.int64)
    np.add.at(ref, f0[nPIs + 1:] >> 1, 1)
    np.add.at(ref, f1[nPIs + 1:] >> 1, 1)
    np.add.at(ref, pos >> 1, 1)
    assert (ref[nPIs + 1:] > 0).all()
    keys = (f0[nPIs + 1:].astype(np.uint64) << 32) | f1[nPIs + 1:]
    assert len(np.unique(keys)) == len(keys)
