"""Batched passes: many AIGs packed into one device arena, each result
bit-identical to the pass on that AIG alone (and to the oracle)."""
import sys

import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


def mixed(ctx, count, seed, max_ands=20000):
    """Seeded mix: random AIGs of log-uniform size, multipliers, a deep geometric
    one, one without ANDs and one whose POs are all constants or PIs."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = i % 5
        if kind == 3:
            out.append(ctx.gen_multiplier(int(rng.integers(3, 10)), 0.1, 100 + i))
        elif kind == 4:
            out.append(ctx.gen_geometric(32, int(rng.integers(500, 4000)), 16, 800, 0.02, 200 + i))
        else:
            n = int(np.exp(rng.uniform(np.log(200), np.log(max_ands))))
            pis = max(16, n // 200)
            out.append(ctx.gen_random(pis, n, max(1, pis // 2), 1000, 1000 + i))
    out.append(ctx.from_arrays(3, np.zeros(4, np.uint32), np.zeros(4, np.uint32), np.array([2, 5, 0, 1], np.uint32)))
    f0 = np.array([0, 0, 0, 2, 2], np.uint32)
    f1 = np.array([0, 0, 0, 3, 4], np.uint32)
    out.append(ctx.from_arrays(2, f0, f1, np.array([6, 1, 4], np.uint32)))   # node3 = a & !a folds to 0
    return out


def test_batch_strash_matches_alone_and_oracle(gpu_ctx):
    log("batch strash")
    aigs = mixed(gpu_ctx, 23, 1)
    inputs = [a.arrays() for a in aigs]
    gpu_ctx.batch_strash(aigs)
    for g, a in zip(inputs, aigs):
        alone = gpu_ctx.from_arrays(*g).strash()
        assert same(a.arrays(), alone.arrays())
        assert a.stats().levels == alone.stats().levels
        want, depth = O.strash(*g)
        assert same(a.arrays(), want) and a.stats().levels == depth


@pytest.mark.parametrize("zeros", [False, True])
def test_batch_rewrite_matches_alone_and_oracle(gpu_ctx, zeros):
    log("batch rw zeros=%s" % zeros)
    aigs = mixed(gpu_ctx, 31, 2)
    gpu_ctx.batch_strash(aigs)
    strashed = [a.arrays() for a in aigs]
    gpu_ctx.batch_rewrite(aigs, use_zeros=zeros)
    info = aigs[0].last_op()
    cand = 0
    for s, a in zip(strashed, aigs):
        alone = gpu_ctx.from_arrays(*s).strash()
        alone.rewrite(use_zeros=zeros)
        cand += alone.last_op()["candidates"]
        assert same(a.arrays(), alone.arrays()), "batched rw differs from rw alone"
        assert a.stats().levels == alone.stats().levels
        want, _ = O.rw(*s, use_zeros=zeros)
        assert same(a.arrays(), want)
    assert info["candidates"] == cand


def test_batch_script_rw_rwz_and_unstrashed_members(gpu_ctx):
    log("batch script")
    aigs = mixed(gpu_ctx, 12, 3)
    raw = [a.arrays() for a in aigs]
    gpu_ctx.batch_rewrite(aigs)            # members are strashed first
    gpu_ctx.batch_rewrite(aigs, use_zeros=True)
    for g, a in zip(raw, aigs):
        alone = gpu_ctx.from_arrays(*g).strash().rewrite().rewrite(use_zeros=True)
        assert same(a.arrays(), alone.arrays())
        assert O.sim_cmp(g, a.arrays()) == -1


def test_batch_rejects_repeated_member(gpu_ctx):
    import culs_amd
    a = gpu_ctx.gen_random(16, 300, 4, 100, 5)
    with pytest.raises(culs_amd.GlsError):
        gpu_ctx.batch_rewrite([a, a])
    assert gpu_ctx.batch_rewrite([]) == []


def test_batch_balance_matches_alone_and_oracle(gpu_ctx):
    log("batch balance")
    aigs = mixed(gpu_ctx, 17, 4)
    raw = [a.arrays() for a in aigs]
    gpu_ctx.batch_balance(aigs, update_level=False)
    for g, a in zip(raw, aigs):
        alone = gpu_ctx.from_arrays(*g).balance(update_level=False)
        assert same(a.arrays(), alone.arrays())
        assert same(a.arrays(), O.balance(*g))
        assert a.stats().levels == -1
    gpu_ctx.batch_strash(aigs)
    for g, a in zip(raw, aigs):
        assert same(a.arrays(), O.strash(*O.balance(*g))[0])
