"""GPU balance vs the oracle, bit for bit, and `b` (balance + strash) end to end."""
import sys

import numpy as np
import pytest

import oracle_ffi as O
from test_balance_oracle import limit_chain

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


CASES = [
    ("rand_small", lambda c: c.gen_random(16, 400, 8, 100, 11)),
    ("cfg0_random_20k", lambda c: c.gen_random(64, 20000, 32, 1000, 1)),
    ("mult8_red", lambda c: c.gen_multiplier(8, 0.1, 2)),
    ("cfg1_mult32_red", lambda c: c.gen_multiplier(32, 0.1, 2)),
    ("geom_50k", lambda c: c.gen_geometric(128, 50000, 256, 5000, 0.01, 3)),
    ("wide_random_100k", lambda c: c.gen_random(2048, 100000, 1024, 50000, 4)),
]


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_balance_parity(gpu_ctx, name, gen):
    log("balance parity: %s" % name)
    a = gen(gpu_ctx)
    a.strash()
    s = a.arrays()
    a.balance(update_level=False)
    want = O.balance(*s)
    assert same(a.arrays(), want), "GPU balance differs from the oracle"
    assert a.stats().levels == -1
    a.strash()
    want_s, depth = O.strash(*want)
    assert same(a.arrays(), want_s) and a.stats().levels == depth


@pytest.mark.parametrize("stack_side,length", [(True, 300), (False, 1500)])
def test_balance_cover_limits(gpu_ctx, stack_side, length):
    log("balance limits stack_side=%s" % stack_side)
    g = limit_chain(length, stack_side)
    a = gpu_ctx.from_arrays(*g)
    a.balance(update_level=False)
    assert same(a.arrays(), O.balance(*g))


def test_balance_on_unstrashed_input_and_script(gpu_ctx):
    """`b; rw; rwz; b` on a raw (not strashed) AIG, against the oracle chain."""
    log("b; rw; rwz; b")
    a = gpu_ctx.gen_multiplier(10, 0.2, 7)
    g = a.arrays()
    a.balance()
    want = O.strash(*O.balance(*g))[0]
    assert same(a.arrays(), want)
    a.rewrite().rewrite(use_zeros=True).balance()
    w, _ = O.rw(*want)
    w, _ = O.rw(*w, use_zeros=True)
    w = O.strash(*O.balance(*w))[0]
    assert same(a.arrays(), w)
    assert O.sim_cmp(g, a.arrays()) == -1


def wide_level(n_covers, nPIs, seed):
    """n_covers POs, each driving AND(AND(a, b), c) of distinct PIs: the PO
    level holds n_covers covers of three inputs."""
    rng = np.random.default_rng(seed)
    f0, f1, pos, seen = [0] * (nPIs + 1), [0] * (nPIs + 1), [], set()
    while len(pos) < n_covers:
        a, b, c = (int(x) for x in rng.choice(np.arange(1, nPIs + 1), 3, replace=False))
        if (min(a, b), max(a, b), c) in seen:
            continue
        seen.add((min(a, b), max(a, b), c))
        inner = len(f0)
        f0.append(2 * a)
        f1.append(2 * b)
        top = len(f0)
        f0.append(2 * inner)
        f1.append(2 * c + int(rng.integers(2)))
        pos.append(2 * top)
    return nPIs, np.array(f0, np.uint32), np.array(f1, np.uint32), np.array(pos, np.uint32)


@pytest.mark.parametrize("narrow_max", [4096, 65536])
def test_balance_level_wider_than_a_batch(gpu_ctx, narrow_max):
    """A level of 40,000 covers is rebuilt in batches of 32,768 covers, as the
    reference does, whatever balance_narrow_max allows the persistent
    workgroup (a level wider than a batch never goes there)."""
    log("balance level wider than a batch, narrow max %d" % narrow_max)
    g = wide_level(40000, 600, 11)
    a = gpu_ctx.from_arrays(*g).strash()
    s = a.arrays()
    objects = len(s[1])
    old = gpu_ctx.get_tuning(objects)["balance_narrow_max"]
    try:
        gpu_ctx.set_tuning(objects, balance_narrow_max=narrow_max)
        a.balance(update_level=False)
    finally:
        gpu_ctx.set_tuning(objects, balance_narrow_max=old)
    assert same(a.arrays(), O.balance(*s)), "GPU balance differs from the oracle"


SHUFFLE_CASES = [c for c in CASES if not c[0].startswith("mult")]


@pytest.mark.parametrize("name,gen", SHUFFLE_CASES, ids=[c[0] for c in SHUFFLE_CASES])
def test_balance_shuffle_parity(gpu_ctx, name, gen):
    """`b -s`: equal-level inputs in thrust's shuffle order, seeded with the
    cover's index in its batch (balance.cu:301, 404-411), bit-exact against
    the oracle's restatement (itself checked against rocThrust's shuffle)."""
    log("balance -s parity: %s" % name)
    a = gen(gpu_ctx)
    a.strash()
    s = a.arrays()
    a.balance(update_level=False, shuffle=True)
    assert same(a.arrays(), O.balance(*s, sort_dec_id=0)), "GPU b -s differs from the oracle"
    assert O.sim_cmp(s, a.arrays()) == -1


def test_balance_shuffle_batches_and_limits(gpu_ctx):
    """`b -s` where the seeds restart: a level of 40,000 covers (two batches),
    and covers cut at the stack and size limits."""
    log("balance -s wide level and limits")
    a = gpu_ctx.from_arrays(*wide_level(40000, 600, 12)).strash()
    s = a.arrays()
    a.balance(update_level=False, shuffle=True)
    assert same(a.arrays(), O.balance(*s, sort_dec_id=0))
    for stack_side, length in [(True, 300), (False, 1500)]:
        g = limit_chain(length, stack_side)
        b = gpu_ctx.from_arrays(*g)
        b.balance(update_level=False, shuffle=True)
        assert same(b.arrays(), O.balance(*g, sort_dec_id=0))


def test_balance_shuffle_not_batched(gpu_ctx):
    """The batched balance refuses `b -s` (a cover's index in a batched arena
    is not its index when balanced alone)."""
    import culs_amd
    a = gpu_ctx.gen_random(32, 500, 16, 100, 1)
    with pytest.raises(culs_amd.GlsError):
        gpu_ctx.batch_balance([a], shuffle=True)
