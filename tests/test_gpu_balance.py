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
