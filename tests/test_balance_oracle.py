"""Balance (`b`) oracle: known answers and properties of the reference's
method (src/algorithms/balance.cu:666-1382) that its formula does not restate."""
import numpy as np
import pytest

import oracle_ffi as O


def chain_graph():
    # a, b, c, d = ids 1..4; n5 = a & b, n6 = n5 & c, n7 = n6 & d; PO = n7
    f0 = np.array([0, 0, 0, 0, 0, 2, 6, 8], np.uint32)
    f1 = np.array([0, 0, 0, 0, 0, 4, 10, 12], np.uint32)
    return 4, f0, f1, np.array([14], np.uint32)


def test_chain_becomes_a_balanced_tree():
    # cover of n7 = {d, c, a, b}, all level 0, ordered by literal descending
    # [d, c, b, a]: the last two (a, b) pair first, then (c, d), then both
    nPIs, f0, f1, pos = O.balance(*chain_graph())
    assert list(f0[5:]) == [2, 6, 10] and list(f1[5:]) == [4, 8, 12]
    assert list(pos) == [14]


def test_complementary_inputs_make_a_constant_cover():
    # n3 = a & b (single fanout, plain edge), n4 = n3 & !a: cover {a, b, !a} is FALSE
    f0 = np.array([0, 0, 0, 2, 3], np.uint32)
    f1 = np.array([0, 0, 0, 4, 6], np.uint32)
    g = O.balance(2, f0, f1, np.array([8, 9, 4], np.uint32))
    assert len(g[1]) == 3                 # no AND left
    assert list(g[3]) == [0, 1, 4]        # FALSE, TRUE, b


def test_shared_node_is_reused():
    # x = a & b drives a PO; y = (a & c) & b has the cover {b, a, c}, whose two
    # lowest inputs (a, b) are paired in the same step as x's cover: the two
    # requests share one new node (smallest id), so a & c is never rebuilt
    # ids: a1 b2 c3, x4 = a&b, n5 = a&c, y6 = n5&b
    f0 = np.array([0, 0, 0, 0, 2, 2, 4], np.uint32)
    f1 = np.array([0, 0, 0, 0, 4, 6, 10], np.uint32)
    pos = np.array([8, 12], np.uint32)
    nPIs, g0, g1, gp = O.balance(3, f0, f1, pos)
    assert len(g0) == 4 + 2               # x and x & c: a & c is not rebuilt
    assert O.sim_cmp((3, f0, f1, pos), (nPIs, g0, g1, gp)) == -1


def limit_chain(length, stack_side):
    """A single-fanout AND chain of `length` nodes.  stack_side: every sibling is
    a multi-fanout AND made just before (grows the DFS stack, hits the 128 limit);
    otherwise every sibling is a fresh PI (grows the cover, hits the 1024 limit)."""
    if stack_side:
        nPIs = 2 * length + 1
        f0, f1, pos = [0] * (nPIs + 1), [0] * (nPIs + 1), []
        prev = 2
        for k in range(length):
            m = len(f0)
            f0.append(2 * (2 + 2 * k)); f1.append(2 * (3 + 2 * k))
            pos.append(2 * m)
            n = len(f0)
            f0.append(min(prev, 2 * m)); f1.append(max(prev, 2 * m))
            prev = 2 * n
        pos.append(prev)
    else:
        nPIs = length + 1
        f0, f1, pos = [0] * (nPIs + 1), [0] * (nPIs + 1), []
        prev = 2
        for k in range(length):
            n = len(f0)
            x = 2 * (2 + k)
            f0.append(min(x, prev)); f1.append(max(x, prev))
            prev = 2 * n
        pos.append(prev)
    return nPIs, np.array(f0, np.uint32), np.array(f1, np.uint32), np.array(pos, np.uint32)


@pytest.mark.parametrize("stack_side,length", [(True, 300), (False, 1500)])
def test_cover_limits_keep_function(stack_side, length):
    g = limit_chain(length, stack_side)
    b = O.balance(*g)
    assert O.sim_cmp(g, b, words=4) == -1
    s, d = O.strash(*b)
    assert d <= O.depth(*g[:3])


GENS = [
    ("random_20k", lambda c: c.gen_random(64, 20000, 32, 1000, 1)),
    ("mult12_red", lambda c: c.gen_multiplier(12, 0.1, 2)),
    ("geom_20k", lambda c: c.gen_geometric(64, 20000, 64, 2000, 0.01, 3)),
]


@pytest.mark.parametrize("name,gen", GENS, ids=[g[0] for g in GENS])
def test_balance_keeps_function_and_does_not_deepen(host_ctx, name, gen):
    g = gen(host_ctx).arrays()
    s, d0 = O.strash(*g)
    b = O.balance(*s)
    assert O.sim_cmp(g, b) == -1
    # created nodes have their fanins below them
    nPIs, f0, f1, pos = b
    ids = np.arange(len(f0))
    assert ((f0[nPIs + 1:] >> 1) < ids[nPIs + 1:]).all() and ((f1[nPIs + 1:] >> 1) < ids[nPIs + 1:]).all()
    _, d1 = O.strash(*b)
    assert d1 <= d0
    # deterministic
    b2 = O.balance(*s)
    assert all(np.array_equal(x, y) for x, y in zip(b[1:], b2[1:]))


# ---- `b -s` (sortDecId = 0): equal levels in thrust's shuffle order --------

def golden_shuffles():
    """tests/golden/thrust_shuffle.txt: permutations written by rocThrust's
    thrust::shuffle(thrust::seq, ..., default_random_engine(seed)) through
    tests/golden/gen_thrust_shuffle.cpp ("len seed p0 p1 ...")."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "thrust_shuffle.txt")
    out = {}
    with open(path) as f:
        for line in f:
            v = [int(x) for x in line.split()]
            out[(v[0], v[1])] = v[2:]
    return out


def test_shuffle_restatement_matches_rocthrust():
    g = golden_shuffles()
    assert len(g) == 184
    for (n, seed), want in g.items():
        assert list(O.thrust_shuffle_perm(n, seed)) == want, (n, seed)


def test_shuffle_order_known_answer():
    # The chain's only cover (index 0 in its batch) gathers its inputs in
    # DFS order, fanin 0 first: [d, c, a, b] (lits 8, 6, 2, 4).  The shuffle
    # seeded with 0 is [2, 1, 0, 3] (golden line "4 0 2 1 0 3"), so the
    # inputs become [a, c, d, b], all on level 0, which the stable level sort
    # keeps.  Popping the last two each step: b & d first (id 5), then a & c
    # (id 6), then the two (id 7).
    assert golden_shuffles()[(4, 0)] == [2, 1, 0, 3]
    nPIs, f0, f1, pos = O.balance(*chain_graph(), sort_dec_id=0)
    assert list(f0[5:]) == [4, 2, 10] and list(f1[5:]) == [8, 6, 12]
    assert list(pos) == [14]
    # `b` pairs by literal instead: a & b, c & d
    nPIs, f0, f1, pos = O.balance(*chain_graph(), sort_dec_id=1)
    assert list(f0[5:]) == [2, 6, 10] and list(f1[5:]) == [4, 8, 12]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_shuffle_balance_keeps_function(host_ctx, seed):
    # random AIGs: covers of several inputs on one level, where the tie order
    # decides the pairing (multipliers' covers mostly have two inputs)
    g = host_ctx.gen_random(64, 5000, 32, 1000, seed).arrays()
    s, _ = O.strash(*g)
    a = O.balance(*s, sort_dec_id=0)
    assert O.sim_cmp(s, a) == -1
    # the tie order matters on such a graph: `b` and `b -s` differ
    b = O.balance(*s, sort_dec_id=1)
    assert not all(np.array_equal(x, y) for x, y in zip(a[1:], b[1:]))
