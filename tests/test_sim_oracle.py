"""Pins of the oracle's simulation (orc_sim_po_words), the checker of the
device simulation behind gls_simulate / gls_cec.

- Pattern words: the random words are splitmix64 outputs; for PI 0 the words
  0, 1, 2 are the generator's first three outputs from state 0 (published test
  values of splitmix64), the exhaustive words are the elementary truth tables
  and the all-0 / all-1 words of the higher PIs.
- AND simulation: every PO bit equals a scalar evaluation of the AIG under that
  one pattern (one assignment at a time, not bit-parallel), on random AIGs,
  with random and exhaustive patterns.
"""
import numpy as np

import oracle_ffi as O

SPLITMIX64_FROM_ZERO = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def one_pi_graph(nPIs, po_lits):
    n = nPIs + 1
    return (nPIs, np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.array(po_lits, np.uint32))


def test_random_words_are_splitmix64():
    g = one_pi_graph(1, [2, 3])
    w = O.sim_po_words(g, 3, seed=0)
    assert [int(x) for x in w[0]] == SPLITMIX64_FROM_ZERO
    assert [int(x) ^ 0xFFFFFFFFFFFFFFFF for x in w[1]] == SPLITMIX64_FROM_ZERO
    # another seed shifts the state: seed = gamma gives the sequence one step later
    w2 = O.sim_po_words(g, 2, seed=0x9E3779B97F4A7C15)
    assert [int(x) for x in w2[0]] == SPLITMIX64_FROM_ZERO[1:3]


def test_exhaustive_words_are_elementary():
    nPIs = 9
    g = one_pi_graph(nPIs, [2 * (j + 1) for j in range(nPIs)])
    words = 1 << (nPIs - 6)
    w = O.sim_po_words(g, words, exhaustive=True)
    el = [0xAAAAAAAAAAAAAAAA, 0xCCCCCCCCCCCCCCCC, 0xF0F0F0F0F0F0F0F0, 0xFF00FF00FF00FF00,
          0xFFFF0000FFFF0000, 0xFFFFFFFF00000000]
    for j in range(nPIs):
        for k in range(words):
            want = el[j] if j < 6 else (0xFFFFFFFFFFFFFFFF if (k >> (j - 6)) & 1 else 0)
            assert int(w[j, k]) == want, (j, k)


def scalar_eval(g, assign):
    """Value of every PO under one assignment (list of PI bits), node by node."""
    nPIs, f0, f1, pos = g
    val = [0] * len(f0)
    for j in range(nPIs):
        val[j + 1] = assign[j]
    for v in range(nPIs + 1, len(f0)):
        a = val[int(f0[v]) >> 1] ^ (int(f0[v]) & 1)
        b = val[int(f1[v]) >> 1] ^ (int(f1[v]) & 1)
        val[v] = a & b
    return [val[int(p) >> 1] ^ (int(p) & 1) for p in pos]


def check_scalar(g, words, seed, exhaustive):
    nPIs = g[0]
    w = O.sim_po_words(g, words, seed=seed, exhaustive=exhaustive)
    pis = O.sim_po_words((nPIs, np.zeros(nPIs + 1, np.uint32), np.zeros(nPIs + 1, np.uint32),
                          np.array([2 * (j + 1) for j in range(nPIs)], np.uint32)), words, seed=seed,
                         exhaustive=exhaustive)
    rng = np.random.default_rng(seed)
    for p in rng.choice(words * 64, size=min(words * 64, 96), replace=False):
        k, b = divmod(int(p), 64)
        assign = [(int(pis[j, k]) >> b) & 1 for j in range(nPIs)]
        if exhaustive:
            assert assign == [(int(p) >> j) & 1 for j in range(nPIs)]
        want = scalar_eval(g, assign)
        got = [(int(w[o, k]) >> b) & 1 for o in range(len(g[3]))]
        assert got == want, (p, got, want)


def test_simulation_matches_scalar_evaluation(host_ctx):
    for seed in (1, 2, 3):
        g = host_ctx.gen_random(12, 300, 9, 100, seed).arrays()
        check_scalar(g, 2, seed, exhaustive=False)
        check_scalar(g, 4, seed, exhaustive=True)
    m = host_ctx.gen_multiplier(4, 0.1, 5).arrays()
    check_scalar(m, 1, 7, exhaustive=True)
    check_scalar(m, 3, 7, exhaustive=False)


def test_sim_cmp_agrees_with_po_words(host_ctx):
    """orc_sim_cmp (the older checker) and orc_sim_po_words agree on exhaustive verdicts."""
    g = host_ctx.gen_multiplier(6, 0.1, 3).arrays()
    s, _ = O.strash(*g)
    assert O.sim_cmp(g, s) == -1
    a, b = O.sim_po_words(g, 64, exhaustive=True), O.sim_po_words(s, 64, exhaustive=True)
    assert np.array_equal(a, b)
    bad = (s[0], s[1], s[2], s[3].copy())
    bad[3][5] ^= 1
    assert O.sim_cmp(g, bad) == 5
    c = O.sim_po_words(bad, 64, exhaustive=True)
    assert not np.array_equal(a[5], c[5]) and np.array_equal(np.delete(a, 5, 0), np.delete(c, 5, 0))
