"""Device simulation (gls_simulate) and simulation-based equivalence checking
(gls_cec) against the oracle's simulation (orc_sim_po_words).

- gls_simulate: PO words bit-exact with the oracle, random and exhaustive
  patterns, on level-major (strashed) and generator-order AIGs, graphs whose
  AND count sits around the 1024-id chunk of the sweep, deep in-chunk chains
  (a ripple multiplier in creation order), more PIs than one chunk.
- gls_cec: the verdict and the reported (po, pattern) equal the smallest
  differing pair computed from the oracle's PO words; equivalent results of
  st / rw / rwz / b are accepted (exhaustively for <= 16 PIs), planted changes
  are caught, interface mismatches are refused.
"""
import sys

import numpy as np
import pytest

import culs_amd
import oracle_ffi as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


GRAPHS = [
    ("random_20k", lambda c: c.gen_random(64, 20000, 32, 1000, 1)),
    ("random_1023", lambda c: c.gen_random(16, 1023 - 17, 8, 100, 2)),   # nObjs = 1023
    ("random_1024", lambda c: c.gen_random(16, 1024 - 17, 8, 100, 3)),
    ("random_2049", lambda c: c.gen_random(16, 2049 - 17, 8, 100, 4)),
    ("mult8_raw", lambda c: c.gen_multiplier(8, 0.1, 2)),                 # deep chains inside a chunk
    ("mult16", lambda c: c.gen_multiplier(16, 0.0, 1)),
    ("geom_50k", lambda c: c.gen_geometric(1500, 50000, 256, 5000, 0.01, 5)),   # PIs span two chunks
    ("one_and", lambda c: c.from_arrays(2, np.array([0, 0, 0, 2], np.uint32), np.array([0, 0, 0, 5], np.uint32),
                                        np.array([6, 7, 1, 0, 3], np.uint32))),
]


@pytest.mark.parametrize("name,make", GRAPHS, ids=[g[0] for g in GRAPHS])
@pytest.mark.parametrize("exhaustive", [False, True])
def test_simulate_matches_oracle(gpu_ctx, name, make, exhaustive):
    a = make(gpu_ctx)
    g = a.arrays()
    words = 3 if not exhaustive else max(1, 1 << max(0, min(g[0], 12) - 6))
    log("simulate %s exhaustive=%s words=%d" % (name, exhaustive, words))
    got = a.simulate(words, seed=7, exhaustive=exhaustive)
    want = O.sim_po_words(g, words, seed=7, exhaustive=exhaustive)
    assert np.array_equal(got, want)
    # the same on the strashed (level-major) graph
    a.strash()
    got = a.simulate(words, seed=7, exhaustive=exhaustive)
    want = O.sim_po_words(a.arrays(), words, seed=7, exhaustive=exhaustive)
    assert np.array_equal(got, want)


def expected_first(ga, gb, words, seed, exhaustive):
    wa = O.sim_po_words(ga, words, seed=seed, exhaustive=exhaustive)
    wb = O.sim_po_words(gb, words, seed=seed, exhaustive=exhaustive)
    nPIs = ga[0]
    d = wa ^ wb
    if exhaustive and nPIs < 6:
        d &= np.uint64((1 << (1 << nPIs)) - 1)
    for po in range(d.shape[0]):
        nz = np.nonzero(d[po])[0]
        if len(nz):
            k = int(nz[0])
            x = int(d[po, k])
            return po, k * 64 + ((x & -x).bit_length() - 1)
    return -1, -1


def check_cec(ctx, ga, gb, words=8, seed=3):
    a, b = ctx.from_arrays(*ga), ctx.from_arrays(*gb)
    info = a.cec(b, words=words, seed=seed)
    exh = ga[0] <= culs_amd.GLS_SIM_EXHAUSTIVE_PIS
    w = max(1, 1 << max(0, ga[0] - 6)) if exh else words
    po, pat = expected_first(ga, gb, w, seed, exh)
    assert info["exhaustive"] == (1 if exh else 0)
    assert info["patterns"] == ((1 << ga[0]) if exh else words * 64)
    assert (info["po"], info["pattern"]) == (po, pat), (info, po, pat)
    assert info["equivalent"] == (1 if po < 0 else 0)
    return info


def mutate(g, which, node=None):
    nPIs, f0, f1, pos = g
    f0, f1, pos = f0.copy(), f1.copy(), pos.copy()
    if which == "po":
        pos[len(pos) // 2] ^= 1
    else:
        v = node if node is not None else (nPIs + 1 + (len(f0) - nPIs - 1) * 2 // 3)
        f0[v] ^= 1
    return nPIs, f0, f1, pos


@pytest.mark.parametrize("name,make", GRAPHS[:7], ids=[g[0] for g in GRAPHS[:7]])
def test_cec_results_of_the_passes(gpu_ctx, name, make):
    a = make(gpu_ctx)
    g = a.arrays()
    log("cec %s" % name)
    a.strash()
    s = a.arrays()
    assert check_cec(gpu_ctx, g, s)["equivalent"] == 1
    a.rewrite()
    a.rewrite(use_zeros=True)
    r = a.arrays()
    assert check_cec(gpu_ctx, g, r)["equivalent"] == 1
    a.balance()
    b = a.arrays()
    assert check_cec(gpu_ctx, g, b)["equivalent"] == 1
    # planted changes
    check_cec(gpu_ctx, g, mutate(r, "po"))
    assert check_cec(gpu_ctx, g, mutate(r, "po"))["equivalent"] == 0
    check_cec(gpu_ctx, g, mutate(r, "fanin"))


def test_cec_refuses_other_interfaces(gpu_ctx):
    a = gpu_ctx.gen_random(16, 500, 8, 100, 1)
    b = gpu_ctx.gen_random(17, 500, 8, 100, 1)
    with pytest.raises(culs_amd.GlsError):
        a.cec(b)
    c = gpu_ctx.gen_random(16, 500, 9, 100, 1)
    with pytest.raises(culs_amd.GlsError):
        a.cec(c)


def test_cec_10m_rw_output(gpu_ctx):
    """The headline workload: the rw output of the 10M-AND AIG against its input, 4,096 patterns."""
    a = gpu_ctx.gen_geometric(1024, 10_000_000, 4096, 50000, 0.001, 3)
    g = a.copy()
    a.strash()
    a.rewrite()
    info = g.cec(a, words=64, seed=5)
    log("10M cec: %s" % info)
    assert info["equivalent"] == 1 and info["patterns"] == 4096
    # a complemented fanin of a PO's driver is visible on about half of the patterns
    r = a.arrays()
    drv = int(r[3][100]) >> 1
    bad = gpu_ctx.from_arrays(*mutate(r, "fanin", node=drv))
    info = g.cec(bad, words=64, seed=5)
    log("10M cec with a planted change: %s" % info)
    assert info["equivalent"] == 0 and info["po"] <= 100
    want = expected_first(g.arrays(), bad.arrays(), 64, 5, False)
    assert (info["po"], info["pattern"]) == want
