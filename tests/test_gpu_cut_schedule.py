"""Cut-enumeration schedules on the GPU, checked against the oracle bit for bit.

The default schedule covers three levels per launch (a node recomputes its
in-launch fanin cone in its workgroup) and falls back to one launch per level
for any group of three levels holding a level wider than 2048 nodes.  These
layered AIGs place wide and narrow levels, widths on both sides of that limit
and long runs of one-node levels at every offset within a group, so groups of
every shape occur: the cut records must equal the oracle's, and rw must give
the same AIG under every schedule, launched from the stream or from one
captured HIP graph, and equal the oracle's rw.
"""
import sys

import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def layered(seed, nPIs, widths, near=0.5):
    """Layer i's nodes take fanin 0 from layer i-1 (so each layer is one level)
    and fanin 1 from layer i-1 with probability `near` (both fanins inside a
    launch), else from any earlier node or PI.  Every AND drives a PO."""
    rng = np.random.default_rng(seed)
    f0, f1 = [0] * (nPIs + 1), [0] * (nPIs + 1)
    prev = np.arange(1, nPIs + 1)
    first = nPIs + 1
    for w in widths:
        base = len(f0)
        a = rng.choice(prev, w)
        far = rng.integers(1, base, w)
        near_pick = rng.choice(prev, w)
        b = np.where(rng.random(w) < near, near_pick, far)
        b = np.where(b == a, np.where(a > 1, a - 1, a + 1), b)
        f0 += list(2 * a + rng.integers(0, 2, w))
        f1 += list(2 * b + rng.integers(0, 2, w))
        prev = np.arange(base, base + w)
    # every AND drives a PO, so all of them stay live
    ids = np.arange(first, len(f0))
    pos = (2 * ids + rng.integers(0, 2, len(ids))).astype(np.uint32)
    return nPIs, np.array(f0, np.uint32), np.array(f1, np.uint32), pos


CASES = [
    # wide levels between narrow ones, at every offset within a group of three
    ("wide_narrow_mix", dict(seed=1, nPIs=256, widths=[3000, 2, 1, 2600, 3, 3, 1, 4000, 1, 2, 2, 2500, 5], near=0.05)),
    # widths on both sides of the 2048-node limit
    ("limit_2047_2049", dict(seed=2, nPIs=256, widths=[2047, 2048, 2049, 2048, 2047, 2047, 2049, 3, 2048])),
    # one-node levels (a chain) with short bursts: groups of 1-node levels,
    # cones of a single bottom node, and a trailing partial group
    ("chain_bursts", dict(seed=3, nPIs=32, widths=[1] * 40 + [300] + [1] * 41 + [2100] + [1] * 20, near=0.9)),
    # narrow and deep, both fanins mostly inside the launch
    ("narrow_deep", dict(seed=4, nPIs=64, widths=[8] * 600, near=0.8)),
]


@pytest.mark.parametrize("name,spec", CASES, ids=[c[0] for c in CASES])
def test_cut_schedule_parity(gpu_ctx, name, spec):
    import culs_amd
    log("cut schedule parity: %s" % name)
    g = layered(**spec)
    a = gpu_ctx.from_arrays(*g).strash()
    s = a.arrays()
    b = a.copy()
    b.reorder()
    r = b.arrays()
    cuts = b.enumerate_cuts()
    want_cuts = O.cuts(*r)
    assert np.array_equal(cuts, want_cuts), "GPU cut sets differ (first node %d)" % (
        int(np.argmax((cuts != want_cuts).reshape(-1, 40).any(axis=1))))
    want, _ = O.rw(*s)
    outs = []
    G = culs_amd.GLS_RW_GRAPH_CUTS
    for mode in (0, culs_amd.GLS_RW_LEVEL_CUTS, culs_amd.GLS_RW_PAIR_CUTS, G, G | culs_amd.GLS_RW_LEVEL_CUTS):
        x = a.copy()
        x.rewrite(cut_mode=mode)
        info = x.last_op()
        assert info["cut_schedule"] == mode
        if mode & G:
            assert info["cut_graph_nodes"] > 0 and info["cut_graph_ms"] > 0
        outs.append(x.arrays())
    for o in outs:
        assert all(np.array_equal(p, q) for p, q in zip(o, want)), "GPU rw differs from the oracle"
