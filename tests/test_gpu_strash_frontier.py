"""Strash sweep boundaries on the GPU, checked against the oracle bit for bit.

Layered AIGs whose wave widths and fanout counts cross the limits of the
in-workgroup sweep: frontiers wider than 1024 nodes, frontiers whose fanout
edges exceed 2048 (at a launch's first wave and in the middle of a launch),
nodes with more fanouts than a node record holds inline (6), duplicates that
merge within one wave and across waves, trivially folding nodes, and graphs
long enough that the sweep's window of in-LDS fanout counters moves.
"""
import sys

import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def layered(seed, nPIs, widths, hot=(), fanin0_pool=None):
    """Layer i's nodes take fanin 0 from layer i-1 (so layer i is wave i+1) and
    fanin 1 from any earlier node or PI.  `hot` lists layers whose fanin 0 is
    one node of the previous layer for 40 of their nodes.  `fanin0_pool`
    maps a layer to how many previous-layer nodes its fanin 0 is drawn from
    (few nodes -> many fanouts each).  About 4% of the nodes repeat an
    earlier node of their layer (same-wave duplicates), 3% fold trivially,
    and 3% are x & y with x folding to a fanin of an existing node of the
    previous layer (a duplicate one wave later)."""
    rng = np.random.default_rng(seed)
    f0, f1 = [0] * (nPIs + 1), [0] * (nPIs + 1)
    prev = list(range(1, nPIs + 1))
    earlier = list(prev)
    prev_nodes = []          # (id, fanin0 lit, fanin1 lit) of the previous layer
    for li, width in enumerate(widths):
        pool = prev
        if fanin0_pool and li in fanin0_pool:
            pool = [prev[int(i)] for i in rng.choice(len(prev), min(len(prev), fanin0_pool[li]), replace=False)]
        hot_node = prev[int(rng.integers(len(prev)))] if li in hot else None
        layer, layer_nodes = [], []
        for j in range(width):
            v = len(f0)
            r = rng.random()
            if hot_node is not None and j < 40:
                a = 2 * hot_node ^ int(rng.integers(2))
                b = 2 * earlier[int(rng.integers(len(earlier)))] ^ int(rng.integers(2))
            elif r < 0.04 and layer_nodes:
                _, a, b = layer_nodes[int(rng.integers(len(layer_nodes)))]
                if rng.random() < 0.5:
                    a, b = b, a
            elif r < 0.07:
                x = pool[int(rng.integers(len(pool)))]
                a = 2 * x
                b = a ^ int(rng.integers(2))        # x & x or x & !x
            elif r < 0.10 and prev_nodes and li > 0:
                # t = a0 & a0 folds to a0; t & b0 repeats the previous layer's a0 & b0
                pid, a0, b0 = prev_nodes[int(rng.integers(len(prev_nodes)))]
                t = len(f0)
                f0.append(a0 & ~1)
                f1.append(a0 & ~1)
                layer.append(t)
                v = len(f0)
                a, b = 2 * t ^ (a0 & 1), b0
            else:
                a = 2 * pool[int(rng.integers(len(pool)))] ^ int(rng.integers(2))
                b = 2 * earlier[int(rng.integers(len(earlier)))] ^ int(rng.integers(2))
            f0.append(a)
            f1.append(b)
            layer.append(v)
            layer_nodes.append((v, a, b))
        earlier += layer
        prev, prev_nodes = layer, layer_nodes
    # every AND drives a PO, so all of them are live
    po_ids = np.arange(nPIs + 1, len(f0))
    pos = (2 * po_ids ^ rng.integers(0, 2, len(po_ids))).astype(np.uint32)
    return nPIs, np.array(f0, np.uint32), np.array(f1, np.uint32), pos


CASES = [
    # entry frontier with 600 nodes and ~3600 edges; a 1500-wide wave emitted
    # mid-launch; a hot node of 40+ fanouts
    ("burst_entry", dict(seed=1, nPIs=64, widths=[600, 3600, 300, 1500, 200, 100, 100, 50, 400, 30],
                         hot=(2, 5, 8), fanin0_pool={1: 600})),
    # narrow start, then a 400-node wave whose fanouts (3000) overflow the edge
    # list in the middle of a launch
    ("edge_overflow_mid", dict(seed=2, nPIs=32, widths=[50, 60, 70, 400, 3000, 80, 90, 1200, 60, 40, 30],
                               hot=(3, 6), fanin0_pool={4: 400})),
    # long narrow chain with hot nodes and folding duplicates
    ("narrow_deep", dict(seed=3, nPIs=16, widths=[20] * 300 + [900, 30] + [15] * 100, hot=(50, 150, 301))),
    # 288k ids in narrow waves, second fanins drawn from anywhere before: the
    # sweep's in-LDS counter window (122,880 ids) moves several times, and
    # releases land both ahead of and behind it
    ("window_moves", dict(seed=4, nPIs=32, widths=[24] * 12000, hot=(100, 5000, 9000))),
]


@pytest.mark.parametrize("name,spec", CASES, ids=[c[0] for c in CASES])
def test_strash_frontier_limits(gpu_ctx, name, spec):
    log("strash frontier limits: %s" % name)
    g = layered(**spec)
    a = gpu_ctx.from_arrays(*g)
    a.strash()
    want, depth = O.strash(*g)
    assert all(np.array_equal(x, y) for x, y in zip(a.arrays(), want)), "GPU strash differs from the oracle"
    assert a.stats().levels == depth


@pytest.mark.parametrize("name,spec", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_rewrite_frontier_limits(gpu_ctx, name, spec):
    log("rw frontier limits: %s" % name)
    g = layered(**spec)
    a = gpu_ctx.from_arrays(*g).strash()
    s = a.arrays()
    a.rewrite()
    want, _ = O.rw(*s)
    assert all(np.array_equal(x, y) for x, y in zip(a.arrays(), want)), "GPU rw differs from the oracle"
