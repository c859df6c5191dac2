"""Cut-enumeration schedules on the GPU, checked against the oracle bit for bit.

The default schedule covers three levels per launch (a node recomputes its
in-launch fanin cone in its workgroup; GLS_RW_QUAD_CUTS four) and falls back to one launch per level
for any group of three levels holding a level wider than 2048 nodes.  These
layered AIGs place wide and narrow levels, widths on both sides of that limit
and long runs of one-node levels at every offset within a group, so groups of
every shape occur: the cut records must equal the oracle's, and rw must give
the same AIG under every schedule, launched from the stream or from one
captured HIP graph, and equal the oracle's rw.

The planned schedule (GLS_RW_PLAN_CUTS) is switched off in the library: its
kernel made an illegal memory access on the GPU in the parity test below (also
after a fix), the cause is not found, and the library refuses the flag.  Its
parity test stays, skipped, and a test checks the refusal.
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
    want_cuts = O.cuts(*r)

    def same_cuts(cuts, what):
        assert np.array_equal(cuts, want_cuts), "GPU cut sets differ under %s (first node %d)" % (
            what, int(np.argmax((cuts != want_cuts).reshape(-1, 40).any(axis=1))))

    same_cuts(b.enumerate_cuts(), "the default schedule")
    same_cuts(b.enumerate_cuts(culs_amd.GLS_RW_QUAD_CUTS), "four levels per launch")
    want, _ = O.rw(*s)
    outs = []
    G = culs_amd.GLS_RW_GRAPH_CUTS
    Q = culs_amd.GLS_RW_QUAD_CUTS
    for mode in (0, culs_amd.GLS_RW_LEVEL_CUTS, culs_amd.GLS_RW_PAIR_CUTS, Q, G, G | culs_amd.GLS_RW_LEVEL_CUTS, G | Q):
        x = a.copy()
        x.rewrite(cut_mode=mode)
        info = x.last_op()
        assert info["cut_schedule"] == mode
        if mode & G:
            assert info["cut_graph_nodes"] > 0 and info["cut_graph_ms"] > 0
        outs.append(x.arrays())
    for o in outs:
        assert all(np.array_equal(p, q) for p, q in zip(o, want)), "GPU rw differs from the oracle"


@pytest.mark.skip(reason="GLS_RW_PLAN_CUTS is disabled: its kernel faulted the GPU (illegal memory access) "
                         "in this test, also after a fix; the cause is not found")
@pytest.mark.parametrize("name,spec", CASES, ids=[c[0] for c in CASES])
def test_planned_schedule_parity(gpu_ctx, name, spec):
    """The planned schedule at every launch depth it offers, record by record."""
    import culs_amd
    log("planned schedule parity: %s" % name)
    a = gpu_ctx.from_arrays(*layered(**spec)).strash()
    b = a.copy()
    b.reorder()
    want_cuts = O.cuts(*b.arrays())
    P = culs_amd.GLS_RW_PLAN_CUTS
    levels = gpu_ctx.get_tuning(1)["cut_levels"]
    try:
        for k in (1, 2, 3, 4, 5, 6):
            gpu_ctx.set_tuning(-1, cut_levels=k)
            assert np.array_equal(b.enumerate_cuts(P), want_cuts), "%d levels per launch" % k
            x = a.copy()
            x.rewrite(cut_mode=P)
            want, _ = O.rw(*a.arrays())
            assert all(np.array_equal(p, q) for p, q in zip(x.arrays(), want))
    finally:
        gpu_ctx.set_tuning(-1, cut_levels=levels)


def test_planned_schedule_refused(gpu_ctx):
    """The disabled schedule is refused before anything is launched."""
    import culs_amd
    log("planned schedule refused")
    a = gpu_ctx.from_arrays(*layered(seed=5, nPIs=16, widths=[8] * 6)).strash()
    b = a.copy()
    b.reorder()
    with pytest.raises(culs_amd.GlsError, match="disabled"):
        b.enumerate_cuts(culs_amd.GLS_RW_PLAN_CUTS)
    with pytest.raises(culs_amd.GlsError, match="disabled"):
        a.copy().rewrite(cut_mode=culs_amd.GLS_RW_PLAN_CUTS)
    # the context still works afterwards
    c = a.copy()
    c.rewrite()
    want, _ = O.rw(*a.arrays())
    assert all(np.array_equal(p, q) for p, q in zip(c.arrays(), want))
