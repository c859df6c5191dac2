"""Independent pins of the oracle functions that the restatement alone left open.

Each check computes the expected answer another way than the oracle does:

- cut sets: on nodes where eviction cannot happen (at most 7 fanin-cut pairs),
  the kept set equals the minimal elements of {v} plus every support-minimised
  union of a fanin-cut pair, with truth tables composed by evaluating the
  fanins' functions on all assignments (no variable-swap code); and a
  hand-built node with ten candidate cuts whose eight kept slots, values and
  truth tables were worked out by hand from the FindCut / CutFindValue /
  CutFilter rules (rewrite.cu:49-61, 84-109, 178-193);
- evaluation: for every node and every cut of >= 3 leaves, the chosen library
  root is instantiated over the cut leaves into the structurally hashed graph
  without the node's MFFC, counting the nodes that must be created; the
  oracle's gain must equal saved - created for its choice and be the largest
  such gain over all cuts and roots (rewrite.cu:369-447);
- reorder: the output is the stable sort of the ANDs by level (ties by old id),
  with fanins remapped (rewrite.cu:1245-1279), computed from levels here;
- gain accounting: a rw pass removes at least the sum of its selected gains.

tests/test_oracle_mutants.py rebuilds the oracle with one planted mistake at a
time and checks that these pins catch each.
"""
import ctypes

import numpy as np
import pytest

import oracle_ffi as O
from test_library import load_forest

# ------------------------------------------------------------------ helpers --


def gen(host_ctx, kind, **kw):
    if kind == "random":
        return host_ctx.gen_random(kw["nPIs"], kw["nAnds"], kw["nPOs"], kw["window"], kw["seed"]).arrays()
    return host_ctx.gen_multiplier(kw["nbits"], kw["redundant"], kw["seed"]).arrays()


def unpack(cuts, v):
    """used slots of node v: slot -> (leaves tuple, tt, value)"""
    out = {}
    for s in range(8):
        w = cuts[(v * 8 + s) * 5:(v * 8 + s) * 5 + 5]
        meta = int(w[4])
        if (meta >> 30) & 1:
            nl = (meta >> 27) & 7
            out[s] = (tuple(int(x) for x in w[:nl]), meta & 0xFFFF, (meta >> 16) & 0x7FF)
    return out


def fn_of(tt, leaves, assign):
    """value of the function tt over `leaves` under a dict leaf -> bit"""
    idx = 0
    for i, l in enumerate(leaves):
        idx |= assign[l] << i
    return (tt >> idx) & 1


def compose(a, ta, ca, b, tb, cb):
    """leaves and 16-bit table of (a ^ ca) & (b ^ cb), by evaluation"""
    leaves = tuple(sorted(set(a) | set(b)))
    tt = 0
    for m in range(16):
        # variables beyond the cut width are don't-cares: the table replicates
        assign = {l: (m >> i) & 1 for i, l in enumerate(leaves)}
        va = fn_of(ta, a, assign) ^ ca
        vb = fn_of(tb, b, assign) ^ cb
        tt |= (va & vb) << m
    return leaves, tt


def minimise(leaves, tt):
    """drop the variables the table does not depend on (tables stay replicated)"""
    keep = [i for i in range(len(leaves))
            if any(((tt >> m) & 1) != ((tt >> (m ^ (1 << i))) & 1) for m in range(16))]
    nl = tuple(leaves[i] for i in keep)
    out = 0
    for m in range(16):
        # new variable j = old variable keep[j]; unused old variables read 0
        old = 0
        for j, i in enumerate(keep):
            old |= ((m >> j) & 1) << i
        out |= ((tt >> old) & 1) << m
    # replicate over the unused high variables
    for m in range(16):
        base = m & ((1 << len(keep)) - 1)
        out = (out & ~(1 << m)) | (((out >> base) & 1) << m)
    return nl, out


def min_elements(cands):
    sets = {c[0]: c for c in cands}
    keys = list(sets)
    return {k: sets[k] for k in keys if not any(set(o) < set(k) for o in keys)}


# ------------------------------------------------------------------ cut sets --

TINY = [dict(kind="random", nPIs=6, nAnds=n, nPOs=3, window=w, seed=s)
        for s, (n, w) in enumerate([(40, 8), (60, 12), (80, 10), (120, 20), (200, 30), (300, 40)], start=31)] + \
       [dict(kind="mult", nbits=3, redundant=0.3, seed=5), dict(kind="mult", nbits=4, redundant=0.2, seed=6)]


def check_cut_sets(g):
    s, _ = O.strash(*g)
    r, _ = O.reorder(*s)
    nPIs, f0, f1, pos = r
    cuts = O.cuts(*r)
    checked = 0
    for v in range(nPIs + 1, len(f0)):
        a, b = int(f0[v]), int(f1[v])
        A, B = unpack(cuts, a >> 1), unpack(cuts, b >> 1)
        S = unpack(cuts, v)
        # a slot is taken (and, when all 8 are used, one evicted) only for pairs
        # that pass the 32-bit leaf-signature prefilter (bit = leaf & 31)
        sig = lambda ls: sum(1 << (l & 31) for l in set(ls))
        passing = sum(1 for i in A for j in B if bin(sig(A[i][0]) | sig(B[j][0])).count("1") <= 4)
        if passing > 7:
            continue                         # eviction possible: covered by the hand-built case
        cands = [((v,), 0xAAAA, None)]
        early = False
        for i in sorted(A):
            for j in sorted(B):
                la, ta, _ = A[i]
                lb, tb, _ = B[j]
                if len(set(la) | set(lb)) > 4:
                    continue
                lv, tt = compose(la, ta, a & 1, lb, tb, b & 1)
                lv, tt = minimise(lv, tt)
                cands.append((lv, tt, None))
                if len(lv) < 2:
                    early = True
        if early:
            continue                         # the enumeration stops at a < 2-leaf cut
        want = min_elements(cands)
        got = {c[0]: c for c in S.values()}
        assert set(got) == set(want), (v, sorted(got), sorted(want))
        for k in want:
            assert got[k][1] == want[k][1], (v, k, hex(got[k][1]), hex(want[k][1]))
        checked += 1
    return checked


@pytest.mark.parametrize("spec", TINY, ids=lambda s: "%s-%s" % (s["kind"], s["seed"]))
def test_cut_sets_equal_minimal_unions(host_ctx, spec):
    assert check_cut_sets(gen(host_ctx, **spec)) >= 5


# node 15 = AND(13, 14), 13 = AND(9, 10), 14 = AND(11, 12), 9..12 = ANDs of PI pairs:
# its fanins have 5 cuts each, ten unions fit 4 leaves, so FindCut must evict
EVICT_G = (8,
           np.array([0] * 9 + [2, 6, 10, 14, 18, 22, 26], np.uint32),
           np.array([0] * 9 + [4, 8, 12, 16, 20, 24, 28], np.uint32),
           np.array([30], np.uint32))
# Worked out by hand.  Every node has one fanout, so CutFindValue gives a
# 2-leaf cut 2, a 3-leaf cut 3, and a 4-leaf cut 5 - 4 = 1 (more than 3
# leaves with one reference), a 1-leaf cut 1001.  Node 15's merges in slot
# order: {13,14} -> 1, {11,12,13} -> 2, {7,8,11,13} -> 3, {5,6,12,13} -> 4,
# {9,10,14} -> 5, {9,10,11,12} -> 6, {3,4,9,14} -> 7; then {1,2,10,14} finds
# no free slot and evicts the first cut of more than 2 leaves with the
# smallest value: slot 3 ({7,8,11,13}, value 1).  Unions of 5+ leaves are cut
# by the signature prefilter before any slot is taken.
EVICT_WANT = {
    0: ((15,), 0xAAAA, 1001),
    1: ((13, 14), 0x8888, 2),
    2: ((11, 12, 13), 0x8080, 3),
    3: ((1, 2, 10, 14), 0x8000, 1),
    4: ((5, 6, 12, 13), 0x8000, 1),
    5: ((9, 10, 14), 0x8080, 3),
    6: ((9, 10, 11, 12), 0x8000, 1),
    7: ((3, 4, 9, 14), 0x8000, 1),
}
EVICT_13 = {0: ((13,), 0xAAAA, 1001), 1: ((9, 10), 0x8888, 2), 2: ((3, 4, 9), 0x8080, 3),
            3: ((1, 2, 10), 0x8080, 3), 4: ((1, 2, 3, 4), 0x8000, 1)}


def check_eviction():
    cuts = O.cuts(*EVICT_G)
    assert unpack(cuts, 13) == EVICT_13
    assert unpack(cuts, 15) == EVICT_WANT


def test_findcut_eviction_known_answer():
    check_eviction()


# n4 = 1&2, n5 = n4&3, n6 = n4&n5: the union {1,2,3,4} is dominated by the kept
# {3,4} and discarded in both of its pairs; slots worked out by hand
DOM_G = (3, np.array([0, 0, 0, 0, 2, 6, 8], np.uint32), np.array([0, 0, 0, 0, 4, 8, 10], np.uint32),
         np.array([12], np.uint32))
DOM_WANT = {0: ((6,), 0xAAAA, 1001), 1: ((4, 5), 0x8888, 3), 2: ((3, 4), 0x8888, 3),
            3: ((1, 2, 5), 0x8080, 3), 4: ((1, 2, 3), 0x8080, 3)}


def check_dominance():
    cuts = O.cuts(*DOM_G)
    assert unpack(cuts, 6) == DOM_WANT


def test_cutfilter_dominance_known_answer():
    check_dominance()


# ---------------------------------------------------------------- evaluation --

def refs(g):
    nPIs, f0, f1, pos = g
    r = np.zeros(len(f0), np.int64)
    np.add.at(r, f0[nPIs + 1:] >> 1, 1)
    np.add.at(r, f1[nPIs + 1:] >> 1, 1)
    np.add.at(r, pos >> 1, 1)
    return r


def mffc_nodes(g, ref, root, leaves):
    """nodes freed when root goes, dereferencing down to the leaves"""
    nPIs, f0, f1, _ = g
    left = {}
    out, stack = {root}, [root]
    while stack:
        v = stack.pop()
        for u in (int(f0[v]) >> 1, int(f1[v]) >> 1):
            if u in leaves or u <= nPIs:
                continue
            left[u] = left.get(u, int(ref[u])) - 1
            if left[u] == 0:
                out.add(u)
                stack.append(u)
    return out


def created(forest, cls, rt, inputs, table, root, deleted):
    """nodes created when library root rt of class cls is built over the input
    literals, reusing existing nodes found in `table` (not deleted, not root)"""
    node_off, nodes, _, _ = forest
    lit = {}
    made = set()

    def build(j):
        if j < 4:
            return inputs[j]
        if j in lit:
            return lit[j]
        w = nodes[node_off[cls] + j - 4]
        a = build(w & 63)
        b = build((w >> 6) & 63)
        la = a if a is None else a ^ ((w >> 12) & 1)
        lb = b if b is None else b ^ ((w >> 13) & 1)
        hit = None
        # a fanin that is a created node (None) or the root cannot be found
        if la is not None and lb is not None and (la >> 1) != root and (lb >> 1) != root:
            hit = table.get((min(la, lb), max(la, lb)))
            if hit is not None and hit in deleted:
                hit = None
        if hit is None:
            made.add(j)
            lit[j] = None
        else:
            lit[j] = 2 * hit
        return lit[j]

    build(rt)
    return len(made), lit.get(rt)


def check_evaluation(g, zeros):
    s, _ = O.strash(*g)
    r, _ = O.reorder(*s)
    nPIs, f0, f1, pos = r
    ref = refs(r)
    cuts = O.cuts(*r)
    best = O.evaluate(*r, cuts, use_zeros=zeros).reshape(-1, 4)
    ph, pe, cl, p4 = O.npn_tables()
    forest = load_forest()
    _, _, root_off, roots = forest
    table = {(int(f0[v]), int(f1[v])): v for v in range(nPIs + 1, len(f0))}
    f0a, f1a, pa = O.u32(f0), O.u32(f1), O.u32(pos)
    checked = chosen = 0
    for v in range(nPIs + 1, len(f0)):
        gains = {}
        exact = True
        for slot, (leaves, tt, _) in unpack(cuts, v).items():
            if len(leaves) < 3:
                continue
            l4 = list(leaves) + [0] * (4 - len(leaves))
            lv = (ctypes.c_int * 4)(*l4)
            tab = ctypes.c_int(0)
            saved = O.lib().orc_mffc(len(f0), nPIs, O.P(f0a), O.P(f1a), len(pos), O.P(pa), v, lv, 4,
                                     ctypes.byref(tab))
            if tab.value >= 8:
                exact = False                 # 8-entry table overflow: saved undercounts
                break
            deleted = mffc_nodes(r, ref, v, set(l4)) - {v}
            assert saved == len(deleted) + 1
            c = int(cl[tt])
            perm = p4[pe[tt]]
            uph = int(ph[tt])
            inputs = [2 * l4[perm[j]] + ((uph >> j) & 1) for j in range(4)]
            for k in range(root_off[c + 1] - root_off[c]):
                rt = roots[root_off[c] + k]
                n_made, out_lit = created(forest, c, rt, inputs, table, v, deleted)
                if out_lit is not None and out_lit >> 1 == v:
                    continue                  # the root found itself: no replacement
                gains[(slot, k)] = saved - n_made
        if not exact:
            continue
        ok = [x for x in gains.values() if x > 0 or (zeros and x == 0)]
        b = best[v]
        if not ok:
            assert b[0] == -1, (v, b)
        else:
            assert b[0] >= 0, (v, gains)
            assert b[2] == max(ok), (v, b, max(ok))
            assert gains[(int(b[0]), int(b[1]))] == b[2], (v, b)
            chosen += 1
        checked += 1
    return checked, chosen


EVAL = [dict(kind="random", nPIs=10, nAnds=600, nPOs=6, window=60, seed=s) for s in (41, 42)] + \
       [dict(kind="mult", nbits=5, redundant=0.2, seed=43)]


@pytest.mark.parametrize("zeros", [False, True])
@pytest.mark.parametrize("spec", EVAL, ids=lambda s: "%s-%s" % (s["kind"], s["seed"]))
def test_evaluation_gain_by_instantiation(host_ctx, spec, zeros):
    checked, chosen = check_evaluation(gen(host_ctx, **spec), zeros)
    assert checked > 100 and chosen > 5


# ------------------------------------------------------------------- reorder --

def check_reorder(g):
    s, _ = O.strash(*g)
    nPIs, f0, f1, pos = s
    n = len(f0)
    lev = np.zeros(n, np.int64)
    for v in range(nPIs + 1, n):
        lev[v] = 1 + max(lev[f0[v] >> 1], lev[f1[v] >> 1])
    ands = sorted(range(nPIs + 1, n), key=lambda v: (lev[v], v))
    new = np.arange(n)
    for k, v in enumerate(ands):
        new[v] = nPIs + 1 + k
    e0, e1 = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    for v in range(nPIs + 1, n):
        a = 2 * int(new[f0[v] >> 1]) | int(f0[v] & 1)
        b = 2 * int(new[f1[v] >> 1]) | int(f1[v] & 1)
        e0[new[v]], e1[new[v]] = min(a, b), max(a, b)
    epos = np.array([2 * int(new[p >> 1]) | int(p & 1) for p in pos], np.uint32)
    (rn, r0, r1, rp), rlev = O.reorder(*s)
    assert np.array_equal(r0, e0) and np.array_equal(r1, e1) and np.array_equal(rp, epos)
    assert np.array_equal(rlev[new], lev[np.arange(n)])
    assert all(np.diff(rlev[nPIs + 1:].astype(np.int64)) >= 0)
    assert O.sim_cmp(s, (rn, r0, r1, rp)) == -1


@pytest.mark.parametrize("spec", TINY[3:] + EVAL, ids=lambda s: "%s-%s" % (s["kind"], s["seed"]))
def test_reorder_is_the_stable_level_sort(host_ctx, spec):
    check_reorder(gen(host_ctx, **spec))


# ----------------------------------------------------------- gain accounting --

def check_gain_accounting(g, zeros):
    s, _ = O.strash(*g)
    out, st = O.rw(*s, use_zeros=zeros)
    assert O.n_ands(s) - O.n_ands(out) >= st[6] >= 0
    assert O.sim_cmp(g, out) == -1


@pytest.mark.parametrize("zeros", [False, True])
@pytest.mark.parametrize("spec", EVAL, ids=lambda s: "%s-%s" % (s["kind"], s["seed"]))
def test_rw_removes_at_least_the_selected_gain(host_ctx, spec, zeros):
    check_gain_accounting(gen(host_ctx, **spec), zeros)
