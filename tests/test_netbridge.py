"""gls_aig_from_net (the ABC patch's gget) against a test double of ABC's
strashed network: objects with ids in arbitrary order, a constant-1 node,
CIs, ANDs made by a structurally hashing `aig_and` (ABC's Abc_AigAnd rules:
trivial cases folded, fanins ordered, one node per fanin pair), COs, and
dangling ANDs no CO reaches.

- numbering: PIs in CI order, ANDs in the depth-first post-order from the COs
  (CO order, fanin 0 first), as the reference's AbcNtkToGpuMan
  (src/abc_patch/abc_patch_transform.cpp:5-87) takes it from Abc_NtkDfs; the
  expected arrays come from a recursive restatement of that order here;
- function: every CO of the network, simulated on the network itself, equals
  the AIG's PO simulated on its arrays;
- gput direction (GpuManToAbcNtk, abc_patch_transform.cpp:89-149): the AIG's
  arrays rebuilt into a network with `aig_and` in id order and read back give
  the same AIG;
- errors: a cycle, a fanin that is a CO or an unlisted CI, a bad kind."""
import sys

import numpy as np
import pytest

import culs_amd

sys.setrecursionlimit(100000)


class MockNtk:
    """Just enough of a strashed Abc_Ntk_t: object kinds, fanins, complements."""

    def __init__(self, rng, n_ci):
        self.kind, self.f0, self.f1, self.c0, self.c1 = [], [], [], [], []
        self.rng = rng
        self.table = {}
        self.const1 = self._obj(culs_amd.GLS_NET_CONST1)
        self.ci = [self._obj(culs_amd.GLS_NET_CI) for _ in range(n_ci)]
        self.co = []

    def _obj(self, kind, f0=-1, f1=-1, c0=0, c1=0):
        self.kind.append(kind)
        self.f0.append(f0)
        self.f1.append(f1)
        self.c0.append(c0)
        self.c1.append(c1)
        return len(self.kind) - 1

    # edges are (object, complement)
    def const(self, value):
        return (self.const1, 0 if value else 1)

    def aig_and(self, a, b):
        """Abc_AigAnd: trivial cases, then one node per ordered fanin pair."""
        if a[0] == b[0]:
            return a if a[1] == b[1] else self.const(False)
        if a[0] == self.const1:
            return b if a[1] == 0 else self.const(False)
        if b[0] == self.const1:
            return a if b[1] == 0 else self.const(False)
        if a[0] > b[0]:
            a, b = b, a
        key = (a, b)
        if key not in self.table:
            self.table[key] = self._obj(culs_amd.GLS_NET_AND, a[0], b[0], a[1], b[1])
        return (self.table[key], 0)

    def add_co(self, e):
        self.co.append(self._obj(culs_amd.GLS_NET_CO, e[0], -1, e[1], 0))

    def arrays(self, perm=None):
        """the network's arrays, its objects renumbered by perm (old -> new)"""
        n = len(self.kind)
        perm = np.arange(n) if perm is None else perm
        kind = np.zeros(n, np.uint8)
        f0 = np.full(n, -1, np.int32)
        f1 = np.full(n, -1, np.int32)
        c0 = np.zeros(n, np.uint8)
        c1 = np.zeros(n, np.uint8)
        for v in range(n):
            w = perm[v]
            kind[w] = self.kind[v]
            f0[w] = perm[self.f0[v]] if self.f0[v] >= 0 else -1
            f1[w] = perm[self.f1[v]] if self.f1[v] >= 0 else -1
            c0[w], c1[w] = self.c0[v], self.c1[v]
        return kind, f0, f1, c0, c1, perm[np.array(self.ci, np.int64)], perm[np.array(self.co, np.int64)]


def random_ntk(seed, n_ci=12, n_and=400, n_co=10, dangling=30):
    rng = np.random.default_rng(seed)
    ntk = MockNtk(rng, n_ci)
    edges = [(c, 0) for c in ntk.ci] + [ntk.const(True)]
    for _ in range(n_and):
        a = edges[int(rng.integers(len(edges)))]
        b = edges[int(rng.integers(len(edges)))]
        e = ntk.aig_and((a[0], a[1] ^ int(rng.integers(2))), (b[0], b[1] ^ int(rng.integers(2))))
        edges.append(e)
    for _ in range(n_co):
        e = edges[len(edges) - 1 - int(rng.integers(min(len(edges), 60)))]
        ntk.add_co((e[0], e[1] ^ int(rng.integers(2))))
    for _ in range(dangling):   # ANDs no CO reaches
        a = edges[int(rng.integers(len(edges)))]
        b = edges[int(rng.integers(len(edges)))]
        ntk.aig_and(a, (b[0], 1 - b[1]))
    return ntk


def expected_gget(kind, f0, f1, c0, c1, ci, co):
    """AbcNtkToGpuMan's numbering, restated recursively: Abc_NtkDfs post-order."""
    lit = {int(v): 2 * (k + 1) for k, v in enumerate(ci)}
    for v in range(len(kind)):
        if kind[v] == culs_amd.GLS_NET_CONST1:
            lit[v] = 1
    order, seen = [], set()

    def visit(v):
        if kind[v] != culs_amd.GLS_NET_AND or v in seen:
            return
        seen.add(v)
        visit(int(f0[v]))
        visit(int(f1[v]))
        order.append(v)

    for c in co:
        visit(int(f0[c]))
    nPIs = len(ci)
    for k, v in enumerate(order):
        lit[v] = 2 * (nPIs + 1 + k)
    n = nPIs + 1 + len(order)
    e0, e1 = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    for k, v in enumerate(order):
        a, b = lit[int(f0[v])] ^ int(c0[v]), lit[int(f1[v])] ^ int(c1[v])
        e0[nPIs + 1 + k], e1[nPIs + 1 + k] = min(a, b), max(a, b)
    pos = np.array([lit[int(f0[c])] ^ int(c0[c]) for c in co], np.uint32)
    return nPIs, e0, e1, pos


def simulate_ntk(kind, f0, f1, c0, c1, ci, co, words):
    """CO values of the network under the given PI words (one uint64 per CI)."""
    val = {}
    for k, v in enumerate(ci):
        val[int(v)] = words[k]
    full = np.uint64(0xFFFFFFFFFFFFFFFF)

    def ev(v):
        if v in val:
            return val[v]
        if kind[v] == culs_amd.GLS_NET_CONST1:
            r = full
        else:
            a = ev(int(f0[v])) ^ (full if c0[v] else np.uint64(0))
            b = ev(int(f1[v])) ^ (full if c1[v] else np.uint64(0))
            r = a & b
        val[v] = r
        return r

    return [ev(int(f0[c])) ^ (full if c0[c] else np.uint64(0)) for c in co]


def simulate_aig(g, words):
    nPIs, f0, f1, pos = g
    full = np.uint64(0xFFFFFFFFFFFFFFFF)
    v = [full] + list(words) + [np.uint64(0)] * (len(f0) - nPIs - 1)
    v[0] = np.uint64(0)   # AIGER: literal 0 = FALSE

    def lit(x):
        r = v[x >> 1]
        return r ^ full if x & 1 else r

    for i in range(nPIs + 1, len(f0)):
        v[i] = lit(int(f0[i])) & lit(int(f1[i]))
    return [lit(int(p)) for p in pos]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gget_numbering_and_function(host_ctx, seed):
    ntk = random_ntk(seed)
    perm = np.random.default_rng(seed + 100).permutation(len(ntk.kind))   # object ids in arbitrary order
    arr = ntk.arrays(perm)
    got = host_ctx.from_net(*arr).arrays()
    want = expected_gget(*arr)
    assert got[0] == want[0]
    for x, y in zip(got[1:], want[1:]):
        assert np.array_equal(x, y)
    # dangling ANDs are left out
    n_and_reached = len(want[1]) - want[0] - 1
    assert n_and_reached < sum(1 for k in ntk.kind if k == culs_amd.GLS_NET_AND)
    words = list(np.random.default_rng(seed).integers(0, 2**63, size=len(ntk.ci), dtype=np.uint64))
    assert all(a == b for a, b in zip(simulate_ntk(*arr, words), simulate_aig(got, words)))


def test_gput_direction_round_trip(host_ctx):
    ntk = random_ntk(7, n_and=600)
    g = host_ctx.from_net(*ntk.arrays()).arrays()
    nPIs, f0, f1, pos = g
    # GpuManToAbcNtk: const, PIs, then one aig_and per AND in id order, POs
    back = MockNtk(np.random.default_rng(0), nPIs)
    edge = {0: back.const(False)}
    for k in range(nPIs):
        edge[k + 1] = (back.ci[k], 0)

    def e(lit):
        o, c = edge[lit >> 1]
        return (o, c ^ (lit & 1))

    for v in range(nPIs + 1, len(f0)):
        edge[v] = back.aig_and(e(int(f0[v])), e(int(f1[v])))
    for p in pos:
        back.add_co(e(int(p)))
    again = host_ctx.from_net(*back.arrays()).arrays()
    for x, y in zip(again, g):
        assert np.array_equal(np.asarray(x), np.asarray(y))


def test_unused_ids(host_ctx):
    """Ids of deleted objects, given as unlisted CIs nothing references, are
    ignored (the ABC patch fills the gaps of Abc_NtkObjNumMax that way)."""
    ntk = random_ntk(11, n_and=80, n_co=4, dangling=0)
    kind, f0, f1, c0, c1, ci, co = ntk.arrays()
    gap = 5
    pad = lambda a, v: np.concatenate([a, np.full(gap, v, a.dtype)])
    kind2 = pad(kind, culs_amd.GLS_NET_CI)
    got = host_ctx.from_net(kind2, pad(f0, -1), pad(f1, -1), pad(c0, 0), pad(c1, 0), ci, co).arrays()
    want = host_ctx.from_net(kind, f0, f1, c0, c1, ci, co).arrays()
    assert all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(got, want))


def test_empty_and_constant_outputs(host_ctx):
    ntk = MockNtk(np.random.default_rng(0), 3)
    ntk.add_co(ntk.const(True))
    ntk.add_co(ntk.const(False))
    ntk.add_co((ntk.ci[1], 1))
    g = host_ctx.from_net(*ntk.arrays()).arrays()
    assert g[0] == 3 and len(g[1]) == 4
    assert list(g[3]) == [1, 0, 5]


def test_errors(host_ctx):
    G = culs_amd
    # a cycle: x = AND(a, y), y = AND(x, a), CO -> x
    kind = np.array([G.GLS_NET_CONST1, G.GLS_NET_CI, G.GLS_NET_AND, G.GLS_NET_AND, G.GLS_NET_CO], np.uint8)
    f0 = np.array([-1, -1, 1, 2, 2], np.int32)
    f1 = np.array([-1, -1, 3, 1, -1], np.int32)
    z = np.zeros(5, np.uint8)
    with pytest.raises(culs_amd.GlsError):
        host_ctx.from_net(kind, f0, f1, z, z, [1], [4])
    ntk = random_ntk(5, n_and=50, n_co=3, dangling=0)
    kind, f0, f1, c0, c1, ci, co = ntk.arrays()
    ands = [v for v in range(len(kind)) if kind[v] == G.GLS_NET_AND]
    # a fanin that is a CO
    bad = f1.copy()
    bad[ands[-1]] = co[0]
    with pytest.raises(culs_amd.GlsError):
        host_ctx.from_net(kind, f0, bad, c0, c1, ci, co)
    # a CI missing from the CI list
    with pytest.raises(culs_amd.GlsError):
        host_ctx.from_net(kind, f0, f1, c0, c1, ci[1:], co)
    # a bad kind
    k2 = kind.copy()
    k2[ands[0]] = 9
    with pytest.raises(culs_amd.GlsError):
        host_ctx.from_net(k2, f0, f1, c0, c1, ci, co)
