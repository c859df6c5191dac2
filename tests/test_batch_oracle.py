"""Why a batched rw pass is exact: on a disjoint union of AIGs the oracle's rw
equals rw on every member alone once each node's cut-signature bit is the one
of its id in its own AIG (the reference's prefilter uses leaf & 31,
rewrite.cu:68, 262, 283), and differs without that.  The HIP batch path
computes those bits; this test pins the argument on the CPU."""
import ctypes

import numpy as np

import oracle_ffi as O


def members(ctx):
    out = []
    for i in range(8):
        if i % 3 == 2:
            out.append(ctx.gen_multiplier(4 + i, 0.1, i))
        else:
            n = 300 + 900 * i
            out.append(ctx.gen_random(16 + i, n, 8, 200, 100 + i))
    return [O.strash(*a.arrays())[0] for a in out]


def pack(mem):
    P = sum(m[0] for m in mem)
    A = sum(len(m[1]) - m[0] - 1 for m in mem)
    N = 1 + P + A
    f0 = np.zeros(N, np.uint32)
    f1 = np.zeros(N, np.uint32)
    tag = np.zeros(N, np.int64)
    pos, offs, pi, an = [], [], 0, 0
    for k, (nPIs, g0, g1, gp) in enumerate(mem):
        def tr(l):
            i, c = l >> 1, l & 1
            if i == 0:
                return l
            if i <= nPIs:
                return 2 * (pi + i) | c
            return 2 * (1 + P + an + i - nPIs - 1) | c
        tag[1 + pi:1 + pi + nPIs] = k
        for v in range(nPIs + 1, len(g0)):
            u = 1 + P + an + v - nPIs - 1
            a, b = tr(int(g0[v])), tr(int(g1[v]))
            f0[u], f1[u] = min(a, b), max(a, b)
            tag[u] = k
        offs.append((pi, len(pos), len(gp)))
        pos += [tr(int(x)) for x in gp]
        pi += nPIs
        an += len(g0) - nPIs - 1
    return (P, f0, f1, np.array(pos, np.uint32)), tag, offs


def split(R, P, tag, mem, offs):
    _, f0, f1, pos = R
    n = len(f0)
    t = np.full(n, -1, np.int64)
    t[1:P + 1] = tag[1:P + 1]
    for v in range(P + 1, n):
        a, b = int(f0[v]) >> 1, int(f1[v]) >> 1
        t[v] = t[a] if a > 0 else t[b]
    out = []
    for k, m in enumerate(mem):
        nPIs = m[0]
        pi, po, npo = offs[k]
        ids = [v for v in range(P + 1, n) if t[v] == k]
        loc = {0: 0}
        loc.update({pi + i: i for i in range(1, nPIs + 1)})
        loc.update({v: nPIs + 1 + r for r, v in enumerate(ids)})
        g0 = np.zeros(nPIs + 1 + len(ids), np.uint32)
        g1 = g0.copy()
        for v in ids:
            a, b = int(f0[v]), int(f1[v])
            la, lb = 2 * loc[a >> 1] | (a & 1), 2 * loc[b >> 1] | (b & 1)
            g0[loc[v]], g1[loc[v]] = min(la, lb), max(la, lb)
        gp = np.array([2 * loc[int(x) >> 1] | (int(x) & 1) for x in pos[po:po + npo]], np.uint32)
        out.append((nPIs, g0, g1, gp))
    return out


def reorder_with_tags(U, tag):
    nPIs, f0, f1, _ = U
    n = len(f0)
    L = np.zeros(n, np.int64)
    for v in range(nPIs + 1, n):
        L[v] = 1 + max(L[int(f0[v]) >> 1], L[int(f1[v]) >> 1])
    order = np.lexsort((np.arange(n), L))
    newid = np.empty(n, np.int64)
    newid[order] = np.arange(n)
    tagR = np.zeros(n, np.int64)
    tagR[newid] = tag
    return O.reorder(*U)[0], tagR


def member_sig_bits(tagR):
    sig = np.zeros(len(tagR), np.uint8)
    seen = {}
    for u in range(1, len(tagR)):
        k = int(tagR[u])
        seen[k] = seen.get(k, 0) + 1
        sig[u] = seen[k] & 31
    return sig


def differing(mem, U, tag, offs, zeros, sig=None):
    Ru, tagR = reorder_with_tags(U, tag)
    bits = member_sig_bits(tagR) if sig else None
    O.lib().orc_set_sig_bits(bits.ctypes.data_as(ctypes.c_void_p) if bits is not None else None)
    try:
        R, _ = O.rw(*Ru, use_zeros=zeros)
    finally:
        O.lib().orc_set_sig_bits(None)
    got = split(R, U[0], tag, mem, offs)
    bad = 0
    for m, g in zip(mem, got):
        want, _ = O.rw(*m, use_zeros=zeros)
        bad += not all(np.array_equal(x, y) for x, y in zip(g[1:], want[1:]))
    return bad


def test_union_rw_equals_members_with_member_signatures(host_ctx):
    mem = members(host_ctx)
    U, tag, offs = pack(mem)
    for zeros in (False, True):
        assert differing(mem, U, tag, offs, zeros, sig=True) == 0
    # the signature bits are what makes it exact
    assert differing(mem, U, tag, offs, False, sig=False) + differing(mem, U, tag, offs, True, sig=False) > 0
