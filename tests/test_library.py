"""Pins the rewriting library: NPN tables and the class forest.

Golden values: sha256 of the reference's int32 tables (src/algorithms/rewrite_library.inc,
layout of rewrite.cu:1281-1307) and the spot values recorded from it.  Both the
oracle's and the product's closed-form generators must reproduce them, and the
forest must satisfy the functional invariant that makes rewriting sound: every
subgraph root, with the (permutation, phase) of a truth table t, computes t or ~t.
"""
import hashlib

import numpy as np
import pytest

import oracle_ffi as O

GOLDEN = {
    "phase": "74c45178e14ecdd909ced74b69d1a796ed592b3a02070b4104d3d54421384753",
    "perm": "e47b3bac12e5548a86953c0b226c97c701bcf12bb39efee63bf4625ebb39fc18",
    "perms4": "9b9f85d9c376c940e92a72e75d95b85567a4bfa36d1617b0ecf2e3ca2e8be638",
    "cls": "cc474920298f5ce89d4d836ee0ce17dbbfd01d5081f975a6018c110f693c3e72",
    "class_section": "16447e3708ac5113e2b32d9abd25e4044d5ec3903e8633ea93e82a05e1777113",
}

SPOT = {0x0001: (0, 0, 1), 0x8000: (15, 0, 1), 0x8888: (12, 16, 5), 0xCAFE: (20, 15, 27),
        0x1234: (0, 15, 39), 0x6996: (0, 0, 221), 0xAAAA: (16, 18, 21), 0x00FF: (0, 0, 21)}


def sha(a):
    return hashlib.sha256(np.asarray(a).astype("<i4").tobytes()).hexdigest()


def check_tables(ph, pe, cl, p4):
    assert sha(ph) == GOLDEN["phase"]
    assert sha(pe) == GOLDEN["perm"]
    assert sha(cl) == GOLDEN["cls"]
    assert sha(p4.reshape(-1)) == GOLDEN["perms4"]
    for t, (a, b, c) in SPOT.items():
        assert (ph[t], pe[t], cl[t]) == (a, b, c), hex(t)


def test_oracle_npn_tables_match_reference():
    check_tables(*O.npn_tables())


def test_product_npn_tables_match_reference():
    import culs_amd
    check_tables(*culs_amd.npn_tables())


def test_npn_class_properties():
    ph, pe, cl, p4 = O.npn_tables()
    # 222 classes; each representative is the smallest truth table of its class
    assert cl.max() == 221 and len(np.unique(cl)) == 222
    first = {}
    for t in range(65536):
        first.setdefault(int(cl[t]), t)
    reps = [first[c] for c in range(1, 222)]
    assert reps == sorted(reps)
    # representatives map to themselves
    for r in reps:
        assert ph[r] == 0 and pe[r] == 0


def load_forest():
    import re
    import os
    path = os.path.join(O.ROOT, "culs_amd", "csrc", "rwlib_forest.h")
    txt = open(path).read()

    def arr(name):
        m = re.search(r"%s\[\d+\] = \{([^}]*)\}" % name, txt)
        return [int(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()]

    return arr("gls_forest_node_off"), arr("gls_forest_nodes"), arr("gls_forest_root_off"), arr("gls_forest_roots")


def test_forest_expands_to_reference_class_section():
    node_off, nodes, root_off, roots = load_forest()
    n_nodes = [node_off[c + 1] - node_off[c] for c in range(222)]
    n_sub = [root_off[c + 1] - root_off[c] for c in range(222)]
    sec = n_nodes + n_sub
    for c in range(222):
        for w in nodes[node_off[c]:node_off[c + 1]]:
            sec += [w & 63, (w >> 6) & 63, (w >> 12) & 1, (w >> 13) & 1]
        sec += roots[root_off[c]:root_off[c + 1]]
    assert len(sec) == 21304
    assert sha(sec) == GOLDEN["class_section"]
    assert sum(n_nodes) == 4947 and sum(n_sub) == 1072
    assert sum(1 for c in range(222) if n_sub[c] > 0) == 221 and n_sub[0] == 0
    assert max(n_nodes) <= 50 and max(n_sub) <= 15


def lib_truths(c, node_off, nodes, perm, phase):
    ev = [0xAAAA, 0xCCCC, 0xF0F0, 0xFF00]
    tt = [(ev[perm[j]] ^ (0xFFFF if (phase >> j) & 1 else 0)) for j in range(4)]
    for w in nodes[node_off[c]:node_off[c + 1]]:
        a = tt[w & 63] ^ (0xFFFF if (w >> 12) & 1 else 0)
        b = tt[(w >> 6) & 63] ^ (0xFFFF if (w >> 13) & 1 else 0)
        tt.append(a & b & 0xFFFF)
    return tt


def test_every_root_computes_t_or_not_t():
    """Soundness of the library + NPN tables, checked for all 65,534 functions."""
    ph, pe, cl, p4 = O.npn_tables()
    node_off, nodes, root_off, roots = load_forest()
    negations = 0
    for t in range(1, 0xFFFF):
        c = int(cl[t])
        tts = lib_truths(c, node_off, nodes, p4[pe[t]], int(ph[t]) & 15)
        rs = roots[root_off[c]:root_off[c + 1]]
        assert rs, hex(t)
        for k, rt in enumerate(rs):
            v = tts[rt]
            assert v == t or v == (~t & 0xFFFF), (hex(t), k)
        negations += tts[rs[0]] != t
    assert negations == 32767


def test_padded_input_unused_for_3_input_functions():
    """3-leaf cuts are padded with the constant node as input 3 (rewrite.cu:386-389);
    no root may depend on the library input mapped to that pad."""
    ph, pe, cl, p4 = O.npn_tables()
    node_off, nodes, root_off, roots = load_forest()
    for t in range(1, 0xFFFF):
        if (t & 0x00FF) != (t >> 8):      # depends on variable 3
            continue
        c = int(cl[t])
        perm = p4[pe[t]]
        pad_input = [j for j in range(4) if perm[j] == 3][0]
        for rt in roots[root_off[c]:root_off[c + 1]]:
            used = {rt}
            for j in range(rt, 3, -1):
                if j in used:
                    w = nodes[node_off[c] + j - 4]
                    used.add(w & 63)
                    used.add((w >> 6) & 63)
            assert pad_input not in used, (hex(t), rt)
