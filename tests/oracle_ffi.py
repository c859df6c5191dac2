"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) — test infrastructure.

Only the tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
and only as the checker.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_P = ctypes.c_void_p
_lib = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build_oracle()
        _lib = _load(ORACLE_LIB)
    return _lib


@contextlib.contextmanager
def using(path):
    """Runs the wrappers against another build of the oracle (a mutant)."""
    global _lib
    old = lib()
    _lib = _load(path)
    try:
        yield
    finally:
        _lib = old


def _load(path):
    L = ctypes.CDLL(path)
    i32, u64 = ctypes.c_int, ctypes.c_uint64
    L.orc_npn_tables.argtypes = [_P] * 4
    L.orc_strash.argtypes = [i32, i32, _P, _P, i32, _P, _P, _P, _P, _P]
    L.orc_reorder.argtypes = [i32, i32, _P, _P, i32, _P, _P, _P, _P, _P]
    L.orc_cuts.argtypes = [i32, i32, _P, _P, i32, _P, _P]
    L.orc_eval.argtypes = [i32, i32, _P, _P, i32, _P, _P, i32, _P]
    L.orc_rw.argtypes = [i32, i32, _P, _P, i32, _P, i32, i32, _P, _P, _P, _P]
    L.orc_rw_culs.argtypes = [i32, i32, _P, _P, i32, _P, i32, i32, _P, _P, _P, _P]
    L.orc_sim_cmp.argtypes = [i32, _P, _P, _P, i32, _P, _P, _P, i32, i32, i32, u64]
    L.orc_depth.argtypes = [i32, i32, _P, _P]
    L.orc_sim_po_words.argtypes = [i32, i32, _P, _P, i32, _P, i32, u64, i32, _P]
    L.orc_sim_po_words.restype = None
    return L


def P(a):
    return None if a is None else a.ctypes.data_as(_P)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def npn_tables():
    ph, pe, cl = (np.zeros(65536, np.uint8) for _ in range(3))
    p4 = np.zeros(96, np.uint8)
    lib().orc_npn_tables(P(ph), P(pe), P(cl), P(p4))
    return ph, pe, cl, p4.reshape(24, 4)


def strash(nPIs, f0, f1, pos):
    f0, f1, pos = u32(f0), u32(f1), u32(pos)
    n = len(f0)
    o0, o1 = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    op = np.zeros(max(len(pos), 1), np.uint32)
    d = ctypes.c_int(0)
    r = lib().orc_strash(n, nPIs, P(f0), P(f1), len(pos), P(pos), P(o0), P(o1), P(op), ctypes.byref(d))
    if r < 0:
        raise RuntimeError("orc_strash failed %d" % r)
    return (nPIs, o0[:r].copy(), o1[:r].copy(), op[:len(pos)].copy()), d.value


def strash_cpu(nPIs, f0, f1, pos):
    """`st -c`: host strash without trivial folding -> (graph, levels, depth)."""
    f0, f1, pos = u32(f0), u32(f1), u32(pos)
    n = len(f0)
    o0, o1, ol = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    op = np.zeros(max(len(pos), 1), np.uint32)
    d = ctypes.c_int(0)
    r = lib().orc_strash_cpu(n, nPIs, P(f0), P(f1), len(pos), P(pos), P(o0), P(o1), P(op), P(ol), ctypes.byref(d))
    if r < 0:
        raise RuntimeError("orc_strash_cpu failed %d" % r)
    return (nPIs, o0[:r].copy(), o1[:r].copy(), op[:len(pos)].copy()), ol[:r].copy(), d.value


def balance(nPIs, f0, f1, pos, sort_dec_id=1):
    """`b` (sort_dec_id 1) or `b -s` (0) without its closing strash."""
    f0, f1, pos = u32(f0), u32(f1), u32(pos)
    n = len(f0)
    cap = 3 * n + 64
    o0, o1 = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
    op = np.zeros(max(len(pos), 1), np.uint32)
    r = lib().orc_balance_ex(n, nPIs, P(f0), P(f1), len(pos), P(pos), cap, P(o0), P(o1), P(op), int(sort_dec_id))
    if r < 0:
        raise RuntimeError("orc_balance failed %d" % r)
    return (nPIs, o0[:r].copy(), o1[:r].copy(), op[:len(pos)].copy())


def thrust_shuffle_perm(length, seed):
    """Output position k of thrust's shuffle holds input element perm[k]."""
    perm = np.zeros(max(length, 1), np.uint32)
    lib().orc_thrust_shuffle_perm(ctypes.c_uint32(length), ctypes.c_uint32(seed), P(perm))
    return perm[:length]


def tt_place(t, phase, shrink):
    """The oracle's truth-table variable placement (tt_stretch / tt_shrink of
    rewrite.cu:213-241) over arrays."""
    t = u32(t)
    ph = np.ascontiguousarray(phase, dtype=np.uint8)
    out = np.zeros(len(t), np.uint32)
    lib().orc_tt_place(len(t), P(t), P(ph), int(shrink), P(out))
    return out


def reorder(nPIs, f0, f1, pos):
    f0, f1, pos = u32(f0), u32(f1), u32(pos)
    n = len(f0)
    o0, o1, lev = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    op = np.zeros(max(len(pos), 1), np.uint32)
    r = lib().orc_reorder(n, nPIs, P(f0), P(f1), len(pos), P(pos), P(o0), P(o1), P(op), P(lev))
    if r < 0:
        raise RuntimeError("orc_reorder failed %d" % r)
    return (nPIs, o0, o1, op[:len(pos)].copy()), lev


def cuts(nPIs, f0, f1, pos):
    f0, f1, pos = u32(f0), u32(f1), u32(pos)
    n = len(f0)
    out = np.zeros(n * 40, np.uint32)
    lib().orc_cuts(n, nPIs, P(f0), P(f1), len(pos), P(pos), P(out))
    return out


def evaluate(nPIs, f0, f1, pos, cut_arr, use_zeros=False):
    f0, f1, pos = u32(f0), u32(f1), u32(pos)
    n = len(f0)
    best = np.zeros(n * 4, np.int32)
    lib().orc_eval(n, nPIs, P(f0), P(f1), len(pos), P(pos), P(u32(cut_arr)), 1 if use_zeros else 0, P(best))
    return best


def _rw(fn, nPIs, f0, f1, pos, use_zeros):
    f0, f1, pos = u32(f0), u32(f1), u32(pos)
    n = len(f0)
    cap = 3 * n + 64
    o0, o1 = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
    op = np.zeros(max(len(pos), 1), np.uint32)
    st = np.zeros(8, np.int64)
    r = fn(n, nPIs, P(f0), P(f1), len(pos), P(pos), 1 if use_zeros else 0, cap, P(o0), P(o1), P(op), P(st))
    if r < 0:
        raise RuntimeError("oracle rw failed %d" % r)
    return (nPIs, o0[:r].copy(), o1[:r].copy(), op[:len(pos)].copy()), st


def rw(nPIs, f0, f1, pos, use_zeros=False):
    """This package's rw pass (strashed in, strashed out)."""
    return _rw(lib().orc_rw, nPIs, f0, f1, pos, use_zeros)


def rw_compose(nPIs, f0, f1, pos, use_zeros=False):
    """The composed graph rw hands to its final strash (old ids keep the reordered numbering)."""
    return _rw(lib().orc_rw_compose, nPIs, f0, f1, pos, use_zeros)


def rw_culs(nPIs, f0, f1, pos, use_zeros=False):
    """The reference's replacement method (sequential choice pass) — QoR baseline."""
    return _rw(lib().orc_rw_culs, nPIs, f0, f1, pos, use_zeros)


def sim_cmp(a, b, words=64, seed=1):
    """-1 if equivalent on all simulated patterns (exhaustive for <= 16 PIs)."""
    nPIs, a0, a1, ap = a
    nPIs2, b0, b1, bp = b
    assert nPIs == nPIs2 and len(ap) == len(bp)
    a0, a1, ap, b0, b1, bp = map(u32, (a0, a1, ap, b0, b1, bp))
    return lib().orc_sim_cmp(len(a0), P(a0), P(a1), P(ap), len(b0), P(b0), P(b1), P(bp), nPIs, len(ap), words, seed)


def sim_po_words(g, words, seed=1, exhaustive=False):
    """PO values [nPOs, words] under the pattern definition of gls_simulate."""
    nPIs, f0, f1, pos = g
    f0, f1, pos = u32(f0), u32(f1), u32(pos)
    out = np.zeros((max(len(pos), 1), words), np.uint64)
    lib().orc_sim_po_words(len(f0), nPIs, P(f0), P(f1), len(pos), P(pos), words, seed, 1 if exhaustive else 0, P(out))
    return out[:len(pos)]


def depth(nPIs, f0, f1):
    f0, f1 = u32(f0), u32(f1)
    return lib().orc_depth(len(f0), nPIs, P(f0), P(f1))


def n_ands(g):
    return len(g[1]) - g[0] - 1
