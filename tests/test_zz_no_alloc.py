"""gls_cec_noalloc (the check on a static device buffer, no allocation)
against gls_cec, through the C ABI: two calls in a row, the second with a
smaller input; a difference found by both at the same (PO, pattern); inputs
beyond the buffer and handles not on the device are refused."""
import ctypes
import sys

import numpy as np
import pytest

import culs_amd
from culs_amd import CecInfo, lib

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noalloc(a, b, words, seed):
    L = lib()
    f = L.gls_cec_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(CecInfo)]
    info = CecInfo()
    rc = f(a._h, b._h, words, seed, ctypes.byref(info))
    return rc, info.as_dict()


def mutated(ctx, g, po):
    nPIs, f0, f1, pos = g
    pos = pos.copy()
    pos[po] ^= 1
    return ctx.from_arrays(nPIs, f0, f1, pos)


def on_device(x):
    x.strash()          # any device operation puts the handle on the device
    return x


def test_noalloc_equals_cec(gpu_ctx):
    ctx = gpu_ctx
    # first call: a 20k-AND AIG against its rw output and against a changed copy
    a = ctx.gen_random(64, 20000, 32, 1000, 1)
    g = a.arrays()
    a.strash()
    r = a.copy()
    r.rewrite()
    bad = on_device(mutated(ctx, r.arrays(), 7))
    log("noalloc: 20k pair")
    for other in (r, bad):
        rc, got = noalloc(a, other, 64, 5)
        assert rc == 0, culs_amd.lib().gls_last_error(ctx._h)
        want = a.cec(other, words=64, seed=5)
        assert got == want, (got, want)
    assert noalloc(a, bad, 64, 5)[1]["equivalent"] == 0
    # second call, smaller: exhaustive over 16 PIs
    m = ctx.gen_multiplier(8, 0.1, 2)
    m.strash()
    mr = m.copy()
    mr.rewrite(use_zeros=True)
    mbad = on_device(mutated(ctx, mr.arrays(), 3))
    log("noalloc: 8-bit multiplier pair")
    for other in (mr, mbad):
        rc, got = noalloc(m, other, 1, 5)
        assert rc == 0
        want = m.cec(other, words=1, seed=5)
        assert got == want, (got, want)
        assert got["exhaustive"] == 1
    assert noalloc(m, mbad, 1, 5)[1]["po"] == 3


def test_noalloc_refuses(gpu_ctx):
    ctx = gpu_ctx
    # not on the device: the call would have to allocate
    h = ctx.gen_random(16, 500, 8, 100, 3)
    d = on_device(ctx.gen_random(16, 500, 8, 100, 3))
    rc, _ = noalloc(h, d, 1, 1)
    assert rc == culs_amd_err("GLS_ERR_ARG")
    # beyond the buffer: (128 MiB - 256) / 8 objects in both AIGs together
    big = ctx.gen_random(64, 8_500_000, 32, 1000, 7)
    big.simulate(1)     # uploads it, unchanged: 8,500,065 objects, twice
    rc, _ = noalloc(big, big, 1, 1)
    assert rc == culs_amd_err("GLS_ERR_ARG")
    assert b"static" in culs_amd.lib().gls_last_error(ctx._h)


def culs_amd_err(name):
    return {"GLS_ERR_ARG": -1}[name]
