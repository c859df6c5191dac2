"""Id order of the results, the difference a maintainer of the ABC patch meets
when `grw` moves to the library: the device strash numbers the ANDs wave by
wave, a wave being the level of an AND in the graph it hashes (the
reference's GPU strash, strash.cu:190-314), while the CPU dedupe (`st -c`,
the reference's CPU-side strash that `grw` used by default,
aig_manager.cu:454-522) keeps the input's id order.  On a graph whose strash
merges only duplicates (which keep their levels) the device result is
therefore level-major; after `rw` the waves are those of the composed graph,
which a merge can leave above an output AND's own level.  The CPU dedupe
folds no trivial AND, so every AND it keeps has its input level, and its
levels in id order are those of the input with the dropped ANDs left out.
The exact arrays of both modes are compared with the oracle in the strash
parity tests as well."""
import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu


def and_levels(g):
    nPIs, f0, f1, _ = g
    f0, f1 = np.asarray(f0), np.asarray(f1)
    lev = np.zeros(len(f0), np.int64)
    for v in range(nPIs + 1, len(f0)):
        lev[v] = 1 + max(lev[f0[v] >> 1], lev[f1[v] >> 1])
    return lev[nPIs + 1:]


def level_major(g):
    return bool(np.all(np.diff(and_levels(g)) >= 0))


def test_device_strash_is_wave_major_cpu_dedupe_keeps_id_order(gpu_ctx):
    src = gpu_ctx.gen_geometric(64, 20000, 32, 1000, 0.01, 5)
    g = src.arrays()
    assert not level_major(g)            # the generator's order interleaves levels

    dev = gpu_ctx.from_arrays(*g)
    dev.strash()
    d = dev.arrays()
    want_d, _ = O.strash(*g)
    assert all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(d[1:], want_d[1:]))
    assert level_major(d)

    cpu = gpu_ctx.from_arrays(*g)
    cpu.strash(cpu=True)
    c = cpu.arrays()
    want, _, _ = O.strash_cpu(*g)
    assert all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(c[1:], want[1:])) and c[0] == want[0]
    assert not level_major(c)            # the input's interleaved order survives
    # id order: the kept ANDs appear in the same order as in the input, so the
    # CPU result's levels are the input's levels with the dropped ANDs removed
    lin, lc = and_levels(g), and_levels(c)
    it = iter(lin.tolist())
    assert all(any(x == y for y in it) for x in lc.tolist())
