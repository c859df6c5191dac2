"""C ABI of libgls.so on a host without a GPU: symbols, host-only behaviour,
AIGER I/O, argument validation and the seeded generators.  No compute calls."""
import ctypes
import hashlib
import json
import os
import re

import numpy as np
import pytest

import culs_amd
from culs_amd import GlsError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "generators.json")


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "gls.h")).read()
    return sorted(set(re.findall(r"\b(gls_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = culs_amd.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


def test_host_only_context_refuses_device_work(host_ctx):
    a = host_ctx.gen_random(8, 50, 4, 20, 1)
    assert not host_ctx.has_device
    with pytest.raises(GlsError):
        a.strash()
    with pytest.raises(GlsError):
        a.rewrite()


def test_batch_needs_a_device(host_ctx):
    a = host_ctx.gen_random(8, 50, 4, 20, 1)
    with pytest.raises(GlsError):
        host_ctx.batch_rewrite([a])
    with pytest.raises(GlsError):
        host_ctx.batch_strash([a])


def test_server_needs_a_device(host_ctx):
    try:
        culs_amd.Context(0).close()
        pytest.skip("a GPU is present")
    except GlsError:
        pass
    with pytest.raises(GlsError):
        culs_amd.Server(0)


def test_aiger_round_trip(host_ctx, tmp_path):
    a = host_ctx.gen_multiplier(6, 0.1, 3)
    p = str(tmp_path / "m.aig")
    a.write_aiger(p)
    b = host_ctx.read_aiger(p)
    for x, y in zip(a.arrays(), b.arrays()):
        assert np.array_equal(x, y)
    data = open(p, "rb").read()
    assert data.startswith(b"aig ")
    c = host_ctx.read_aiger_bytes(data)
    assert np.array_equal(c.arrays()[1], a.arrays()[1])


def test_aiger_hand_written_file(host_ctx):
    # and gate: 2 inputs, 1 output = !(a & !b); AIGER lit 2 = a, 4 = b, 6 = and
    body = b"aig 3 2 0 1 1\n7\n" + bytes([6 - 5, 5 - 2])
    a = host_ctx.read_aiger_bytes(body)
    nPIs, f0, f1, pos = a.arrays()
    assert nPIs == 2 and list(f0) == [0, 0, 0, 2] and list(f1) == [0, 0, 0, 5] and list(pos) == [7]
    assert a.ps() == "AIG stats: i/o = 2/1 and = 1 level = -1"


@pytest.mark.parametrize("bad", [b"aag 3 2 0 1 1\n", b"aig 3 2 1 1 1\n", b"aig 4 2 0 1 1\n7\n\x01\x03",
                                 b"aig 3 2 0 1 1\n7\n", b"aig 3 2 0 1 1 0\n7\n\x01\x03",
                                 # headers that promise far more than the file holds: rejected
                                 # before any allocation (terabytes of outputs / 16 GB of ANDs)
                                 b"aig 2147483646 0 0 1099511627775 2147483646\n",
                                 b"aig 2147483646 0 0 0 2147483646\n\x01\x01"])
def test_aiger_rejects_malformed(host_ctx, bad):
    with pytest.raises(GlsError) as e:
        host_ctx.read_aiger_bytes(bad)
    assert "(-4)" in str(e.value)


def test_from_arrays_validates_topological_order(host_ctx):
    f0 = np.array([0, 0, 0, 2, 8], np.uint32)
    f1 = np.array([0, 0, 0, 4, 4], np.uint32)   # node 4 uses itself
    with pytest.raises(GlsError):
        host_ctx.from_arrays(2, f0, f1, np.array([8], np.uint32))


def test_culs_literal_convention(host_ctx):
    # CULS arrays: literal 0 = TRUE.  AND(a, TRUE) with PO = FALSE (CULS 1)
    f0 = np.array([0, 0, 0, 0], np.uint32)
    f1 = np.array([0, 0, 0, 2], np.uint32)
    a = host_ctx.from_arrays(2, f0, f1, np.array([1, 6], np.uint32), convention=culs_amd.GLS_LIT_CULS)
    nPIs, g0, g1, pos = a.arrays()
    assert g0[3] == 1 and g1[3] == 2 and list(pos) == [0, 6]      # AIGER: TRUE = 1, FALSE = 0
    _, c0, c1, cpos = a.arrays(convention=culs_amd.GLS_LIT_CULS)
    assert c0[3] == 0 and list(cpos) == [1, 6]


def sim(g, nwords=2, seed=0):
    nPIs, f0, f1, pos = g
    rng = np.random.default_rng(seed)
    v = np.zeros((len(f0), nwords), np.uint64)
    v[1:nPIs + 1] = rng.integers(0, 2**63, size=(nPIs, nwords), dtype=np.uint64)
    F = np.uint64(0xFFFFFFFFFFFFFFFF)
    for i in range(nPIs + 1, len(f0)):
        v[i] = (v[f0[i] >> 1] ^ (F if f0[i] & 1 else 0)) & (v[f1[i] >> 1] ^ (F if f1[i] & 1 else 0))
    return v


@pytest.mark.parametrize("nbits,red", [(4, 0.0), (8, 0.1), (12, 0.3)])
def test_multiplier_multiplies(host_ctx, nbits, red):
    g = host_ctx.gen_multiplier(nbits, red, 5).arrays()
    nPIs, f0, f1, pos = g
    assert nPIs == 2 * nbits and len(pos) == 2 * nbits
    v = sim(g, 1)
    for bit in range(64):
        a = sum(((int(v[1 + i, 0]) >> bit) & 1) << i for i in range(nbits))
        b = sum(((int(v[1 + nbits + i, 0]) >> bit) & 1) << i for i in range(nbits))
        p = 0
        for k, lit in enumerate(pos):
            x = (int(v[lit >> 1, 0]) >> bit) & 1
            p |= (x ^ (lit & 1)) << k
        assert p == a * b


def test_random_generator_structure(host_ctx):
    nPIs, f0, f1, pos = host_ctx.gen_random(64, 5000, 32, 1000, 1).arrays()
    n = len(f0)
    for v in range(nPIs + 1, n):
        a, b = int(f0[v]) >> 1, int(f1[v]) >> 1
        assert a != b and a < v and b < v
        for u in (a, b):
            assert u <= nPIs or v - u <= 1000
    assert list(pos) == [2 * (n - 32 + i) for i in range(32)]


def test_generators_match_golden(host_ctx, tmp_path):
    gold = json.load(open(GOLDEN))
    for name, spec in gold.items():
        kind = spec["kind"]
        if kind == "random":
            a = host_ctx.gen_random(spec["nPIs"], spec["nAnds"], spec["nPOs"], spec["window"], spec["seed"])
        elif kind == "geometric":
            a = host_ctx.gen_geometric(spec["nPIs"], spec["nAnds"], spec["nPOs"], spec["window"], spec["p"],
                                       spec["seed"])
        else:
            a = host_ctx.gen_multiplier(spec["nbits"], spec["redundant"], spec["seed"])
        p = str(tmp_path / (name + ".aig"))
        a.write_aiger(p)
        assert hashlib.sha256(open(p, "rb").read()).hexdigest() == spec["sha256"], name
