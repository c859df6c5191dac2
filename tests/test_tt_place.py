"""The cut kernels' truth-table variable placement (culs_amd/csrc/tt_place.h:
delta swaps from a 16-entry table) against the reference's adjacent-swap
chains restated in the oracle (tt_stretch / tt_shrink, rewrite.cu:213-241),
exhaustively over every table the kernels can meet:

- placement (the merge): a table over k leaves does not depend on variables
  k..3; for each of the 16 phases (k = its popcount) every such table, with
  the upper 16 bits of the word set in three ways (a complemented fanin table
  carries them), compared on the low 16 bits;
- reduction (the support minimisation): for each phase every table that does
  not depend on the variables outside it.

The header is compiled for the host with g++ (its functions are plain
integer arithmetic) into a helper library in a temporary directory."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as O

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(HERE), "culs_amd", "csrc", "tt_place.h")

HELPER = r"""
#define __host__
#define __device__
#define __forceinline__ inline
#include "%s"
extern "C" void tp_place(int n, const uint32_t *t, const uint8_t *phase, int shrink, uint32_t *out) {
    for (int i = 0; i < n; i++) {
        const uint64_t e = tt_place_entry(phase[i]);
        out[i] = shrink ? tt_place_shrink(t[i], e) : tt_place_stretch(t[i], e);
    }
}
"""

M0 = np.array([0x5555, 0x3333, 0x0F0F, 0x00FF], np.uint32)
M1 = np.array([0xAAAA, 0xCCCC, 0xF0F0, 0xFF00], np.uint32)


@pytest.fixture(scope="module")
def helper(tmp_path_factory):
    d = tmp_path_factory.mktemp("tt_place")
    src, so = d / "tp.cpp", d / "libtp.so"
    src.write_text(HELPER % HEADER)
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", str(so), str(src)], check=True)
    L = ctypes.CDLL(str(so))
    vp = ctypes.c_void_p

    def place(t, phase, shrink):
        t = np.ascontiguousarray(t, np.uint32)
        ph = np.ascontiguousarray(phase, np.uint8)
        out = np.zeros(len(t), np.uint32)
        L.tp_place(ctypes.c_int(len(t)), vp(t.ctypes.data), vp(ph.ctypes.data), ctypes.c_int(int(shrink)),
                   vp(out.ctypes.data))
        return out

    return place


def independent(t, v):
    """t (16-bit tables) does not depend on variable v"""
    return (t & M0[v]) == ((t & M1[v]) >> np.uint32(1 << v))


ALL = np.arange(65536, dtype=np.uint32)


def test_placement_matches_adjacent_swaps(helper):
    total = 0
    for ph in range(16):
        k = bin(ph).count("1")
        ok = np.ones(65536, bool)
        for v in range(k, 4):
            ok &= independent(ALL, v)
        for hi in (0, 0xFFFF0000, 0x5A5A0000):
            t = ALL[ok] | np.uint32(hi)
            phase = np.full(len(t), ph, np.uint8)
            got = helper(t, phase, False) & 0xFFFF
            want = O.tt_place(t, phase, False) & 0xFFFF
            assert np.array_equal(got, want), ph
            total += len(t)
    assert total > 100_000


def test_reduction_matches_adjacent_swaps(helper):
    for ph in range(16):
        ok = np.ones(65536, bool)
        for v in range(4):
            if not (ph >> v) & 1:
                ok &= independent(ALL, v)
        t = ALL[ok]
        phase = np.full(len(t), ph, np.uint8)
        assert np.array_equal(helper(t, phase, True) & 0xFFFF, O.tt_place(t, phase, True) & 0xFFFF), ph


def test_known_placements(helper):
    # x0 (0xAAAA) placed at variable 3: 0xFF00; x0 & x1 placed at 1 and 3:
    # the table of x1 & x3, 0xCC00; and shrinking them back
    assert helper([0xAAAA], [8], False)[0] & 0xFFFF == 0xFF00
    assert helper([0x8888], [0b1010], False)[0] & 0xFFFF == 0xCC00
    assert helper([0xCC00], [0b1010], True)[0] & 0xFFFF == 0x8888
    assert helper([0xFF00], [8], True)[0] & 0xFFFF == 0xAAAA
    # every variable kept: no swap
    assert helper([0x1234], [0xF], False)[0] == 0x1234
