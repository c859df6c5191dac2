"""The gpuls shell on the GPU: the script `read; st; rw; rwz; b; ps; write`
gives the same AIG as the same chain through the C ABI, `ps` prints the
reference's line with the final counts (src/aig_manager.cu:604-612), and the
flag spellings of the reference's handlers (`rw -z`, `rewrite`, `st -c`) run
the same passes (src/command_manager.cpp:64-120)."""
import os
import re
import subprocess

import numpy as np
import pytest

import culs_amd

pytestmark = pytest.mark.gpu

GPULS = os.path.join(os.path.dirname(culs_amd.__file__), "gpuls")
PS = re.compile(r"AIG stats: i/o = (\d+)/(\d+) and = (\d+) level = (-?\d+)")


def run(script):
    print("gpuls -c %r" % script, flush=True)
    r = subprocess.run([GPULS, "-c", script], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "failed" not in r.stdout, r.stdout
    return r.stdout


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("spell", ["rw; rwz", "rewrite; rw -z"])
def test_shell_chain_equals_c_abi(gpu_ctx, tmp_path, spell):
    src, out = str(tmp_path / "in.aig"), str(tmp_path / "out.aig")
    a = gpu_ctx.gen_random(64, 20000, 32, 1000, 1)
    a.write_aiger(src)
    so = run("read %s; st; %s; b; ps; write %s" % (src, spell, out))
    a.strash()
    a.rewrite()
    a.rewrite(use_zeros=True)
    a.balance()
    st = a.stats()
    got = PS.findall(so)
    assert len(got) == 1 and tuple(map(int, got[0])) == (st.nPIs, st.nPOs, st.nAnds, st.levels)
    assert "rewrite: (zeros)" in so
    b = gpu_ctx.read_aiger(out)
    assert same(b.arrays(), a.arrays())


def test_shell_cpu_strash_then_rw(gpu_ctx, tmp_path):
    src, out = str(tmp_path / "in.aig"), str(tmp_path / "out.aig")
    a = gpu_ctx.gen_multiplier(8, 0.1, 3)
    a.write_aiger(src)
    run("read %s; st -c; rw; write %s" % (src, out))
    a.strash(cpu=True)
    a.rewrite()
    assert same(gpu_ctx.read_aiger(out).arrays(), a.arrays())


def test_shell_balance_shuffle(gpu_ctx, tmp_path):
    """`b -s` runs the balance with thrust's shuffle tie order, as the
    reference's handler does (command_manager.cpp:61-74)."""
    src, out = str(tmp_path / "in.aig"), str(tmp_path / "out.aig")
    a = gpu_ctx.gen_random(64, 20000, 32, 1000, 5)
    a.write_aiger(src)
    so = run("read %s; st; b -s; write %s" % (src, out))
    assert "without using id as tie break" in so
    a.strash()
    a.balance(shuffle=True)
    assert same(gpu_ctx.read_aiger(out).arrays(), a.arrays())
