"""The gpuls shell (culs_amd/gpuls): script mode, flag parsing, `ps` format,
`write`, and `st -c`, on the host (no GPU: the shell falls back to a host-only
context, as the reference's read/write/ps work without a device).

Reference behaviour followed: src/main.cpp:56-83 (script mode splits on ';'),
src/command_manager.cpp:254-306 (registry, "not registered, ignored",
"not recognized or in wrong format!" for a handler that returns 1),
src/aig_manager.cu:604-612 (the ps line), :231-267 (write .aig only).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import culs_amd
import oracle_ffi as O

GPULS = os.path.join(os.path.dirname(culs_amd.__file__), "gpuls")
PS = re.compile(r"AIG stats: i/o = (\d+)/(\d+) and = (\d+) level = (-?\d+)")


def run(script, env=None):
    r = subprocess.run([GPULS, "-c", script], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120, env=env)
    return r.returncode, r.stdout, r.stderr


@pytest.fixture
def no_gpu_env():
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = "-1"     # a host-only context even where a GPU exists
    return env


def test_shell_host_commands(host_ctx, tmp_path, no_gpu_env):
    a = host_ctx.gen_random(24, 3000, 12, 300, 9)
    src = str(tmp_path / "in.aig")
    out = str(tmp_path / "out.aig")
    a.write_aiger(src)
    rc, so, _ = run("read %s; ps; st -c; print_stats; write %s; time" % (src, out), no_gpu_env)
    assert rc == 0
    assert so.startswith('Running in script mode: "read')
    stats = PS.findall(so)
    assert len(stats) == 2
    assert tuple(map(int, stats[0])) == (24, 12, 3000, -1)
    g = a.arrays()
    want, _, depth = O.strash_cpu(*g)
    n_and = O.n_ands(want)
    assert tuple(map(int, stats[1])) == (24, 12, n_and, depth)
    assert "Output AIG file saved at path: %s" % out in so
    b = host_ctx.read_aiger(out)
    assert all(np.array_equal(x, y) for x, y in zip(b.arrays(), want))
    assert re.search(r"\{time\} prev cmd: alg [0-9.]+ s, full [0-9.]+ s; total: alg [0-9.]+ s, full [0-9.]+ s\.", so)


def test_shell_format_errors(host_ctx, tmp_path, no_gpu_env):
    a = host_ctx.gen_random(8, 200, 4, 50, 1)
    src = str(tmp_path / "in.aig")
    a.write_aiger(src)
    rc, so, _ = run("read %s; rw -q; st -x; b -z; frobnicate; write %s; ps" % (src, str(tmp_path / "o.txt")),
                    no_gpu_env)
    assert rc == 0                    # the reference's script mode always returns 0
    for cmd in ("rw", "st", "b", "write"):
        assert "Command %s not recognized or in wrong format!" % cmd in so
    assert "Command frobnicate not registered, ignored." in so
    assert "write: only support .aig format!" in so
    assert PS.search(so)


def test_shell_without_aig(no_gpu_env):
    rc, so, _ = run("ps; st -c; rw", no_gpu_env)
    assert rc == 0
    assert "ps: AIG is null!" in so and "strash: AIG is null!" in so and "rewrite: AIG is null!" in so


def test_shell_usage():
    r = subprocess.run([GPULS, "--bogus"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr
