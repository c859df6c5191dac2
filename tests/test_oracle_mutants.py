"""Planted mistakes in the oracle must be caught by its pins.

The oracle is built once per mutant (an ORC_MUT_* macro in oracle/oracle.c
plants one mistake) into a temporary directory, and the pin checks of
test_oracle_pins.py run against that build.  Each mutant must make at least
one of the checks it targets fail:

- DOMINANCE: CutFilter tests containment the wrong way round (rewrite.cu:178-193)
- FINDCUT: FindCut evicts the last smallest-value cut instead of the first (:84-109)
- VALUE: CutFindValue without its many-single-fanout-leaves rule (:49-61)
- ADDED: the created-node count also counts library nodes found in the graph (:369-374)
- ZERO: plain rw accepts zero-gain replacements (:429-436)
- REORDER: nodes of one level in descending instead of ascending old id (:1245-1279)
"""
import os
import subprocess

import pytest

import oracle_ffi as O
import test_oracle_pins as P

MUTANTS = {
    "DOMINANCE": [lambda ctx: P.check_dominance(),
                  lambda ctx: [P.check_cut_sets(P.gen(ctx, **s)) for s in P.TINY]],
    "FINDCUT": [lambda ctx: P.check_eviction()],
    "VALUE": [lambda ctx: P.check_eviction()],
    "ADDED": [lambda ctx: [P.check_evaluation(P.gen(ctx, **s), z) for s in P.EVAL for z in (False, True)]],
    "ZERO": [lambda ctx: [P.check_evaluation(P.gen(ctx, **s), False) for s in P.EVAL]],
    "REORDER": [lambda ctx: [P.check_reorder(P.gen(ctx, **s)) for s in P.EVAL]],
}


def build_mutant(name, out_dir):
    path = os.path.join(str(out_dir), "liboracle_%s.so" % name.lower())
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-std=c11", "-DORC_MUT_%s" % name, "-o", path,
                    os.path.join(O.ORACLE_DIR, "oracle.c")], check=True)
    return path


def test_pins_pass_on_the_real_oracle(host_ctx):
    for checks in MUTANTS.values():
        for c in checks:
            c(host_ctx)


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_catch_mutant(host_ctx, tmp_path, name):
    path = build_mutant(name, tmp_path)
    caught = 0
    with O.using(path):
        for c in MUTANTS[name]:
            try:
                c(host_ctx)
            except AssertionError:
                caught += 1
    assert caught > 0, "the pins did not notice mutant %s" % name
