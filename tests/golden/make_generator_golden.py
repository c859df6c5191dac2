"""Writes tests/golden/generators.json: sha256 of the AIGER files the seeded
generators produce for the benchmark configurations (inputs, not outputs).

Run from the repository root after building libgls.so:
    python tests/golden/make_generator_golden.py
"""
import hashlib
import json
import os
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import culs_amd  # noqa: E402

SPECS = {
    "cfg0_random_20k": dict(kind="random", nPIs=64, nAnds=20000, nPOs=32, window=1000, seed=1),
    "cfg1_mult32_red10": dict(kind="mult", nbits=32, redundant=0.1, seed=2),
    "cfg2_geometric_200k_sample": dict(kind="geometric", nPIs=1024, nAnds=200000, nPOs=4096, window=50000,
                                       p=0.001, seed=3),
    "cfg3_member_1000": dict(kind="random", nPIs=16, nAnds=1000, nPOs=8, window=1000, seed=1000),
}


def main():
    ctx = culs_amd.Context(-1)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, spec in SPECS.items():
            if spec["kind"] == "random":
                a = ctx.gen_random(spec["nPIs"], spec["nAnds"], spec["nPOs"], spec["window"], spec["seed"])
            elif spec["kind"] == "geometric":
                a = ctx.gen_geometric(spec["nPIs"], spec["nAnds"], spec["nPOs"], spec["window"], spec["p"],
                                      spec["seed"])
            else:
                a = ctx.gen_multiplier(spec["nbits"], spec["redundant"], spec["seed"])
            p = os.path.join(d, name + ".aig")
            a.write_aiger(p)
            out[name] = dict(spec, sha256=hashlib.sha256(open(p, "rb").read()).hexdigest(), ands=a.n_ands)
    json.dump(out, open(os.path.join(ROOT, "tests", "golden", "generators.json"), "w"), indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
