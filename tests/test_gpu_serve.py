"""Serving: jobs submitted while others run are batched with them, and every
result equals running the job's script alone."""
import sys
import time

import numpy as np
import pytest

import culs_amd
import oracle_ffi as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


def run_alone(ctx, g, script):
    a = ctx.from_arrays(*g)
    for cmd in [c.strip() for c in script.split(";") if c.strip()]:
        if cmd == "b":
            a.balance()
        elif cmd == "rw":
            a.rewrite()
        elif cmd in ("rwz", "rw -z"):
            a.rewrite(use_zeros=True)
        elif cmd == "st":
            a.strash()
        else:
            raise ValueError(cmd)
    return a.arrays()


def job_mix(ctx, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 2:
            out.append(ctx.gen_multiplier(int(rng.integers(8, 20)), 0.0, 0).arrays())
        else:
            n = int(np.exp(rng.uniform(np.log(1000), np.log(30000))))
            pis = max(16, n // 200)
            out.append(ctx.gen_random(pis, n, max(1, pis // 2), 1000, 2000 + i).arrays())
    return out


def test_serving_results_equal_alone(gpu_ctx):
    log("serving: mixed jobs, arrivals while others run")
    script = "b; rw; rwz; b"
    jobs = job_mix(gpu_ctx, 16, 5)
    srv = culs_amd.Server(0, hbm_budget=8 << 30, max_active=64)
    ids = []
    for k, g in enumerate(jobs):
        ids.append(srv.submit(g, script))
        if k == 7:
            time.sleep(0.05)          # the second half arrives while the first runs
    infos = [srv.wait(j) for j in ids]
    for j, g, info in zip(ids, jobs, infos):
        got = srv.result(j)
        assert info["status"] == 2
        assert same(got, run_alone(gpu_ctx, g, script)), "served result differs from the script alone"
        assert O.sim_cmp(g, got, words=4) == -1
        srv.release(j)
    srv.close()


def test_serving_mixed_scripts_and_budget(gpu_ctx):
    log("serving: different scripts, a tight budget")
    jobs = job_mix(gpu_ctx, 8, 6)
    scripts = ["rw", "st; rw; rwz", "b", "rw; b; rw", "", "rwz", "b; b", "st"]
    srv = culs_amd.Server(0, hbm_budget=64 << 20, max_active=3)   # forces queueing
    ids = [srv.submit(g, s) for g, s in zip(jobs, scripts)]
    for j, g, s in zip(ids, jobs, scripts):
        assert same(srv.result(j), run_alone(gpu_ctx, g, s))
    with pytest.raises(culs_amd.GlsError):
        srv.submit(jobs[0], "rw; frobnicate")
    big = gpu_ctx.gen_random(64, 200000, 32, 1000, 1).arrays()
    with pytest.raises(culs_amd.GlsError):
        srv.submit(big, "rw")           # needs more than the whole budget
    srv.close()
