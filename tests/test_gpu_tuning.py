"""Pass settings on the GPU: every setting changes only how a pass is
scheduled or how much room its tables have, so rw / rwz / strash / balance
give bit-identical results under each setting at its extremes, single and
batched; and the autotuner returns candidate values that are no slower than
the settings it started from, for the bucket of its inputs."""
import sys

import numpy as np
import pytest

import culs_amd

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


EXTREMES = [
    {"cut_wide": 1},                 # every group with a level of 2+ nodes launches per level
    {"cut_wide": 1 << 30},           # every group three levels per launch
    {"rw_sweep_width": 1},           # the final strash of rw is always the wave sweep
    {"rw_sweep_width": 1 << 30},     # always the incremental strash
    {"strash_table_factor": 8},
    {"strash_wave_batch": 1},
    {"strash_wave_batch": 64},
    {"balance_narrow_max": 1},       # every balance level by the grid-wide phase kernels
    {"balance_narrow_max": 1 << 24}, # every level in the persistent workgroup
]


def members(ctx, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 4 == 3:
            out.append(ctx.gen_multiplier(int(rng.integers(4, 12)), 0.1, 300 + i))
        else:
            n = int(np.exp(rng.uniform(np.log(300), np.log(30000))))
            out.append(ctx.gen_random(max(16, n // 200), n, max(8, n // 400), 1000, 500 + i))
    return out


def run_all(ctx, aigs):
    """rw; rwz on a copy of the first, balance on another, batched rw; rwz and
    balance on copies of all -> list of result arrays."""
    res = []
    x = aigs[0].copy()
    x.strash().rewrite().rewrite(use_zeros=True)
    res.append(x.arrays())
    y = aigs[1].copy()
    y.balance()
    res.append(y.arrays())
    b = [a.copy() for a in aigs]
    ctx.batch_rewrite(b)
    ctx.batch_rewrite(b, use_zeros=True)
    res += [a.arrays() for a in b]
    c = [a.copy() for a in aigs]
    ctx.batch_balance(c)
    res += [a.arrays() for a in c]
    for a in [x, y] + b + c:
        a.free()
    return res


@pytest.mark.parametrize("setting", EXTREMES, ids=lambda d: "%s=%d" % next(iter(d.items())))
def test_results_do_not_depend_on_settings(gpu_ctx, setting):
    log("tuning extreme %s" % setting)
    aigs = members(gpu_ctx, 12, 5)
    want = run_all(gpu_ctx, aigs)
    gpu_ctx.set_tuning(-1, **setting)
    try:
        got = run_all(gpu_ctx, aigs)
    finally:
        gpu_ctx.set_tuning(-1, cut_wide=2048, rw_sweep_width=1024, strash_table_factor=4, strash_wave_batch=4,
                           balance_narrow_max=4096)
    assert len(got) == len(want)
    for k, (g, w) in enumerate(zip(got, want)):
        assert same(g, w), "result %d differs under %s" % (k, setting)


@pytest.mark.parametrize("command", [culs_amd.GLS_TUNE_RW, culs_amd.GLS_TUNE_BALANCE])
def test_autotune_batch(gpu_ctx, command):
    log("autotune command %d" % command)
    aigs = members(gpu_ctx, 16, 9)
    gpu_ctx.batch_strash(aigs)
    before = [a.arrays() for a in aigs]
    objects = 1 + sum(a.stats().nObjs for a in aigs)
    start = gpu_ctx.get_tuning(objects)
    best, best_ms, start_ms = gpu_ctx.autotune(aigs, command, reps=1)
    log("autotune: %s %.2f ms (from %.2f ms)" % (best, best_ms, start_ms))
    assert best_ms <= start_ms
    assert gpu_ctx.get_tuning(objects) == best
    assert best["strash_table_factor"] in (4, 8)
    assert best["strash_wave_batch"] in (1, 2, 4, 8, 16) or best["strash_wave_batch"] == start["strash_wave_batch"]
    # the inputs are unchanged
    assert all(same(a.arrays(), b) for a, b in zip(aigs, before))
    # results under the tuned settings equal those under the defaults
    tuned = [a.copy() for a in aigs]
    gpu_ctx.batch_rewrite(tuned)
    gpu_ctx.set_tuning(objects, **start)
    plain = [a.copy() for a in aigs]
    gpu_ctx.batch_rewrite(plain)
    assert all(same(a.arrays(), b.arrays()) for a, b in zip(tuned, plain))
