"""Batched rw at full size: 4,096 seeded random AIGs with AND counts
log-uniform in [1,000, 200,000] (PIs = max(16, ANDs/200), POs = PIs/2,
fanins uniform over the PIs and the previous 1,000 ANDs; about 63M ANDs in
all), packed into one device arena and rewritten by one batched pass.  Every
member must be bit-identical to the same pass on that AIG alone, and a
sample of members bit-identical to the oracle's rw."""
import sys

import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu

COUNT = 4096
MAX_ANDS = 200_000


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def members(ctx):
    rng = np.random.default_rng(12345)
    out = []
    for i in range(COUNT):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(MAX_ANDS))))
        pis = max(16, n // 200)
        out.append(ctx.gen_random(pis, n, max(1, pis // 2), 1000, 1000 + i))
    return out


@pytest.mark.parametrize("zeros", [False, True])
def test_batch_4096_members_equal_alone(gpu_ctx, zeros):
    log("batch 4096 zeros=%s: generate" % zeros)
    aigs = members(gpu_ctx)
    gpu_ctx.batch_strash(aigs)
    total = sum(a.n_ands for a in aigs)
    assert total > 50_000_000
    strashed = [a.arrays() for a in aigs]
    # the alone passes start from device copies of the strashed members: an
    # upload of their arrays would be strashed again first, and a second
    # strash removes the nodes the first one left dangling (dangling nodes
    # are removed before hashing, strash.cu:175-188)
    alones = [a.copy() for a in aigs]
    log("batch 4096: %d ANDs after strash; batched rw" % total)
    gpu_ctx.batch_rewrite(aigs, use_zeros=zeros)
    log("batch 4096: each member alone")
    for k, (alone, a) in enumerate(zip(alones, aigs)):
        alone.rewrite(use_zeros=zeros)
        got, want = a.arrays(), alone.arrays()
        alone.free()
        assert len(got[1]) == len(want[1]), "member %d: %d vs %d objects" % (k, len(got[1]), len(want[1]))
        assert all(np.array_equal(x, y) for x, y in zip(got, want)), "member %d differs from its pass alone" % k
    log("batch 4096: oracle sample")
    for k in np.linspace(0, COUNT - 1, 6).astype(int):
        want, _ = O.rw(*strashed[k], zeros)
        assert all(np.array_equal(x, y) for x, y in zip(aigs[k].arrays(), want)), "member %d differs from the oracle" % k
    for a in aigs:
        a.free()
