"""out-of-gpu-join: host-resident inputs bucketed on the device through a
pinned host arena and joined bucket by bucket, against the oracle."""
import json

import numpy as np
import pytest

import proteus_amd as pa
from proteus_amd import relbuilder as rb
from oracle import pyoracle as O


def _inputs(nr, ns, seed):
    rk = O.gen_perm_i32(nr, seed)
    rv = O.gen_uniform_i32(nr, seed, 1, 0, 1 << 31)
    sk = O.gen_uniform_i32(ns, seed, 2, 0, int(1.33 * nr))
    sv = O.gen_uniform_i32(ns, seed, 3, 0, 1 << 31)
    return rk, rv, sk, sv


def test_relbuilder_out_of_gpu_join_shape():
    b = rb.RelBuilder()
    R = b.scan("R", ["k", "v"])
    S = b.scan("S", ["k", "v"])
    j = S.out_of_gpu_join(R, [rb.col("R", "k"), rb.col("R", "v")], [rb.col("S", "k")], 16,
                          100, 200).plan()
    assert j["operator"] == "out-of-gpu-join" and j["numOfBuckets"] == 16
    assert j["build_e"][0]["original"] == rb.col("R", "k")
    assert j["probe_e"][0]["original"] == rb.col("S", "k")


@pytest.mark.gpu
@pytest.mark.parametrize("buckets,budget", [(0, 64 << 20), (3, 128 << 20), (1, 0), (0, 16 << 20)])
def test_host_inputs_match_oracle(ctx, buckets, budget):
    nr, ns = 1 << 20, 1 << 23
    rk, rv, sk, sv = _inputs(nr, ns, 21)
    got = ctx.join_out_of_gpu_reduce(rk, rv, sk, sv, num_buckets=buckets, hbm_budget=budget)
    rc, want = O.join_dense(rk, rv, sk, sv)
    assert got == want, (buckets, budget)


@pytest.mark.gpu
def test_duplicates_null_values_empty_and_skew(ctx):
    rs = np.random.default_rng(5)
    rk = rs.integers(-3000, 3000, 40000).astype(np.int32)   # duplicate build keys
    rv = rs.integers(-(1 << 31), 1 << 31, len(rk)).astype(np.int32)
    sk = rs.integers(-4000, 4000, 300000).astype(np.int32)
    sv = rs.integers(-(1 << 31), 1 << 31, len(sk)).astype(np.int32)
    rc, want = O.join_chained(rk, rv, sk, sv, 14)
    assert ctx.join_out_of_gpu_reduce(rk, rv, sk, sv, num_buckets=7, hbm_budget=8 << 20) == want
    # no values: the same as the device join over device columns
    dev = [ctx.to_device(a) for a in (rk, sk)]
    try:
        ref = ctx.join_partitioned_reduce(dev[0], None, dev[1], None)
    finally:
        for d in dev:
            d.free()
    assert ctx.join_out_of_gpu_reduce(rk, None, sk, None, num_buckets=5, hbm_budget=8 << 20) == ref
    # an empty side gives the identities
    empty = ctx.join_out_of_gpu_reduce(rk[:0], rv[:0], sk, sv)
    assert empty["count"] == 0 and empty["min_b"] == np.iinfo(np.int64).max
    # one key everywhere cannot be split by buckets: capacity error, not a wrong answer
    one = np.zeros(1 << 20, np.int32)
    with pytest.raises(pa.PjError) as e:
        ctx.join_out_of_gpu_reduce(one[:4], None, one, None, num_buckets=4, hbm_budget=8 << 20)
    assert e.value.code == pa.PJ_E_CAPACITY


@pytest.mark.gpu
def test_plan_operator(ctx):
    nr, ns = 1 << 14, 1 << 18
    rk, rv, sk, sv = _inputs(nr, ns, 8)
    ctx.clear_catalog()
    for rel, k, v in (("R", rk, rv), ("S", sk, sv)):
        ctx.register_column(rel, "k", "int", k)
        ctx.register_column(rel, "v", "int", v)
    b = rb.RelBuilder(ctx)
    R = b.scan("R", ["k", "v"])
    S = b.scan("S", ["k", "v"])
    j = S.out_of_gpu_join(R, [rb.col("R", "k"), rb.col("R", "v")],
                          [rb.col("S", "k"), rb.col("S", "v")], 16, nr, ns)
    tot = rb.add(rb.cast(rb.col("R", "v"), "int64"), rb.cast(rb.col("S", "v"), "int64"))
    q = j.reduce([rb.int64(1), tot], ["sum", "sum"])
    got = [int(x) for x in str(q.prepare().execute()).strip().split(",")]
    rc, want = O.join_dense(rk, rv, sk, sv)
    assert got == [want["count"], want["sum_b"] + want["sum_p"]]
    ctx.clear_catalog()


@pytest.mark.gpu
def test_bucket_limits(ctx):
    """One device bucketing pass makes at most 256 buckets: 300 explicit
    buckets are an argument error, and a budget that would need more than
    256 buckets a capacity error, both before any work; 256 buckets run."""
    nr, ns = 1 << 14, 1 << 16
    rk, rv, sk, sv = _inputs(nr, ns, 4)
    with pytest.raises(pa.PjError) as e:
        ctx.join_out_of_gpu_reduce(rk, rv, sk, sv, num_buckets=300)
    assert e.value.code == pa.PJ_E_ARG and "at most 256" in str(e.value)
    # (2^14 + 2^16) tuples x 48 B / 4 KiB budget: 960 buckets
    with pytest.raises(pa.PjError) as e:
        ctx.join_out_of_gpu_reduce(rk, rv, sk, sv, hbm_budget=4096)
    assert e.value.code == pa.PJ_E_CAPACITY and "buckets" in str(e.value)
    got = ctx.join_out_of_gpu_reduce(rk, rv, sk, sv, num_buckets=256)
    rc, want = O.join_dense(rk, rv, sk, sv)
    assert got == want
