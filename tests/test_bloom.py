"""Bloom filters (bloomfilter_build / _probe / _repack): the oracle pinned
against set-based semi-joins on the masked keys, then the HIP path (C ABI,
plan operators, the fused pipeline) against the oracle."""
import json
import os

import numpy as np
import pytest

import proteus_amd as pa
from proteus_amd import relbuilder as rb
from proteus_amd import star
from tests import helpers as H
from oracle import pyoracle as O


def _keys(rs, n, build=None, frac=0.5):
    """int32 keys over the whole range (negatives included), part of them
    drawn from `build`."""
    k = rs.integers(-(1 << 31), 1 << 31, n, dtype=np.int64).astype(np.int32)
    if build is not None and len(build):
        take = rs.random(n) < frac
        k[take] = rs.choice(build, int(take.sum()))
    return k


def _semijoin(build, probe, size):
    """Row ids of probe whose masked key equals some masked build key: the
    definition of a one-hash filter, written with a Python set."""
    mask = size - 1
    s = {int(b) & mask for b in build.astype(np.int64)}
    return np.array([i for i, k in enumerate(probe.astype(np.int64)) if (int(k) & mask) in s],
                    dtype=np.uint32)


# ---- CPU: the oracle -----------------------------------------------------
@pytest.mark.parametrize("size", [1, 2, 64, 1 << 10, 1 << 16])
def test_oracle_matches_set_semijoin(size):
    rs = np.random.default_rng(size)
    build = _keys(rs, 700)
    probe = _keys(rs, 5000, build, 0.3)
    got = O.bloom_probe(O.bloom_build(build, size), probe)
    assert np.array_equal(got, _semijoin(build, probe, size))


def test_oracle_special_cases():
    rs = np.random.default_rng(1)
    build = _keys(rs, 1000)
    probe = _keys(rs, 3000, build, 0.5)
    # no false negatives: every build key passes, in row order
    f = O.bloom_build(build, 1 << 12)
    assert np.array_equal(O.bloom_probe(f, build), np.arange(len(build), dtype=np.uint32))
    # one entry: everything passes after any build, nothing after an empty one
    assert len(O.bloom_probe(O.bloom_build(build, 1), probe)) == len(probe)
    assert len(O.bloom_probe(O.bloom_build(build[:0], 1), probe)) == 0
    # a filter covering every key of a small range is the exact semi-join
    small_b = rs.integers(0, 1 << 20, 5000).astype(np.int32)
    small_p = rs.integers(0, 1 << 20, 20000).astype(np.int32)
    want = np.flatnonzero(np.isin(small_p, small_b)).astype(np.uint32)
    assert np.array_equal(O.bloom_probe(O.bloom_build(small_b, 1 << 20), small_p), want)
    # negative keys select high entries: -1 & (2^k - 1) = 2^k - 1
    f = O.bloom_build(np.array([-1], np.int32), 16)
    assert f[15] == 1 and f.sum() == 1


def test_relbuilder_bloom_plan_shape():
    b = rb.RelBuilder()
    f = b.scan("F", ["k", "v"]).to_gpu().unpack()
    p = f.bloomfilter_repack(rb.col("F", "k"), 1024, 7, [rb.col("F", "v")]).plan()
    assert p["operator"] == "bloomfilter-repack" and p["filterSize"] == 1024 and p["bloomId"] == 7
    assert p["projections"] == [rb.col("F", "v")] and p["e"] == rb.col("F", "k")
    assert "projections" not in f.bloomfilter_repack(rb.col("F", "k"), 4, 1).plan()
    sp = star.plan(bloom_size=1 << 12)
    text = json.dumps(sp)
    assert text.count("bloomfilter-build") == 2 and text.count("bloomfilter-probe") == 2


# ---- GPU: C ABI against the oracle ---------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("size", [1, 32, 1 << 10, 1 << 20])
@pytest.mark.parametrize("nb", [0, 1, 5000])
def test_probe_matches_oracle(ctx, size, nb):
    rs = np.random.default_rng(size * 7 + nb)
    build = _keys(rs, nb)
    probe = _keys(rs, 100003, build, 0.4)
    db, dp = ctx.to_device(build), ctx.to_device(probe)
    try:
        ctx.bloom_build(11, size, db)
        got = ctx.bloom_probe(11, dp)
        want = O.bloom_probe(O.bloom_build(build, size), probe)
        assert np.array_equal(got, want), (size, nb)
    finally:
        ctx.bloom_release(11)
        db.free()
        dp.free()


@pytest.mark.gpu
def test_full_size_filter_is_exact_semijoin(ctx):
    """2^32 entries: the mask keeps every bit, so the probe is the exact
    semi-join (the oracle's byte array would be 4 GiB)."""
    rs = np.random.default_rng(3)
    build = _keys(rs, 20000)
    probe = _keys(rs, 200000, build, 0.5)
    db, dp = ctx.to_device(build), ctx.to_device(probe)
    try:
        ctx.bloom_build(12, 1 << 32, db)
        got = ctx.bloom_probe(12, dp)
        assert np.array_equal(got, np.flatnonzero(np.isin(probe, build)).astype(np.uint32))
    finally:
        ctx.bloom_release(12)
        db.free()
        dp.free()


@pytest.mark.gpu
def test_repack_rebuild_and_errors(ctx):
    rs = np.random.default_rng(4)
    build = _keys(rs, 3000)
    probe = _keys(rs, 77777, build, 0.25)
    c64 = rs.integers(-(1 << 62), 1 << 62, len(probe), dtype=np.int64)
    c8 = rs.integers(0, 2, len(probe)).astype(np.uint8)
    dev = [ctx.to_device(a) for a in (build, probe, c64, c8)]
    db, dp, d64, d8 = dev
    try:
        ctx.bloom_build(5, 1 << 14, db)
        sel = O.bloom_probe(O.bloom_build(build, 1 << 14), probe)
        got = ctx.bloom_repack(5, dp, [dp, d64, d8])
        for g, w in zip(got, (probe, c64, c8)):
            assert np.array_equal(g, w[sel])
        # a rebuild replaces the filter: only the new keys' entries are set
        ctx.bloom_build(5, 1 << 14, dp)
        assert len(ctx.bloom_probe(5, dp)) == len(probe)
        assert np.array_equal(ctx.bloom_probe(5, db),
                              O.bloom_probe(O.bloom_build(probe, 1 << 14), build))
        # empty probe side
        assert len(ctx.bloom_probe(5, ctx.to_device(np.zeros(0, np.int32)))) == 0
        with pytest.raises(pa.PjError) as e:
            ctx.bloom_probe(99, dp)
        assert e.value.code == pa.PJ_E_ARG
        with pytest.raises(pa.PjError) as e:
            ctx.bloom_build(6, 1000, db)
        assert e.value.code == pa.PJ_E_ARG
        ctx.bloom_release(5)
        with pytest.raises(pa.PjError):
            ctx.bloom_release(5)
    finally:
        for d in dev:
            d.free()


# ---- GPU: plan operators -------------------------------------------------
def _with_pipeline(on, fn):
    old = os.environ.get("PROTEUS_AMD_NO_PIPELINE")
    try:
        os.environ["PROTEUS_AMD_NO_PIPELINE"] = "0" if on else "1"
        return fn()
    finally:
        if old is None:
            os.environ.pop("PROTEUS_AMD_NO_PIPELINE", None)
        else:
            os.environ["PROTEUS_AMD_NO_PIPELINE"] = old


@pytest.mark.gpu
@pytest.mark.parametrize("bloom_size", [1 << 22, 1 << 12])
@pytest.mark.parametrize("pipeline", [True, False])
def test_star_plan_with_bloom_filters(ctx, bloom_size, pipeline):
    """ssb100_bloom's shape: filters built on the filtered dimensions, the
    fact side probed before the joins; a small filter (false positives) must
    give the same result."""
    ctx.clear_catalog()
    s = star.Star(ctx, 1 << 22, seed=6)
    try:
        want = H.star_expected(s)
        text = json.dumps(s.plan(bloom_size=bloom_size))
        got = _with_pipeline(pipeline, lambda: ctx.execute_plan(text))
        assert [[int(x) for x in ln.split(",")] for ln in got.strip().splitlines()] == want
    finally:
        s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("pipeline", [True, False])
def test_repack_plan_under_a_join(ctx, pipeline):
    """bloomfilter-repack keeps the projected attributes of the passing rows;
    a join above makes it a stage of the fused pipeline."""
    ctx.clear_catalog()
    rs = np.random.default_rng(8)
    n = 300000
    fk = rs.integers(0, 5000, n).astype(np.int32)
    fv = rs.integers(-1000, 1000, n).astype(np.int32)
    fw = rs.integers(0, 7, n).astype(np.int32)
    dk = np.arange(5000, dtype=np.int32)
    dg = (dk % 11).astype(np.int32)
    for rel, cols in (("F", {"k": fk, "v": fv, "w": fw}), ("D", {"k": dk, "g": dg})):
        for name, v in cols.items():
            ctx.register_column(rel, name, "int", v)
    b = rb.RelBuilder(ctx)
    col = rb.col
    D = b.scan("D", ["k", "g"]).to_gpu().unpack() \
         .filter(rb.lt(col("D", "g"), rb.lit(3))) \
         .bloomfilter_build(col("D", "k"), 1 << 10, 4)
    F = b.scan("F", ["k", "v", "w"]).to_gpu().unpack() \
         .bloomfilter_repack(col("F", "k"), 1 << 10, 4, [col("F", "k"), col("F", "v")])
    q = F.join(D, col("D", "k"), col("F", "k"), 13, 5000) \
         .reduce([rb.int64(1), rb.cast(col("F", "v"), "int64"), col("D", "g")],
                 ["sum", "sum", "max"])
    got = [int(x) for x in
           _with_pipeline(pipeline, lambda: str(q.prepare().execute())).strip().split(",")]
    keep = (fk % 11) < 3
    assert got == [int(keep.sum()), int(fv[keep].astype(np.int64).sum()), 2]
    # w is not among the repacked attributes
    bad = F.join(D, col("D", "k"), col("F", "k"), 13, 5000).reduce([col("F", "w")], ["sum"])
    with pytest.raises(pa.PjError):
        _with_pipeline(pipeline, lambda: bad.prepare().execute())
    # a probe with another size than the build's is an error
    wrong = b.scan("F", ["k", "v", "w"]).to_gpu().unpack() \
             .bloomfilter_probe(col("F", "k"), 1 << 11, 4) \
             .join(D, col("D", "k"), col("F", "k"), 13, 5000).reduce([rb.int64(1)], ["sum"])
    with pytest.raises(pa.PjError) as e:
        _with_pipeline(pipeline, lambda: wrong.prepare().execute())
    assert e.value.code == pa.PJ_E_ARG
    ctx.clear_catalog()
