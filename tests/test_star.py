"""Star-schema plan (four-way join chain -> group-by) through the plan
executor: the fused probe pipeline and operator-at-a-time execution both
equal the numpy oracle; builds with duplicate keys fall back correctly."""
import json
import os

import numpy as np
import pytest

from proteus_amd import relbuilder as rb
from proteus_amd import star
from tests import helpers as H


def _rows(text):
    return [[int(x) for x in ln.split(",")] for ln in text.strip().splitlines()]


@pytest.mark.gpu
@pytest.mark.parametrize("pipeline", [True, False])
def test_star_plan_matches_oracle(ctx, pipeline):
    ctx.clear_catalog()
    s = star.Star(ctx, 1 << 23, seed=3)
    want = H.star_expected(s)
    assert len(want) == 800
    old = os.environ.get("PROTEUS_AMD_NO_PIPELINE")
    try:
        os.environ["PROTEUS_AMD_NO_PIPELINE"] = "0" if pipeline else "1"
        got = _rows(ctx.execute_plan(json.dumps(s.plan())))
    finally:
        if old is None:
            os.environ.pop("PROTEUS_AMD_NO_PIPELINE", None)
        else:
            os.environ["PROTEUS_AMD_NO_PIPELINE"] = old
    assert got == want
    s.close()


@pytest.mark.gpu
def test_pipeline_with_filters_projects_and_duplicate_builds(ctx):
    ctx.clear_catalog()
    rs = np.random.default_rng(5)
    n = 50000
    f_a = rs.integers(0, 300, n).astype(np.int32)
    f_b = rs.integers(0, 40, n).astype(np.int32)
    f_v = rs.integers(-1000, 1000, n).astype(np.int32)
    a_key = np.arange(250, dtype=np.int32)          # unique, some fks miss
    a_w = (a_key * 7 % 13).astype(np.int32)
    b_key = np.concatenate([np.arange(40), np.arange(10)]).astype(np.int32)  # duplicates
    b_w = np.arange(50, dtype=np.int32)
    for rel, cols in (("F", {"a": f_a, "b": f_b, "v": f_v}), ("A", {"k": a_key, "w": a_w}),
                      ("B", {"k": b_key, "w": b_w})):
        for name, v in cols.items():
            ctx.register_column(rel, name, "int", v)
    b = rb.RelBuilder(ctx)
    F = b.scan("F", ["a", "b", "v"]).to_gpu().unpack()
    A = b.scan("A", ["k", "w"]).to_gpu().unpack()
    Bb = b.scan("B", ["k", "w"]).to_gpu().unpack()
    col = rb.col
    # F join A (unique) -> select -> project -> groupby
    q = F.join(A, col("A", "k"), col("F", "a"), 9, 250) \
         .filter(rb.gt(col("F", "v"), col("A", "w"))) \
         .project([rb.as_(rb.mod(col("A", "w"), rb.lit(4)), "p", "g"),
                   rb.as_(rb.add(col("F", "v"), col("A", "w")), "p", "x")], "p") \
         .groupby([col("p", "g")], [(rb.as_(rb.cast(col("p", "x"), "int64"), "o", "s"), "sum"),
                                    (rb.as_(rb.int64(1), "o", "c"), "sum")], 4, 16) \
         .sort([col("p", "g"), col("o", "s"), col("o", "c")], ["ASC", "NONE", "NONE"])
    got = _rows(str(q.prepare().execute()))
    hit = f_a < 250
    w = a_w[np.where(hit, f_a, 0)]
    keep = hit & (f_v > w)
    g = w[keep] % 4
    x = f_v[keep].astype(np.int64) + w[keep]
    want = [[k, int(x[g == k].sum()), int((g == k).sum())] for k in range(4) if (g == k).any()]
    assert got == want
    # F join B (duplicate build keys): pipeline falls back, every match counted
    q2 = F.join(Bb, col("B", "k"), col("F", "b"), 7, 64) \
          .project([rb.as_(col("B", "w"), "p", "w")], "p") \
          .reduce([rb.int64(1), rb.cast(col("p", "w"), "int64")], ["sum", "sum"])
    cnt, sm = (int(x) for x in str(q2.prepare().execute()).strip().split(","))
    mult = np.bincount(b_key, minlength=40)
    wsum = np.bincount(b_key, weights=b_w, minlength=40).astype(np.int64)
    assert cnt == int(mult[f_b].sum())
    assert sm == int(wsum[f_b].sum())
    ctx.clear_catalog()


def test_star_plan_shape():
    p = star.plan()
    s = json.dumps(p)
    assert s.count('"hashjoin-chained"') == 4 and '"groupby"' in s and '"sort"' in s


def _chain_tables(ctx, n, seed):
    """Fact F(a, b, v) and dimensions A (keys 0..999, all present: a join on
    A keeps every fact row) and B (multiples of 3 below 5000: a join on B
    keeps a third), registered on the device."""
    rs = np.random.default_rng(seed)
    cols = {
        ("F", "a"): rs.integers(0, 1000, n).astype(np.int32),
        ("F", "b"): rs.integers(0, 5000, n).astype(np.int32),
        ("F", "v"): rs.integers(-100, 100, n).astype(np.int32),
        ("A", "k"): rs.permutation(1000).astype(np.int32),
        ("B", "k"): rs.permutation(np.arange(0, 5000, 3)).astype(np.int32),
    }
    cols[("A", "w")] = (cols[("A", "k")] % 17).astype(np.int32)
    cols[("B", "w")] = (cols[("B", "k")] % 11).astype(np.int32)
    devs = []
    for (rel, name), v in cols.items():
        d = ctx.to_device(v)
        devs.append(d)
        ctx.register_column(rel, name, "int", d)
    return cols, devs


@pytest.mark.gpu
def test_probe_chain_duplicates_only_among_rejected_rows(ctx):
    """A filtered dimension whose key repeats only in rows the filter
    rejects: the table leaves those rows out, so the fused chain runs (no
    duplicate-key fallback) and joins the kept rows, as numpy says."""
    ctx.clear_catalog()
    n = 120_001
    c, devs = _chain_tables(ctx, n, 9)
    rs = np.random.default_rng(9)
    keys = np.concatenate([np.arange(1000), rs.integers(0, 1000, 500)]).astype(np.int32)
    flag = np.concatenate([np.ones(1000), np.zeros(500)]).astype(np.int32)  # 1 = kept
    wv = (keys % 13).astype(np.int32)
    perm = rs.permutation(len(keys))
    keys, flag, wv = keys[perm], flag[perm], wv[perm]
    for name, v in (("k", keys), ("f", flag), ("w", wv)):
        d = ctx.to_device(v)
        devs.append(d)
        ctx.register_column("D", name, "int", d)
    b = rb.RelBuilder(ctx)
    col = rb.col
    F = b.scan("F", ["a", "b", "v"]).to_gpu().unpack()
    D = b.scan("D", ["k", "f", "w"]).to_gpu().unpack().filter(rb.eq(col("D", "f"), rb.lit(1)))
    q = F.join(D, col("D", "k"), col("F", "a"), 11, 2000) \
         .groupby([col("D", "w")], [(rb.as_(rb.cast(col("F", "v"), "int64"), "o", "s"), "sum"),
                                    (rb.as_(rb.int64(1), "o", "c"), "sum")], 6, 64)
    ctx.prof_enable(True)
    ctx.prof_reset()
    got = sorted(_rows(str(q.prepare().execute())))
    chain = ctx.prof_get("rel_pipeline_chain")[1]
    ctx.prof_enable(False)
    assert chain >= 1
    fa, fv = c[("F", "a")], c[("F", "v")]
    g = fa % 13  # every fact key 0..999 meets its one kept row, whose w is key % 13
    want = sorted([int(k), int(fv[g == k].astype(np.int64).sum()), int((g == k).sum())]
                  for k in np.unique(g))
    assert got == want
    for d in devs:
        d.free()
    ctx.clear_catalog()


@pytest.mark.gpu
def test_probe_chain_dimension_filter_keeps_nothing(ctx):
    """A fused chain whose filtered dimension keeps no row (the filter's
    mask rejects every build row, so its table holds no key): every fact
    row misses and the grouped result is empty, as numpy says."""
    ctx.clear_catalog()
    n = 100_003
    c, devs = _chain_tables(ctx, n, 5)
    b = rb.RelBuilder(ctx)
    col = rb.col
    F = b.scan("F", ["a", "b", "v"]).to_gpu().unpack()
    A = b.scan("A", ["k", "w"]).to_gpu().unpack().filter(rb.gt(col("A", "w"), rb.lit(100)))
    B = b.scan("B", ["k", "w"]).to_gpu().unpack()
    q = F.join(B, col("B", "k"), col("F", "b"), 13, 2000) \
         .join(A, col("A", "k"), col("F", "a"), 11, 1000) \
         .groupby([col("A", "w")], [(rb.as_(rb.int64(1), "o", "c"), "sum")], 6, 64)
    assert _rows(str(q.prepare().execute())) == []
    for d in devs:
        d.free()
    ctx.clear_catalog()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["all_survive_first", "filter_first", "single_join",
                                  "join_filter_join", "join_join_filter"])
def test_probe_chain_stage0_compaction(ctx, case):
    """The fused probe chain (the rel_pipeline_chain kernel) when stage 0
    keeps every row (each wave's queue holds its whole tile and the later
    stages run in several rounds), when stage 0 is a filter on a fact column
    (no build row to carry), with a single join, and with three stages whose
    stage 0 is a join: a filter at stage 1 (no build row for the second
    queue), or a join at stage 1 and a filter last, grouped by a column of
    the stage-0 table and summing one of the stage-1 table (both build rows
    come back from the second queue).  300,001 fact rows (a partial last
    tile, and rows left in both queues at the end).  Grouped results equal
    numpy's."""
    ctx.clear_catalog()
    n = 300_001
    c, devs = _chain_tables(ctx, n, 21)
    b = rb.RelBuilder(ctx)
    col = rb.col
    F = b.scan("F", ["a", "b", "v"]).to_gpu().unpack()
    A = b.scan("A", ["k", "w"]).to_gpu().unpack()
    B = b.scan("B", ["k", "w"]).to_gpu().unpack()
    fa, fb, fv = c[("F", "a")], c[("F", "b")], c[("F", "v")]
    aw = np.zeros(1000, np.int64)
    aw[c[("A", "k")]] = c[("A", "w")]
    bw = np.full(5000, -1, np.int64)
    bw[c[("B", "k")]] = c[("B", "w")]
    if case == "all_survive_first":
        j = F.join(A, col("A", "k"), col("F", "a"), 11, 1000) \
             .join(B, col("B", "k"), col("F", "b"), 13, 2000)
        keep = bw[fb] >= 0
        g, x = aw[fa][keep], (fv.astype(np.int64) + bw[fb])[keep]
        gkey, xval = col("A", "w"), rb.add(col("F", "v"), col("B", "w"))
    elif case == "filter_first":
        j = F.filter(rb.gt(col("F", "v"), rb.lit(50))) \
             .join(B, col("B", "k"), col("F", "b"), 13, 2000) \
             .join(A, col("A", "k"), col("F", "a"), 11, 1000)
        keep = (fv > 50) & (bw[fb] >= 0)
        g, x = aw[fa][keep], (fv.astype(np.int64) + bw[fb])[keep]
        gkey, xval = col("A", "w"), rb.add(col("F", "v"), col("B", "w"))
    elif case == "single_join":
        j = F.join(B, col("B", "k"), col("F", "b"), 13, 2000)
        keep = bw[fb] >= 0
        g, x = bw[fb][keep], fv[keep].astype(np.int64)
        gkey, xval = col("B", "w"), col("F", "v")
    elif case == "join_filter_join":
        j = F.join(B, col("B", "k"), col("F", "b"), 13, 2000) \
             .filter(rb.gt(col("F", "v"), rb.lit(-20))) \
             .join(A, col("A", "k"), col("F", "a"), 11, 1000)
        keep = (bw[fb] >= 0) & (fv > -20)
        g, x = bw[fb][keep], (fv.astype(np.int64) + aw[fa])[keep]
        gkey, xval = col("B", "w"), rb.add(col("F", "v"), col("A", "w"))
    else:
        j = F.join(B, col("B", "k"), col("F", "b"), 13, 2000) \
             .join(A, col("A", "k"), col("F", "a"), 11, 1000) \
             .filter(rb.lt(col("F", "v"), rb.lit(60)))
        keep = (bw[fb] >= 0) & (fv < 60)
        g, x = bw[fb][keep], aw[fa][keep]
        gkey, xval = col("B", "w"), col("A", "w")
    q = j.groupby([gkey], [(rb.as_(rb.cast(xval, "int64"), "o", "s"), "sum"),
                           (rb.as_(rb.int64(1), "o", "c"), "sum")], 6, 64) \
         .sort([gkey, col("o", "s"), col("o", "c")], ["ASC", "NONE", "NONE"])
    ctx.prof_enable(True)
    ctx.prof_reset()
    got = _rows(str(q.prepare().execute()))
    chain_launches = ctx.prof_get("rel_pipeline_chain")[1]
    ctx.prof_enable(False)
    want = [[int(k), int(x[g == k].sum()), int((g == k).sum())] for k in np.unique(g)]
    assert got == want
    assert chain_launches >= 1
    ctx.clear_catalog()
    for d in devs:
        d.free()
