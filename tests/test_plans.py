"""Plans in the reference's JSON plan language, run unchanged through
pj_plan_execute: the reference's own plan fixtures against their golden
outputs (tests/olap/tests-plan-parsing.cpp, tests/olap/outputs/), and
synthetic plans built with the RelBuilder against the oracle."""
import json
import os

import numpy as np
import pytest

import proteus_amd as pa
from proteus_amd import relbuilder as rb
from oracle import pyoracle as O
from tests import helpers as H

GOLD = json.load(open(os.path.join(H.GOLDEN, "reference_outputs.json")))
PLANS = os.path.join(H.GOLDEN, "inputs", "plans")


def plan_text(name):
    with open(os.path.join(PLANS, name + ".json")) as f:
        return f.read()


def run(ctx, name):
    return ctx.execute_plan(plan_text(name), H.GOLDEN).strip()


# ---- CPU: plan construction ---------------------------------------------
def test_relbuilder_builds_reference_shapes():
    r = rb.RelBuilder()
    s = r.scan("S", ["k", "v"]).to_gpu().unpack()
    b = r.scan("R", ["k", "v"]).to_gpu().unpack()
    q = s.join(b, rb.col("R", "k"), rb.col("S", "k"), hash_bits=10, max_build=1024) \
         .reduce([rb.as_(rb.int64(1), "agg", "cnt")], ["sum"]).to_cpu().print()
    p = q.plan()
    assert p["operator"] == "print"
    j = p["input"]["input"]["input"]
    assert j["operator"] == "hashjoin-chained"
    assert set(j) >= {"build_input", "probe_input", "build_k", "probe_k", "hash_bits",
                      "maxBuildInputSize"}
    assert j["build_input"]["input"]["input"]["plugin"]["name"] == "R"
    json.dumps(p)


def test_reference_plan_fixtures_parse_as_json():
    n = 0
    for d, _, files in os.walk(PLANS):
        for f in files:
            with open(os.path.join(d, f)) as fh:
                json.load(fh)
            n += 1
    assert n >= 25


# ---- GPU: the reference's plan fixtures -----------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["reduce-scan", "reduce-twofields-scan", "reduce-join",
                                  "reduce-select", "reduce-scan-bin"])
def test_reference_plan_golden(ctx, name):
    assert run(ctx, name) == GOLD[name]["expected"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["reduce-join-record", "reduce-join-record-nonselective"])
def test_reference_join_record_plans(ctx, name):
    # same rows as the golden; the row order of a parallel join is not fixed
    got = json.loads(run(ctx, name))
    want = json.loads(GOLD[name]["expected"])
    key = lambda r: (r["A3"], r["B3"])
    assert sorted(got, key=key) == sorted(want, key=key)


def _tbl(name, cols):
    rows = [ln.rstrip("\n").split("|") for ln in open(os.path.join(H.GOLDEN, "inputs", "tpch", name))
            if ln.strip()]
    return [np.array([int(r[c]) for r in rows]) for c in cols]


@pytest.mark.gpu
def test_reference_tpch_plans(ctx):
    (pk,) = _tbl("part_0.001.tbl", [0])
    (ps,) = _tbl("partsupp_0.001.tbl", [0])
    assert run(ctx, "reduce-scan-part") == str(len(pk))
    assert run(ctx, "reduce-scan-partsupp") == str(len(ps))
    assert run(ctx, "reduce-join-parts") == str(int(np.isin(ps, pk).sum()))


def _register_sailors(ctx):
    s = H.sailors()
    ctx.clear_catalog()
    ctx.register_column("inputs/sailors.csv", "sid", "int", s["sid"])
    ctx.register_column("inputs/sailors.csv", "rating", "int", s["rating"])
    ctx.register_column("inputs/sailors.csv", "age", "float", s["age"].astype(np.float64))
    return s


@pytest.mark.gpu
def test_reference_gpu_plans_on_registered_columns(ctx):
    s = _register_sailors(ctx)
    # reduce-scan-gpu: sum(1), max(sid) -> sailorsScan's count and max
    cnt, mx = run(ctx, "reduce-scan-gpu").split(",")
    want = GOLD["sailorsScan"]["expected"].split(",")
    assert (mx, cnt) == (want[0], want[2])
    # gpu-group-by: sum(sid) group by rating
    rows = [ln.split(",") for ln in run(ctx, "gpu-group-by").splitlines()]
    got = {int(r[0]): int(r[1]) for r in rows}
    want = {}
    for k, v in zip(s["rating"], s["sid"]):
        want[int(k)] = want.get(int(k), 0) + int(v)
    assert got == want
    ctx.clear_catalog()


@pytest.mark.gpu
def test_sailors_relbuilder_golden(ctx):
    """tests-join-sailors.cpp JoinLeft3 and the scans, built with RelBuilder."""
    _register_sailors(ctx)
    r = H.reserves()
    ctx.register_column("inputs/reserves.csv", "sid", "int", r["sid"])
    ctx.register_column("inputs/reserves.csv", "bid", "int", r["bid"])
    b = rb.RelBuilder(ctx, H.GOLDEN)
    sail = b.scan("inputs/sailors.csv", ["sid", "rating", "age"]).to_gpu().unpack()
    res = b.scan("inputs/reserves.csv", ["sid", "bid"]).to_gpu().unpack()
    # scan: max(sid), sum(sid), count
    q = sail.reduce([rb.col("inputs/sailors.csv", "sid"), rb.col("inputs/sailors.csv", "sid"),
                     rb.lit(1)], ["max", "sum", "sum"]).to_cpu()
    assert str(q.prepare().execute()).strip() == GOLD["sailorsScan"]["expected"]
    # select age > 40 (sailorsSel)
    q = sail.filter(rb.gt(rb.col("inputs/sailors.csv", "age"), rb.lit(40.0))) \
            .reduce([rb.col("inputs/sailors.csv", "sid"), rb.col("inputs/sailors.csv", "sid"),
                     rb.lit(1)], ["max", "sum", "sum"])
    assert str(q.prepare().execute()).strip() == GOLD["sailorsSel"]["expected"]
    # JoinLeft3: reserves JOIN sailors, max(sid), sum(sid), count
    for fused in (True, False):
        j = res.join(sail, rb.col("inputs/sailors.csv", "sid"), rb.col("inputs/reserves.csv", "sid"),
                     hash_bits=10, max_build=16)
        e = rb.col("inputs/sailors.csv", "sid")
        # a projection between join and reduce turns the fused path off
        if not fused:
            j = j.project([rb.as_(e, "p", "sid")], "p")
            e = rb.col("p", "sid")
        q = j.reduce([e, e, rb.lit(1)], ["max", "sum", "sum"])
        assert str(q.prepare().execute()).strip() == GOLD["sailorsJoinLeft3"]["expected"], fused
    ctx.clear_catalog()


# ---- GPU: synthetic plans vs the oracle ---------------------------------
def _synthetic(ctx, nr, ns, seed):
    rk = O.gen_perm_i32(nr, seed)
    rv = O.gen_uniform_i32(nr, seed, 1, 0, 1 << 31)
    sk = O.gen_uniform_i32(ns, seed, 2, 0, int(1.33 * nr))
    sv = O.gen_uniform_i32(ns, seed, 3, 0, 1 << 31)
    ctx.clear_catalog()
    for rel, k, v in (("R", rk, rv), ("S", sk, sv)):
        ctx.register_column(rel, "k", "int", ctx.to_device(k))
        ctx.register_column(rel, "v", "int", ctx.to_device(v))
    return rk, rv, sk, sv


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["hashjoin-chained", "partitioned-hashjoin-chained"])
def test_join_reduce_plan_matches_oracle(ctx, kind):
    nr, ns = 1 << 16, 1 << 20
    rk, rv, sk, sv = _synthetic(ctx, nr, ns, 5)
    rc, want = O.join_dense(rk, rv, sk, sv)
    b = rb.RelBuilder(ctx)
    R = b.scan("R", ["k", "v"]).to_gpu().unpack()
    S = b.scan("S", ["k", "v"]).to_gpu().unpack()
    if kind == "hashjoin-chained":
        j = S.join(R, rb.col("R", "k"), rb.col("S", "k"), hash_bits=17, max_build=nr)
    else:
        j = S.partitioned_join(R, rb.col("R", "k"), rb.col("S", "k"), nr, ns)
    tot = rb.add(rb.cast(rb.col("R", "v"), "int64"), rb.cast(rb.col("S", "v"), "int64"))
    q = j.reduce([rb.int64(1), tot, rb.col("R", "v"), rb.col("S", "v")],
                 ["sum", "sum", "max", "min"])
    got = [int(x) for x in str(q.prepare().execute()).strip().split(",")]
    assert got == [want["count"], want["sum_b"] + want["sum_p"], want["max_b"], want["min_p"]]
    # int32 sum wraps like the reference's i32 accumulator
    q32 = j.reduce([rb.add(rb.col("R", "v"), rb.col("S", "v"))], ["sum"])
    w32 = (want["sum_b"] + want["sum_p"]) & 0xffffffff
    w32 = w32 - (1 << 32) if w32 >= 1 << 31 else w32
    assert int(str(q32.prepare().execute()).strip()) == w32
    ctx.clear_catalog()


@pytest.mark.gpu
def test_materialized_join_filter_groupby_plan(ctx):
    """join -> project -> select -> groupby -> sort, the non-fused path."""
    nr, ns = 1 << 12, 1 << 16
    rk, rv, sk, sv = _synthetic(ctx, nr, ns, 9)
    b = rb.RelBuilder(ctx)
    R = b.scan("R", ["k", "v"]).to_gpu().unpack()
    S = b.scan("S", ["k", "v"]).to_gpu().unpack()
    g = rb.mod(rb.col("R", "k"), rb.lit(13))
    q = S.join(R, rb.col("R", "k"), rb.col("S", "k"), hash_bits=13, max_build=nr) \
         .project([rb.as_(g, "p", "g"), rb.as_(rb.col("S", "v"), "p", "sv"),
                   rb.as_(rb.col("R", "v"), "p", "rv")], "p") \
         .filter(rb.gt(rb.col("p", "sv"), rb.lit(1 << 30))) \
         .groupby([rb.col("p", "g")],
                  [(rb.as_(rb.int64(1), "g", "cnt"), "sum"),
                   (rb.as_(rb.cast(rb.col("p", "rv"), "int64"), "g", "s"), "sum"),
                   (rb.as_(rb.col("p", "sv"), "g", "mx"), "max")], 4, 64) \
         .sort([rb.col("p", "g"), rb.col("g", "cnt"), rb.col("g", "s"), rb.col("g", "mx")],
               ["ASC", "NONE", "NONE", "NONE"])
    rows = [[int(x) for x in r] for r in q.prepare().execute().rows()]
    # expected with numpy
    pos = np.full(nr, -1, np.int64)
    pos[rk] = np.arange(nr)
    hit = (sk < nr)
    bi = pos[sk[hit]]
    svh = sv[hit].astype(np.int64)
    keep = svh > (1 << 30)
    gk = rk[bi][keep] % 13
    rvv = rv[bi][keep].astype(np.int64)
    want = []
    for k in range(13):
        m = gk == k
        if m.any():
            want.append([k, int(m.sum()), int(rvv[m].sum()), int(svh[keep][m].max())])
    assert rows == want
    ctx.clear_catalog()


@pytest.mark.gpu
def test_hash_rearrange_plan_keeps_rows(ctx):
    nr, ns = 1 << 10, 1 << 14
    rk, rv, sk, sv = _synthetic(ctx, nr, ns, 3)
    b = rb.RelBuilder(ctx)
    S = b.scan("S", ["k", "v"]).to_gpu().unpack()
    q = S.pack([rb.col("S", "k"), rb.col("S", "v")], rb.col("S", "k"), 7) \
         .reduce([rb.int64(1), rb.cast(rb.col("S", "v"), "int64"), rb.col("S", "k")],
                 ["sum", "sum", "max"])
    got = [int(x) for x in str(q.prepare().execute()).strip().split(",")]
    assert got == [ns, int(sv.astype(np.int64).sum()), int(sk.max())]
    # the rows themselves: the input rows, grouped by bucket = hashInt32(k)
    # mod 7 in ascending bucket order (the oracle's bucket function)
    q2 = S.pack([rb.col("S", "k"), rb.col("S", "v")], rb.col("S", "k"), 7).print()
    rows = [[int(x) for x in ln.split(",")] for ln in
            str(q2.prepare().execute()).strip().splitlines()]
    assert len(rows) == ns
    assert sorted(map(tuple, rows)) == sorted(zip(sk.tolist(), sv.tolist()))
    buckets = [O.rearrange_bucket(k, 7) for k, _ in rows]
    assert buckets == sorted(buckets)
    assert len(set(buckets)) == 7
    ctx.clear_catalog()


@pytest.mark.gpu
def test_plan_errors(ctx):
    with pytest.raises(pa.PjError) as e:
        ctx.execute_plan("{not json", ".")
    assert e.value.code == pa.PJ_E_PARSE
    with pytest.raises(pa.PjError) as e:
        ctx.execute_plan(json.dumps({"operator": "nest", "input": {}}), ".")
    assert e.value.code == pa.PJ_E_UNSUPPORTED
    with pytest.raises(pa.PjError) as e:
        ctx.execute_plan(json.dumps({"operator": "scan", "plugin": {"name": "nope", "type": "block",
                                                                     "projections": []}}), ".")
    assert e.value.code == pa.PJ_E_PARSE


# ---- GPU: the device sort --------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("n", [1500, 50000])
def test_device_sort_matches_stable_lexsort(ctx, n):
    """sort over int32 (ties, negatives), FP64 (-0.0 and +0.0 equal,
    negatives) and int64 (NONE: carried, ties keep input order); 1500 rows
    take the one-workgroup rank sort, 50000 the radix passes."""
    rs = np.random.default_rng(12)
    a = rs.integers(-50, 50, n).astype(np.int32)
    c = rs.choice(np.array([-2.5, -0.0, 0.0, 1.0, 3.25, -1e300, 7e-300]), n)
    b = np.arange(n, dtype=np.int64) * 7 - 3
    ctx.clear_catalog()
    ctx.register_column("T", "a", "int", a)
    ctx.register_column("T", "c", "float", c)
    ctx.register_column("T", "b", "int64", b)
    r = rb.RelBuilder(ctx).scan("T", ["a", "c", "b"]).to_gpu().unpack()
    q = r.sort([rb.col("T", "a"), rb.col("T", "c"), rb.col("T", "b")], ["ASC", "DESC", "NONE"])
    rows = q.print([rb.col("T", "a"), rb.col("T", "c"), rb.col("T", "b")]).prepare().execute().rows()
    order = np.lexsort((-c, a))  # stable: last key primary
    assert [int(x[0]) for x in rows] == a[order].tolist()
    assert [float(x[1]) for x in rows] == [float(v) for v in c[order]]
    assert [int(x[2]) for x in rows] == b[order].tolist()
    ctx.clear_catalog()


@pytest.mark.gpu
def test_device_sort_dictionary_strings(ctx, tmp_path):
    """string columns sort by their text, not their dictionary codes."""
    names = ["pear", "apple", "fig", "apple", "kiwi", "fig", "banana", "pear", "apple"]
    score = [3, 1, 4, 1, 5, 9, 2, 6, 5]
    (tmp_path / "fruit.csv").write_text("".join("%s;%d\n" % x for x in zip(names, score)))
    ctx.clear_catalog()
    b = rb.RelBuilder(ctx, str(tmp_path))
    r = b.scan_csv("fruit.csv", [("name", "string", 1), ("score", "int", 2)])
    q = r.sort([rb.col("fruit.csv", "name"), rb.col("fruit.csv", "score")], ["DESC", "ASC"]) \
         .print([rb.col("fruit.csv", "name"), rb.col("fruit.csv", "score")])
    rows = [(x[0], int(x[1])) for x in q.prepare().execute().rows()]
    want = sorted(zip(names, score), key=lambda t: t[1])
    want = sorted(want, key=lambda t: t[0], reverse=True)
    assert rows == want
    ctx.clear_catalog()


@pytest.mark.gpu
def test_reduce_with_predicate(ctx):
    """reduce's predicate p: only the rows where it holds are accumulated,
    over a scan and over a join (which then is not fused)."""
    nr, ns = 1 << 12, 1 << 16
    rk, rv, sk, sv = _synthetic(ctx, nr, ns, 13)
    b = rb.RelBuilder(ctx)
    S = b.scan("S", ["k", "v"]).to_gpu().unpack()
    q = S.reduce([rb.int64(1), rb.cast(rb.col("S", "v"), "int64"), rb.col("S", "k")],
                 ["sum", "sum", "max"], pred=rb.lt(rb.col("S", "k"), rb.lit(1000)))
    got = [int(x) for x in str(q.prepare().execute()).strip().split(",")]
    m = sk < 1000
    assert got == [int(m.sum()), int(sv[m].astype(np.int64).sum()), int(sk[m].max())]
    R = b.scan("R", ["k", "v"]).to_gpu().unpack()
    j = S.join(R, rb.col("R", "k"), rb.col("S", "k"), hash_bits=13, max_build=nr)
    qj = j.reduce([rb.int64(1), rb.cast(rb.col("R", "v"), "int64")], ["sum", "sum"],
                  pred=rb.gt(rb.col("S", "v"), rb.lit(1 << 30)))
    got = [int(x) for x in str(qj.prepare().execute()).strip().split(",")]
    pos = np.full(nr, -1, np.int64)
    pos[rk] = np.arange(nr)
    hit = (sk < nr) & (sv > (1 << 30))
    assert got == [int(hit.sum()), int(rv[pos[sk[hit]]].astype(np.int64).sum())]
    ctx.clear_catalog()


def _gb_plan(ctx, n, groups, dist, seed):
    """A groupby plan over registered device columns C3(k int32, v int64,
    w float32): COUNT, SUM/MIN/MAX(v), SUM(w) per k; returns the plan text,
    the host columns and the device arrays."""
    if dist == "uniform":
        k = ctx.gen_uniform_i32(n, seed, 4, 0, groups)
    else:
        k = ctx.gen_zipf_i32(n, seed, 4, groups, 0.8, seed)
    v = ctx.gen_uniform_i64(n, seed, 5, 0, 1 << 20)
    w = ctx.gen_unit_f32(n, seed, 6)
    ctx.clear_catalog()
    ctx.register_column("C3", "k", "int", k)
    ctx.register_column("C3", "v", "int64", v)
    ctx.register_column("C3", "w", "float32", w)
    b = rb.RelBuilder(ctx)
    q = b.scan("C3", ["k", "v", "w"]).to_gpu().unpack() \
         .groupby([rb.col("C3", "k")],
                  [(rb.as_(rb.int64(1), "g", "cnt"), "sum"),
                   (rb.as_(rb.col("C3", "v"), "g", "s"), "sum"),
                   (rb.as_(rb.col("C3", "v"), "g", "mn"), "min"),
                   (rb.as_(rb.col("C3", "v"), "g", "mx"), "max"),
                   (rb.as_(rb.col("C3", "w"), "g", "fs"), "sum")],
                  max(1, (groups - 1).bit_length()), groups)
    return json.dumps(q.plan()), (k, v, w)


@pytest.mark.gpu
@pytest.mark.parametrize("groups,dist", [(1 << 10, "uniform"), (1 << 16, "zipf")])
def test_groupby_plan_runs_lds_kernel(ctx, groups, dist):
    """A one-int32-key groupby plan with COUNT / SUM / MIN / MAX of an int64
    column and the FP64 SUM of a float32 column runs the LDS group-by (no
    global-atomic kernel) and equals the dense oracle: integers bit-exact,
    FP64 sums exact for the u/2^24 inputs (1e-9 relative written as the
    bound)."""
    n = 1 << 20
    text, (k, v, w) = _gb_plan(ctx, n, groups, dist, 31)
    ctx.prof_enable(True)
    ctx.prof_reset()
    out = ctx.execute_plan(text)
    lds = ctx.prof_get("gb_agg")[1]
    generic = ctx.prof_get("rel_groupby")[1]
    ctx.prof_enable(False)
    assert lds > 0 and generic == 0
    rows = [ln.split(",") for ln in out.splitlines() if ln]
    rc, d = O.groupby_dense(k.numpy(), v.numpy(), w.numpy(), groups)
    assert rc == O.OR_OK
    present = np.nonzero(d["cnt"])[0]
    assert len(rows) == len(present)
    got = {int(r[0]): r[1:] for r in rows}
    for g in present:
        r = got[int(g)]
        assert [int(x) for x in r[:4]] == [int(d["cnt"][g]), int(d["sum"][g]), int(d["min"][g]),
                                           int(d["max"][g])]
        assert abs(float(r[4]) - d["fsum"][g]) <= 1e-9 * abs(d["fsum"][g])
    for x in (k, v, w):
        x.free()
    ctx.clear_catalog()


@pytest.mark.gpu
@pytest.mark.parametrize("na,nb", [(12, 10), (50, 100)])
def test_groupby_two_keys_integer_aggregates_lds_tables(ctx, na, nb):
    """A two-key groupby with integer COUNT / SUM / MIN / MAX (signed values)
    runs the workgroup-private LDS tables of the general group-by
    (rel_groupby_lds): 120 groups fit every workgroup's table; 5000 groups
    overflow the local tables, whose rows go to the global table.  Equals
    numpy."""
    n = 300_003
    rs = np.random.default_rng(na * nb)
    a = rs.integers(0, na, n).astype(np.int32)
    b_ = rs.integers(-nb // 2, nb - nb // 2, n).astype(np.int32)
    v = rs.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    dev = [ctx.to_device(x) for x in (a, b_, v)]
    ctx.clear_catalog()
    ctx.register_column("G", "a", "int", dev[0])
    ctx.register_column("G", "b", "int", dev[1])
    ctx.register_column("G", "v", "int64", dev[2])
    q = rb.RelBuilder(ctx).scan("G", ["a", "b", "v"]).to_gpu().unpack() \
        .groupby([rb.col("G", "a"), rb.col("G", "b")],
                 [(rb.as_(rb.int64(1), "g", "cnt"), "sum"),
                  (rb.as_(rb.col("G", "v"), "g", "s"), "sum"),
                  (rb.as_(rb.col("G", "v"), "g", "mn"), "min"),
                  (rb.as_(rb.col("G", "v"), "g", "mx"), "max")],
                 13, 8192)
    ctx.prof_enable(True)
    ctx.prof_reset()
    out = ctx.execute_plan(json.dumps(q.plan()))
    lds = ctx.prof_get("rel_groupby_lds")[1]
    ctx.prof_enable(False)
    assert lds == 1
    got = sorted(tuple(int(x) for x in ln.split(",")) for ln in out.splitlines() if ln)
    uniq, inv = np.unique(np.stack([a, b_], axis=1), axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    g = len(uniq)
    cnt = np.bincount(inv, minlength=g)
    sm = np.zeros(g, np.int64)
    np.add.at(sm, inv, v)
    mn = np.full(g, np.iinfo(np.int64).max)
    np.minimum.at(mn, inv, v)
    mx = np.full(g, np.iinfo(np.int64).min)
    np.maximum.at(mx, inv, v)
    want = sorted((int(uniq[i, 0]), int(uniq[i, 1]), int(cnt[i]), int(sm[i]), int(mn[i]), int(mx[i]))
                  for i in range(g))
    assert got == want
    for x in dev:
        x.free()
    ctx.clear_catalog()


@pytest.mark.gpu
def test_groupby_boolean_aggregates_lds_tables(ctx):
    """AND / OR aggregates with a count through the LDS group-by at 700001
    rows: 342 workgroups, whose groups merge into 11 copies of the global
    accumulators that the emit folds with each aggregate's monoid.
    Even groups hold no v = 0 (AND(v > 0) true) and odd groups no v = 999
    (OR(v > 998) false).  Equals numpy."""
    n, ng = 700_001, 40
    rs = np.random.default_rng(7)
    a = rs.integers(0, ng, n).astype(np.int32)
    v = rs.integers(0, 1000, n).astype(np.int32)
    v[(a % 2 == 0) & (v == 0)] = 1
    v[(a % 2 == 1) & (v == 999)] = 998
    dev = [ctx.to_device(x) for x in (a, v)]
    ctx.clear_catalog()
    ctx.register_column("B", "a", "int", dev[0])
    ctx.register_column("B", "v", "int", dev[1])
    va = rb.col("B", "v")
    q = rb.RelBuilder(ctx).scan("B", ["a", "v"]).to_gpu().unpack() \
        .groupby([rb.col("B", "a")],
                 [(rb.as_(rb.int64(1), "g", "cnt"), "sum"),
                  (rb.as_(rb.gt(va, rb.lit(0)), "g", "all_pos"), "and"),
                  (rb.as_(rb.gt(va, rb.lit(998)), "g", "any_top"), "or")],
                 10, 1024)
    ctx.prof_enable(True)
    ctx.prof_reset()
    out = ctx.execute_plan(json.dumps(q.plan()))
    lds = ctx.prof_get("rel_groupby_lds")[1]
    ctx.prof_enable(False)
    assert lds == 1
    got = sorted(tuple(ln.split(",")) for ln in out.splitlines() if ln)
    want = sorted((str(g), str(int((a == g).sum())),
                   "true" if (v[a == g] > 0).all() else "false",
                   "true" if (v[a == g] > 998).any() else "false") for g in range(ng))
    assert {w[2] for w in want} == {"true", "false"} and {w[3] for w in want} == {"true", "false"}
    assert got == want
    for x in dev:
        x.free()
    ctx.clear_catalog()


@pytest.mark.gpu
def test_groupby_plan_falls_back_to_general_kernel(ctx):
    """Shapes the LDS group-by does not compute (two integer value columns,
    an FP64 value, two keys) run the general kernel with the same results as
    the LDS path on the shared aggregates."""
    n, groups = 1 << 16, 1 << 8
    text, (k, v, w) = _gb_plan(ctx, n, groups, "uniform", 7)
    lds_rows = sorted(ln for ln in ctx.execute_plan(text).splitlines() if ln)
    import os as _os
    _os.environ["PROTEUS_AMD_NO_LDS_GROUPBY"] = "1"
    try:
        ctx.prof_enable(True)
        ctx.prof_reset()
        gen_rows = sorted(ln for ln in ctx.execute_plan(text).splitlines() if ln)
        assert ctx.prof_get("rel_groupby")[1] > 0 and ctx.prof_get("gb_agg")[1] == 0
        ctx.prof_enable(False)
    finally:
        del _os.environ["PROTEUS_AMD_NO_LDS_GROUPBY"]
    assert [r.split(",")[:5] for r in gen_rows] == [r.split(",")[:5] for r in lds_rows]
    for a, b in zip(gen_rows, lds_rows):
        fa, fb = float(a.split(",")[5]), float(b.split(",")[5])
        assert abs(fa - fb) <= 1e-9 * abs(fb)
    for x in (k, v, w):
        x.free()
    ctx.clear_catalog()
