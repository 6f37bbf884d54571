"""The benchmark workloads at their full sizes against the oracle, on the GPU:
the hash group-by over 2^29 rows (2^10 / 2^16 / 2^22 groups, uniform and
Zipf(0.8) keys), the star-schema plan over 6e8 fact rows, and the 4096-query
served mix.  Integer results are bit-exact; FP64 sums of the u/2^24 float32
inputs are exact (the tolerance written, 1e-9 relative, is the bound)."""
import json

import numpy as np
import pytest

import proteus_amd as pa
from oracle import pyoracle as O
from proteus_amd import star, workload
from tests import helpers as H

pytestmark = pytest.mark.gpu

SEED = 20251203
C3_ROWS = 1 << 29


@pytest.fixture(scope="module")
def c3_values(ctx):
    """The group-by's value columns (int64 v uniform [0, 2^20), float32
    w = u/2^24), on the device and as host copies."""
    v = ctx.gen_uniform_i64(C3_ROWS, SEED, 5, 0, 1 << 20)
    w = ctx.gen_unit_f32(C3_ROWS, SEED, 6)
    yield v, w, v.numpy(), w.numpy()
    v.free()
    w.free()


@pytest.mark.parametrize("groups", [1 << 10, 1 << 16, 1 << 22])
@pytest.mark.parametrize("dist", ["uniform", "zipf0.8"])
def test_groupby_c3_full_size(ctx, c3_values, groups, dist):
    """pj_groupby_i32 over 2^29 rows: COUNT, SUM, MIN, MAX of v bit-exact and
    SUM of w exact, per group, against the dense oracle."""
    v, w, vh, wh = c3_values
    if dist == "uniform":
        k = ctx.gen_uniform_i32(C3_ROWS, SEED, 4, 0, groups)
    else:
        k = ctx.gen_zipf_i32(C3_ROWS, SEED, 4, groups, 0.8, SEED)
    got = ctx.groupby_i32(k, v, w, groups)
    rc, want = O.groupby_dense(k.numpy(), vh, wh, groups)
    k.free()
    assert rc == O.OR_OK
    keys = got["keys"]
    assert len(keys) == int((want["cnt"] > 0).sum())
    assert len(np.unique(keys)) == len(keys)
    for name in ("cnt", "sum", "min", "max"):
        assert np.array_equal(got[name], want[name][keys]), name
    rel = np.abs(got["fsum"] - want["fsum"][keys]) / np.maximum(np.abs(want["fsum"][keys]), 1e-300)
    assert float(rel.max()) <= 1e-9
    assert int(got["cnt"].sum()) == C3_ROWS


def test_star_c4_full_size(ctx):
    """The star-schema plan over 6e8 fact rows (fused probe pipeline) and
    its bloom-filter variant against the numpy result, computed in slices."""
    ctx.clear_catalog()
    s = star.Star(ctx, 600_000_000, SEED)
    want = H.star_expected(s)
    assert len(want) == 800
    rows = [[int(x) for x in ln.split(",")] for ln in
            ctx.execute_plan(json.dumps(s.plan())).strip().splitlines()]
    assert rows == want
    rows = [[int(x) for x in ln.split(",")] for ln in
            ctx.execute_plan(json.dumps(s.plan(bloom_size=1 << 22))).strip().splitlines()]
    assert rows == want
    s.close()


def _expected_query(mix, i):
    """Query i of the mix on the host: the generators' probe columns, the
    dimension's (key, value), then per group COUNT and SUM(R.value + S.val)
    through the oracle's dense group-by (exact int64)."""
    u, n, groups = int(mix.u[i]), int(mix.n[i]), int(mix.groups[i])
    seed, rng = int(mix.gen_seed[i]), int(mix.fk_range[i])
    dk, dv = (c.numpy() for c in mix.dim_cols[u])
    fk = O.gen_uniform_i32(n, seed, 1, 0, rng)
    val = O.gen_uniform_i32(n, seed, 2, 0, 1 << 31)
    grp = O.gen_uniform_i32(n, seed, 3, 0, groups)
    pos = np.empty(len(dk), np.int64)
    pos[dk] = np.arange(len(dk))
    hit = fk < len(dk)
    b = dv.astype(np.int64)[pos[fk[hit]]]
    rc, d = O.groupby_dense(grp[hit], b + val[hit].astype(np.int64), None, groups)
    assert rc == O.OR_OK
    return d["cnt"], d["sum"]


def test_served_mix_c5_full(ctx):
    """The 4096-query seeded mix (builds 2^10..2^24, probes 8-64x, 2^4..2^16
    groups): every query's result is the same run alone (one launch
    sequence per query), in closed-loop batches and served with Poisson
    arrivals; a seeded sample, the 2^24-row builds among it, equals the
    oracle."""
    nq = 4096
    mix = workload.Mix(ctx, nq, seed=SEED, qps=2000.0)
    budget = 48 << 30
    zeros = np.zeros(nq)
    ctx.serve_run(mix.queries(), zeros, mode=1, slots=3, hbm_budget=budget)
    alone = mix.results()
    ctx.serve_run(mix.queries(), zeros, mode=0, slots=3, hbm_budget=budget)
    batched = mix.results()
    ctx.serve_run(mix.queries(), np.asarray(mix.arrivals), mode=0, slots=3, hbm_budget=budget)
    served = mix.results()
    for i in range(nq):
        for got in (batched[i], served[i]):
            assert np.array_equal(got[0], alone[i][0]) and np.array_equal(got[1], alone[i][1]), i
    rs = np.random.default_rng(SEED)
    big = np.flatnonzero(mix.u == 24)
    sample = sorted(set(rs.choice(nq, 6, replace=False).tolist()) |
                    set(rs.choice(big, 2, replace=False).tolist()))
    for i in sample:
        cnt, sm = _expected_query(mix, i)
        assert np.array_equal(alone[i][0], cnt) and np.array_equal(alone[i][1], sm), i
    mix.close()
