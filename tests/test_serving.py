"""Resident dimensions, batched join -> group-by queries and the serving
loop: every query's result is the same alone, in a batch and in the served
mix, and equals the oracle's."""
import numpy as np
import pytest

import proteus_amd as pa
from proteus_amd import workload
from oracle import pyoracle as O


def expected(dk, dv, n, groups, seed, fk_range):
    fk = O.gen_uniform_i32(n, seed, 1, 0, fk_range).astype(np.int64)
    val = O.gen_uniform_i32(n, seed, 2, 0, 1 << 31).astype(np.int64)
    grp = O.gen_uniform_i32(n, seed, 3, 0, groups).astype(np.int64)
    pos = np.full(len(dk), -1, np.int64)
    pos[dk] = np.arange(len(dk))
    hit = fk < len(dk)
    b = dv.astype(np.int64)[pos[fk[hit]]]
    cnt = np.zeros(groups, np.int64)
    s = np.zeros(groups, np.int64)
    np.add.at(cnt, grp[hit], 1)
    np.add.at(s, grp[hit], b + val[hit])
    return cnt, s


def _inputs(ctx, n):
    return [ctx.empty(n, np.int32) for _ in range(3)]


@pytest.mark.gpu
def test_dimension_errors(ctx):
    k = ctx.to_device(np.array([1, 2, 2, 5], np.int32))
    with pytest.raises(pa.PjError) as e:
        ctx.dim_create(k)
    assert e.value.code == pa.PJ_E_ARG
    k2 = ctx.to_device(np.array([1, np.iinfo(np.int32).min], np.int32))
    with pytest.raises(pa.PjError) as e:
        ctx.dim_create(k2)
    assert e.value.code == pa.PJ_E_ARG
    # dense keys: a direct table of one slot per key in the span
    d = ctx.dim_create(ctx.to_device(np.arange(100, dtype=np.int32)))
    assert ctx.dim_info(d) == (100, 100 * 8)
    ctx.dim_destroy(d)
    # sparse keys: open addressing at load <= 1/2
    d = ctx.dim_create(ctx.to_device(np.arange(100, dtype=np.int32) * 3))
    assert ctx.dim_info(d) == (100, 256 * 8)
    ctx.dim_destroy(d)
    # a duplicate in a dense span is still an error
    with pytest.raises(pa.PjError) as e:
        ctx.dim_create(ctx.to_device(np.array([5, 6, 7, 6], np.int32)))
    assert e.value.code == pa.PJ_E_ARG
    # above 2048 keys a dense span keeps the 4-byte values alone; a duplicate
    # that leaves the span as wide as the key count is still caught
    d = ctx.dim_create(ctx.to_device(np.arange(-7, 4993, dtype=np.int32)))
    assert ctx.dim_info(d) == (5000, 5000 * 4)
    ctx.dim_destroy(d)
    k = np.arange(5000, dtype=np.int32)
    k[10] = k[20]
    with pytest.raises(pa.PjError) as e:
        ctx.dim_create(ctx.to_device(k))
    assert e.value.code == pa.PJ_E_ARG


@pytest.mark.gpu
@pytest.mark.parametrize("u,groups", [(10, 16), (10, 1 << 13), (13, 16), (13, 1 << 13),
                                      (11, 2048), (12, 2049)])
def test_query_matches_oracle(ctx, u, groups):
    nb = 1 << u
    dk = ctx.gen_perm_i32(nb, 77 + u)
    dv = ctx.gen_uniform_i32(nb, 77 + u, 7, 0, 1 << 31)
    dim = ctx.dim_create(dk, dv)
    n = 37 * nb + 5
    seed, fk_range = 1234 + u, int(1.33 * nb)
    fk, val, grp = _inputs(ctx, n)
    cnt, sm = ctx.empty(groups, np.int64), ctx.empty(groups, np.int64)
    q = pa.Query(dim, fk.ptr, val.ptr, grp.ptr, n, groups, 0, seed, fk_range, cnt.ptr, sm.ptr)
    wc, ws = expected(dk.numpy(), dv.numpy(), n, groups, seed, fk_range)
    # LDS or HBM dimension table (lds_dim 2048 / -1); more than 2048 groups
    # on the HBM table take the staged aggregation
    for tile, lds in ((0, 2048), (4096, 2048), (0, -1), (4096, -1)):
        ctx.query_run(q, pa.BatchCfg(tile, 0, lds))
        assert np.array_equal(cnt.numpy(), wc), (tile, lds)
        assert np.array_equal(sm.numpy(), ws), (tile, lds)
    # the same columns again, given (no generation)
    q.gen_seed = 0
    ctx.query_run(q)
    assert np.array_equal(cnt.numpy(), wc) and np.array_equal(sm.numpy(), ws)
    for a in (fk, val, grp, cnt, sm):
        a.free()
    ctx.dim_destroy(dim)


@pytest.mark.gpu
def test_staged_groups_with_one_hot_group(ctx):
    """Every probe row in one group of 65536: that group's staged run
    overflows and the rest of its matches fall back to atomics; the result
    is exact, alone and batched."""
    nb = 1 << 14
    dk = ctx.gen_perm_i32(nb, 3)
    dv = ctx.gen_uniform_i32(nb, 3, 7, 0, 1 << 31)
    dim = ctx.dim_create(dk, dv)
    n, groups = 300001, 1 << 16
    fk_h = O.gen_uniform_i32(n, 9, 1, 0, int(1.33 * nb))
    val_h = O.gen_uniform_i32(n, 9, 2, -(1 << 31), 1 << 32)
    grp_h = np.full(n, 4099, np.int32)
    fk, val, grp = (ctx.to_device(a) for a in (fk_h, val_h, grp_h))
    cnt, sm = ctx.empty(groups, np.int64), ctx.empty(groups, np.int64)
    q = pa.Query(dim, fk.ptr, val.ptr, grp.ptr, n, groups, 0, 0, 0, cnt.ptr, sm.ptr)
    pos = np.full(nb, -1, np.int64)
    pos[dk.numpy()] = np.arange(nb)
    hit = fk_h < nb
    b = dv.numpy().astype(np.int64)[pos[fk_h[hit]]]
    wc = np.zeros(groups, np.int64)
    ws = np.zeros(groups, np.int64)
    wc[4099] = int(hit.sum())
    ws[4099] = int((b + val_h[hit].astype(np.int64)).sum())
    ctx.query_run(q, pa.BatchCfg(8192, 0, -1))
    assert np.array_equal(cnt.numpy(), wc) and np.array_equal(sm.numpy(), ws)
    ctx.batch_slots(1)
    ctx.batch_launch(0, [q], pa.BatchCfg(65536, 0, -1))
    ctx.batch_wait(0)
    assert np.array_equal(cnt.numpy(), wc) and np.array_equal(sm.numpy(), ws)
    for a in (fk, val, grp, cnt, sm):
        a.free()
    ctx.dim_destroy(dim)


@pytest.mark.gpu
def test_bad_group_key_is_an_error(ctx):
    dim = ctx.dim_create(ctx.to_device(np.arange(64, dtype=np.int32)))
    fk = ctx.to_device(np.arange(10, dtype=np.int32))
    val = ctx.to_device(np.zeros(10, np.int32))
    grp = ctx.to_device(np.array([0, 1, 2, 3, 9, 0, 1, 2, 3, 0], np.int32))
    cnt, sm = ctx.empty(4, np.int64), ctx.empty(4, np.int64)
    q = pa.Query(dim, fk.ptr, val.ptr, grp.ptr, 10, 4, 0, 0, 0, cnt.ptr, sm.ptr)
    with pytest.raises(pa.PjError) as e:
        ctx.query_run(q)
    assert e.value.code == pa.PJ_E_ARG
    ctx.dim_destroy(dim)


def _alloc_inputs(ctx, mix):
    bufs, qs = [], mix.queries()
    for q in qs:
        b = _inputs(ctx, q.n)
        bufs += b
        q.d_fk, q.d_val, q.d_grp = b[0].ptr, b[1].ptr, b[2].ptr
    return qs, bufs


@pytest.mark.gpu
def test_batch_and_serving_equal_alone(ctx):
    mix = workload.Mix(ctx, 48, seed=11, u_lo=10, u_hi=15, qps=4000)
    qs, bufs = _alloc_inputs(ctx, mix)
    alone = []
    for q in qs:
        ctx.query_run(q)
    alone = mix.results()
    # spot-check three queries against the oracle
    for i in (0, 17, 47):
        u = int(mix.u[i])
        dk, dv = (c.numpy() for c in mix.dim_cols[u])
        wc, ws = expected(dk, dv, int(mix.n[i]), int(mix.groups[i]), int(mix.gen_seed[i]),
                          int(mix.fk_range[i]))
        assert np.array_equal(alone[i][0], wc) and np.array_equal(alone[i][1], ws)
    # two batches on two slots, graph replays included
    ctx.batch_slots(2)
    for rep in range(3):
        ctx.batch_launch(0, qs[:20])
        ctx.batch_launch(1, qs[20:])
        ctx.batch_wait(0)
        ctx.batch_wait(1)
        got = mix.results()
        for i in range(mix.nq):
            assert np.array_equal(got[i][0], alone[i][0]), (rep, i)
            assert np.array_equal(got[i][1], alone[i][1]), (rep, i)
        ctx.sync()  # eager work between replays
        _ = ctx.gen_uniform_i32(1 << 16, rep, 1, 0, 100)
    for b in bufs:
        b.free()
    # served mix (inputs allocated by the server), batched, batched with at
    # most 2^14 probe rows per batch (one query per batch when larger), and
    # one-by-one
    for mode, cap in ((0, 0), (0, 1 << 14), (1, 0)):
        done, st = ctx.serve_run(mix.queries(), mix.arrivals, mode=mode, slots=3,
                                 hbm_budget=64 << 20, max_batch_rows=cap)
        assert all(d >= a for d, a in zip(done, mix.arrivals))
        assert st["max_in_flight_bytes"] <= 64 << 20
        got = mix.results()
        for i in range(mix.nq):
            assert np.array_equal(got[i][0], alone[i][0]), (mode, cap, i)
            assert np.array_equal(got[i][1], alone[i][1]), (mode, cap, i)
    mix.close()


@pytest.mark.gpu
def test_autotune_and_table_placement_keep_results(ctx):
    """The tuned configuration and every LDS-vs-HBM table threshold give the
    results of the default configuration."""
    mix = workload.Mix(ctx, 24, seed=13, u_lo=10, u_hi=14)
    qs, bufs = _alloc_inputs(ctx, mix)
    ctx.batch_slots(1)
    ctx.batch_launch(0, qs, pa.BatchCfg(65536, 2, 0))
    ctx.batch_wait(0)
    want = mix.results()
    tuned = mix.autotune(k=12, max_rows=1 << 20)
    assert tuned["tile_rows"] in (16384, 32768, 65536, 131072, 262144)
    assert tuned["wg_per_cu"] in (1, 2, 4) and tuned["lds_dim"] in (0, -1, 512)
    for cfg in (None, pa.BatchCfg(16384, 1, -1), pa.BatchCfg(32768, 4, 512),
                pa.BatchCfg(4096, 2, 1 << 11)):
        ctx.batch_launch(0, qs, cfg)
        ctx.batch_wait(0)
        got = mix.results()
        for i in range(mix.nq):
            assert np.array_equal(got[i][0], want[i][0]) and np.array_equal(got[i][1], want[i][1])
    for q in qs[:4]:
        ctx.query_run(q, pa.BatchCfg(8192, 1, -1))
    got = mix.results()
    for i in range(4):
        assert np.array_equal(got[i][0], want[i][0]) and np.array_equal(got[i][1], want[i][1])
    for b in bufs:
        b.free()
    mix.close()


@pytest.mark.gpu
@pytest.mark.parametrize("lo,step,groups", [(7, 3, 16), (-50000, 5, 3000), (-(1 << 31) + 1, 1, 16),
                                            (11, 100, 16), (-50000, 97, 3000)])
def test_direct_and_hashed_dimensions(ctx, lo, step, groups):
    """Dimensions of dense keys (step 1), of keys with holes (steps 3 and 5:
    direct tables with empty slots) and of sparse keys (steps 97 and 100:
    open addressing), negative keys and the lowest key above the reserved
    INT32_MIN included, probed alone and batched, against numpy: probe keys
    cover the key span and beyond on both sides."""
    rs = np.random.default_rng(abs(lo) + step)
    nb = 5000
    keys = (lo + step * rs.permutation(nb)).astype(np.int64)
    vals = rs.integers(-(1 << 31), 1 << 31, nb).astype(np.int32)
    dim = ctx.dim_create(ctx.to_device(keys.astype(np.int32)), ctx.to_device(vals))
    span = step * nb
    direct = span - step + 1 <= (1 << 17)  # the hash table of 5000 keys: 2^17 slots
    slot = 4 if step == 1 else 8  # a dense span above 2048 keys: 4-byte values alone
    assert ctx.dim_info(dim)[1] == ((span - step + 1) * slot if direct else (1 << 17) * 8)
    n = 200003
    fk = np.clip(lo - 1000 + rs.integers(0, span + 2000, n), -(1 << 31), (1 << 31) - 1)
    fk[:3] = [-(1 << 31), (1 << 31) - 1, lo]
    fk = fk.astype(np.int32)
    val = rs.integers(-(1 << 31), 1 << 31, n).astype(np.int32)
    grp = rs.integers(0, groups, n).astype(np.int32)
    look = dict(zip(keys.tolist(), vals.astype(np.int64).tolist()))
    hit = np.array([int(k) in look for k in fk.tolist()])
    b = np.array([look.get(int(k), 0) for k in fk.tolist()], np.int64)
    wc = np.zeros(groups, np.int64)
    ws = np.zeros(groups, np.int64)
    np.add.at(wc, grp[hit], 1)
    np.add.at(ws, grp[hit], b[hit] + val[hit].astype(np.int64))
    cols = [ctx.to_device(a) for a in (fk, val, grp)]
    cnt, sm = ctx.empty(groups, np.int64), ctx.empty(groups, np.int64)
    q = pa.Query(dim, cols[0].ptr, cols[1].ptr, cols[2].ptr, n, groups, 0, 0, 0, cnt.ptr, sm.ptr)
    for tile, lds in ((0, 2048), (0, -1), (4096, -1)):
        ctx.query_run(q, pa.BatchCfg(tile, 0, lds))
        assert np.array_equal(cnt.numpy(), wc), (tile, lds)
        assert np.array_equal(sm.numpy(), ws), (tile, lds)
    ctx.batch_slots(1)
    ctx.batch_launch(0, [q, q])  # the same query twice: twice the sums
    ctx.batch_wait(0)
    assert np.array_equal(cnt.numpy(), 2 * wc) and np.array_equal(sm.numpy(), 2 * ws)
    for a in cols + [cnt, sm]:
        a.free()
    ctx.dim_destroy(dim)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_serving_cancel_and_caller_columns(ctx, mode):
    """A host flag set while queries are still arriving cancels the serving
    loop: batches on the device finish, the queries run keep results equal to
    running them alone, the rest report done = -1, and the call raises
    PJ_E_CANCELLED.  Probe columns the caller supplied are left as given; the
    ones the loop placed in its arena are reset."""
    import ctypes
    import threading
    mix = workload.Mix(ctx, 64, seed=17, u_lo=10, u_hi=13, qps=200)
    qs, bufs = _alloc_inputs(ctx, mix)
    for q in qs:
        ctx.query_run(q)
    alone = mix.results()
    # queries 0..7 with caller columns (regenerated from gen_seed), the rest placed
    served = [qs[i] if i < 8 else mix.queries()[i] for i in range(mix.nq)]
    flag = ctypes.c_int32(0)
    timer = threading.Timer(0.1, lambda: setattr(flag, "value", 1))
    timer.start()
    with pytest.raises(pa.PjError) as e:
        ctx.serve_run(served, mix.arrivals, mode=mode, slots=3, hbm_budget=64 << 20, cancel=flag)
    timer.join()
    assert e.value.code == pa.PJ_E_CANCELLED
    done = list(ctx.last_done)
    ran = [i for i, d in enumerate(done) if d >= 0]
    assert 0 < len(ran) < mix.nq
    got = mix.results()
    for i in ran:
        assert done[i] >= mix.arrivals[i]
        assert np.array_equal(got[i][0], alone[i][0]) and np.array_equal(got[i][1], alone[i][1])
    arr = ctx.last_queries
    for i in range(8):
        assert arr[i].d_fk == qs[i].d_fk
    for i in range(8, mix.nq):
        assert not arr[i].d_fk
    for b in bufs:
        b.free()
    mix.close()
