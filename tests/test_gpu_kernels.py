"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same
seeded inputs.  Integer results are bit-exact; FP64 sums of float inputs
agree within 1e-9 relative (and exactly for u/2^24 inputs)."""
import numpy as np
import pytest

import proteus_amd as pa
from oracle import pyoracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

I32MIN, I32MAX = np.iinfo(np.int32).min, np.iinfo(np.int32).max


def _dev(ctx, *arrs):
    return [None if a is None else ctx.to_device(a) for a in arrs]


def _free(*arrs):
    for a in arrs:
        if a is not None:
            a.free()


# ---- generators: device == oracle, bit-exact --------------------------------------
def test_generators_match_oracle(ctx):
    n = 100003
    for dev, host in [
        (ctx.gen_perm_i32(n, 42), O.gen_perm_i32(n, 42)),
        (ctx.gen_uniform_i32(n, 7, 3, -5, 1 << 31), O.gen_uniform_i32(n, 7, 3, -5, 1 << 31)),
        (ctx.gen_uniform_i64(n, 7, 4, 0, 1 << 20), O.gen_uniform_i64(n, 7, 4, 0, 1 << 20)),
        (ctx.gen_unit_f32(n, 7, 5), O.gen_unit_f32(n, 7, 5)),
    ]:
        assert np.array_equal(dev.numpy(), host)
        dev.free()
    # Zipf uses libm (log/exp/expm1/log1p) on both sides; they may round
    # differently in the last bit, which can move a rare draw by one rank.
    for dev, host in [
        (ctx.gen_zipf_i32(n, 3, 1, 1 << 16, 0.8, 9), O.gen_zipf_i32(n, 3, 1, 1 << 16, 0.8, 9)),
        (ctx.gen_zipf_fk_i32(n, 3, 2, 1 << 20, int(1.33 * (1 << 20)), 1.0, 5),
         O.gen_zipf_fk_i32(n, 3, 2, 1 << 20, int(1.33 * (1 << 20)), 1.0, 5)),
    ]:
        assert np.mean(dev.numpy() == host) > 0.9999
        dev.free()


# ---- hashjoin-chained ------------------------------------------------------------------
def _join_inputs(seed, nr, ns, key_range, dup=False):
    rs = np.random.default_rng(seed)
    if dup:
        rk = rs.integers(-key_range, key_range, nr).astype(np.int32)
    else:
        rk = rs.permutation(np.arange(-(nr // 2), nr - nr // 2)).astype(np.int32) * 3
    rv = rs.integers(-2**31, 2**31, nr).astype(np.int32)
    sk = rs.integers(-key_range, key_range, ns).astype(np.int32)
    sv = rs.integers(-2**31, 2**31, ns).astype(np.int32)
    if nr > 2:
        rk[:2] = [I32MIN, I32MAX]
    if ns > 4:
        sk[:4] = [I32MIN, I32MAX, I32MIN, 0]
    return rk, rv, sk, sv


@pytest.mark.parametrize("case", [
    dict(seed=1, nr=1000, ns=5000, kr=3000, dup=False, hb=11),
    dict(seed=2, nr=4000, ns=20000, kr=500, dup=True, hb=6),
    dict(seed=3, nr=1 << 16, ns=1 << 18, kr=1 << 17, dup=False, hb=17),
    dict(seed=4, nr=1, ns=100, kr=2, dup=True, hb=0),
])
def test_join_chained_reduce_parity(ctx, case):
    rk, rv, sk, sv = _join_inputs(case["seed"], case["nr"], case["ns"], case["kr"], case["dup"])
    rc, want = O.join_chained(rk, rv, sk, sv, case["hb"], threads=8)
    assert rc == O.OR_OK
    d = _dev(ctx, rk, rv, sk, sv)
    got = ctx.join_chained_reduce(d[0], d[1], d[2], d[3], case["hb"])
    assert got == want
    got_nv = ctx.join_chained_reduce(d[0], None, d[2], None, case["hb"])
    assert got_nv["count"] == want["count"] and got_nv["sum_b"] == 0
    _free(*d)


def test_join_chained_c1_full_size(ctx):
    """The plain chained join at full size: R = 2^20 permuted keys, S = 2^24 uniform over R's keys;
    COUNT = 2^24 exactly and the int64 checksum equals the dense oracle's."""
    nr, ns, seed = 1 << 20, 1 << 24, 2024
    rk = ctx.gen_perm_i32(nr, seed)
    rv = ctx.gen_uniform_i32(nr, seed, 1, 0, 1 << 31)
    sk = ctx.gen_uniform_i32(ns, seed, 2, 0, nr)
    sv = ctx.gen_uniform_i32(ns, seed, 3, 0, 1 << 31)
    got = ctx.join_chained_reduce(rk, rv, sk, sv, 21, int(nr * 1.1))
    rc, want = O.join_dense(rk.numpy(), rv.numpy(), sk.numpy(), sv.numpy())
    assert rc == O.OR_OK
    assert got == want
    assert got["count"] == ns
    _free(rk, rv, sk, sv)


def test_join_chained_dispatch(ctx):
    """pj_join_chained_reduce tries the direct table first (its kernels run on
    every call and do nothing when the keys are not dense) and keeps its
    result for dense build keys (a permutation of 0..n-1); other keys take the
    chained table at small probe sides and the radix-partitioned join for
    large ones (the C1 size, 2^20 x 2^24, keys spread 3 apart); a dense span
    with a repeated key is found by the direct table and goes to the chained
    table.  Every result equals the chained oracle's."""
    fallbacks = ("join_chained_probe", "pjoin_probe")
    for nr, ns, step, repeat, want_fb in ((1 << 12, 1 << 16, 1, False, set()),
                                          (1 << 12, 1 << 16, 3, False, {"join_chained_probe"}),
                                          (1 << 20, 1 << 24, 3, False, {"pjoin_probe"}),
                                          (1 << 12, 1 << 16, 1, True, {"join_chained_probe"})):
        rs = np.random.default_rng(nr + step + repeat)
        rk = (rs.permutation(nr) * step).astype(np.int32)
        if repeat:
            rk[7] = rk[8]  # the span still holds nr keys' worth of slots, one of them twice
        rv = rs.integers(-2**31, 2**31, nr).astype(np.int32)
        sk = rs.integers(-5, int(1.33 * nr * step), ns).astype(np.int32)
        sv = rs.integers(-2**31, 2**31, ns).astype(np.int32)
        d = _dev(ctx, rk, rv, sk, sv)
        hb = (nr - 1).bit_length() + 1
        ctx.prof_enable(True)
        ctx.prof_reset()
        got = ctx.join_chained_reduce(d[0], d[1], d[2], d[3], hb)
        ran = {n for n in fallbacks if ctx.prof_get(n)[1] > 0}
        direct = ctx.prof_get("join_direct_probe")[1]
        ctx.prof_enable(False)
        assert ran == want_fb and direct == 1, (nr, step, repeat, ran, direct)
        rc, want = O.join_chained(rk, rv, sk, sv, hb, threads=8)
        assert rc == O.OR_OK and got == want, (nr, step, repeat)
        if not repeat:
            assert ctx.join_chained_reduce_noalloc(d[0], d[1], d[2], d[3], hb) == want
        _free(*d)


def test_join_direct_smallest_sizes(ctx):
    """The direct table at its smallest: two dense build keys (INT32_MIN and
    INT32_MIN + 1, then 5 and 6) and probe sides of 1 to 7 rows (below one
    four-row vector); each result equals the chained oracle's."""
    for lo in (I32MIN, 5):
        rk = np.array([lo + 1, lo], np.int32)
        rv = np.array([-7, 11], np.int32)
        for ns in range(1, 8):
            sk = np.array([lo, lo + 1, lo + 2, I32MAX, lo, 0, lo + 1][:ns], np.int32)
            sv = np.arange(ns, dtype=np.int32) * 3 - 4
            rc, want = O.join_chained(rk, rv, sk, sv, 2)
            assert rc == O.OR_OK
            d = _dev(ctx, rk, rv, sk, sv)
            assert ctx.join_chained_reduce(d[0], d[1], d[2], d[3], 2) == want, (lo, ns)
            _free(*d)


@pytest.mark.parametrize("case", [
    dict(seed=5, nr=3000, ns=50_001, lo=-1500, off=0),
    dict(seed=6, nr=1 << 16, ns=1 << 18, lo=I32MAX - (1 << 16) + 1, off=0),
    dict(seed=7, nr=5000, ns=40_003, lo=10, off=1),
])
def test_join_direct_dense_keys(ctx, case):
    """The direct table of pj_join_chained_reduce over dense build keys (a
    permutation of lo .. lo + nr - 1, negative keys, a span ending at
    INT32_MAX): probe keys inside and outside the span, INT32_MIN and
    INT32_MAX among them, a probe side that is not a multiple of 4 rows, and
    columns 4-byte but not 16-byte aligned (one row at a time); with and
    without values.  COUNT, SUMs, MIN and MAX equal the chained oracle's, and
    only the direct join's kernels run."""
    rs = np.random.default_rng(case["seed"])
    nr, ns, lo, off = case["nr"], case["ns"], case["lo"], case["off"]
    rk = (lo + rs.permutation(nr).astype(np.int64)).astype(np.int32)
    rv = rs.integers(-2**31, 2**31, nr).astype(np.int32)
    sk = np.clip(lo - nr // 2 + rs.integers(0, 2 * nr, ns), I32MIN, I32MAX).astype(np.int32)
    sk[:4] = [I32MIN, I32MAX, lo, lo + nr - 1]
    sv = rs.integers(-2**31, 2**31, ns).astype(np.int32)
    hb = (nr - 1).bit_length() + 1
    rc, want = O.join_chained(rk, rv, sk, sv, hb, threads=8)
    assert rc == O.OR_OK
    # each column off elements into its allocation
    pad = [np.concatenate([np.zeros(off, np.int32), a]) for a in (rk, rv, sk, sv)]
    d = _dev(ctx, *pad)
    v = [pa.DeviceArray(ctx, x.ptr + 4 * off, len(a), np.int32) for x, a in zip(d, (rk, rv, sk, sv))]
    ctx.prof_enable(True)
    ctx.prof_reset()
    got = ctx.join_chained_reduce(v[0], v[1], v[2], v[3], hb)
    got_nv = ctx.join_chained_reduce(v[0], None, v[2], None, hb)
    ran = {n: ctx.prof_get(n)[1] for n in ("join_direct_probe", "join_chained_probe", "pjoin_probe")}
    ctx.prof_enable(False)
    assert ran["join_direct_probe"] == 2 and ran["join_chained_probe"] == 0 and ran["pjoin_probe"] == 0
    assert got == want
    assert got_nv["count"] == want["count"] and got_nv["sum_b"] == 0
    assert got_nv["sum_p"] == 0 and got_nv["min_b"] == 0 == got_nv["max_b"]
    _free(*d)


def test_join_chained_single_duplicate(ctx):
    """One duplicated key among 2^16 unique build keys must take the probe's
    all-matches walk (the build reports the duplicate); probes of that key
    count both build rows."""
    nr, ns = 1 << 16, 1 << 18
    rs = np.random.default_rng(77)
    rk = rs.permutation(nr).astype(np.int32)
    rk[-1] = rk[12345]  # one key twice, one key absent
    rv = rs.integers(-2**31, 2**31, nr).astype(np.int32)
    sk = rs.integers(0, nr, ns).astype(np.int32)
    sk[:64] = rk[12345]
    sv = rs.integers(-2**31, 2**31, ns).astype(np.int32)
    rc, want = O.join_chained(rk, rv, sk, sv, 17, threads=8)
    assert rc == O.OR_OK
    d = _dev(ctx, rk, rv, sk, sv)
    assert ctx.join_chained_reduce(d[0], d[1], d[2], d[3], 17) == want
    assert ctx.join_chained_reduce_noalloc(d[0], d[1], d[2], d[3], 17) == want
    _free(*d)


def test_join_chained_edges(ctx):
    e = ctx.to_device(np.zeros(0, np.int32))
    k = ctx.to_device(np.arange(100, dtype=np.int32))
    assert ctx.join_chained_reduce(e, None, k, None, 4)["count"] == 0
    assert ctx.join_chained_reduce(k, None, e, None, 4)["count"] == 0
    with pytest.raises(pa.PjError) as ex:
        ctx.join_chained_reduce(k, None, k, None, 4, max_build=99)
    assert ex.value.code == pa.PJ_E_CAPACITY
    _free(e, k)


def test_join_chained_pairs_parity(ctx):
    rk, rv, sk, sv = _join_inputs(5, 3000, 9000, 800, dup=True)
    rc, ob, op = O.join_pairs(rk, sk, 12, cap=1 << 22)
    assert rc == O.OR_OK
    d = _dev(ctx, rk, sk)
    rc2, n, gb, gp = ctx.join_chained_pairs(d[0], d[1], 12, cap=len(ob) + 10)
    assert rc2 == pa.PJ_OK and n == len(ob)
    got = sorted(zip(gb.numpy()[:n].tolist(), gp.numpy()[:n].tolist()))
    assert got == sorted(zip(ob.tolist(), op.tolist()))
    # truncated output reports the full count and PJ_E_CAPACITY
    rc3, n3, gb3, gp3 = ctx.join_chained_pairs(d[0], d[1], 12, cap=10)
    assert rc3 == pa.PJ_E_CAPACITY and n3 == len(ob)
    _free(*d, gb, gp, gb3, gp3)


def test_sailors_join_left3_through_kernels(ctx):
    """The reference's JoinLeft3 golden (74,383,10) through the HIP join."""
    s, r, b = H.sailors(), H.reserves(), H.boats()
    ds, dr, db = ctx.to_device(s["sid"]), ctx.to_device(r["sid"]), ctx.to_device(b["bid"])
    rc, n, ob, op = ctx.join_chained_pairs(dr, ds, 16, cap=1024, max_build=128)
    ob_, op_ = ob.numpy()[:n], op.numpy()[:n]
    bid = ctx.to_device(r["bid"][ob_])
    sid = s["sid"][op_]
    rc2, n2, ob2, op2 = ctx.join_chained_pairs(db, bid, 16, cap=1024, max_build=128)
    sid2 = sid[op2.numpy()[:n2]].astype(np.int64)
    assert [int(sid2.max()), int(sid2.sum()), len(sid2)] == [74, 383, 10]
    _free(ds, dr, db, ob, op, bid, ob2, op2)


# ---- partitioned-hashjoin-chained ---------------------------------------------------
CFGS = [
    None,                                                        # planner default
    dict(bits1=0, bits2=0, lds_log=13, probe_unit=4096, grid_mul=1),  # default re-derived
    dict(bits1=3, bits2=0, lds_log=11, probe_unit=2048, grid_mul=1),  # single pass
    dict(bits1=4, bits2=3, lds_log=12, probe_unit=8192, grid_mul=1),  # two passes
    dict(bits1=8, bits2=7, lds_log=12, probe_unit=65536, grid_mul=0), # max fan-out
    dict(bits1=7, bits2=8, lds_log=13, probe_unit=65536, grid_mul=0), # fan-out 256 in pass 2
    dict(bits1=1, bits2=0, lds_log=11, probe_unit=1024, grid_mul=2),  # LDS rounds (overflow)
]


@pytest.mark.parametrize("ci", range(len(CFGS)))
@pytest.mark.parametrize("dup", [False, True])
def test_join_partitioned_parity(ctx, ci, dup):
    rk, rv, sk, sv = _join_inputs(10 + ci, 20000, 150000, 30000 if not dup else 3000, dup)
    rc, want = O.join_chained(rk, rv, sk, sv, 16, threads=8)
    d = _dev(ctx, rk, rv, sk, sv)
    got = ctx.join_partitioned_reduce(d[0], d[1], d[2], d[3], cfg=CFGS[ci])
    assert got == want
    _free(*d)


def _shifted(ctx, a):
    """A device copy of `a` that starts one element past a 16-byte boundary
    (the owner, to free, and the view)."""
    d = ctx.to_device(np.concatenate([a[:1], a]))
    return d, pa.DeviceArray(ctx, d.ptr + a.itemsize, len(a), a.dtype)


def test_unaligned_columns_and_ragged_sizes(ctx):
    """Row counts that are not multiples of 4 and columns that start off a
    16-byte boundary take the element-wise load paths of the radix passes and
    of the group-by; the results equal the oracle's."""
    rk, rv, sk, sv = _join_inputs(41, 20001, 150003, 30000, False)
    rc, want = O.join_chained(rk, rv, sk, sv, 16, threads=8)
    d = _dev(ctx, rk, rv, sk, sv)
    assert ctx.join_partitioned_reduce(d[0], d[1], d[2], d[3]) == want
    _free(*d)
    sh = [_shifted(ctx, a) for a in (rk, rv, sk, sv)]
    got = ctx.join_partitioned_reduce(*[v for _, v in sh])
    assert got == want
    _free(*[o for o, _ in sh])
    n = 50001
    k = O.gen_uniform_i32(n, 3, 1, 0, 3000)
    v = O.gen_uniform_i64(n, 3, 2, -(1 << 40), 1 << 41)
    w = O.gen_unit_f32(n, 3, 3)
    sh = [_shifted(ctx, a) for a in (k, v, w)]
    got = ctx.groupby_i32(sh[0][1], sh[1][1], sh[2][1], 4096)
    _check_groups(got, k, v, w)
    _free(*[o for o, _ in sh])


def test_join_partitioned_zipf_and_skewed_build(ctx):
    nr, ns, seed = 1 << 18, 1 << 21, 77
    rk = ctx.gen_perm_i32(nr, seed)
    rv = ctx.gen_uniform_i32(nr, seed, 1, 0, 1 << 31)
    sk = ctx.gen_zipf_fk_i32(ns, seed, 2, nr, int(1.33 * nr), 1.0, seed)
    sv = ctx.gen_uniform_i32(ns, seed, 3, 0, 1 << 31)
    rc, want = O.join_dense(rk.numpy(), rv.numpy(), sk.numpy(), sv.numpy())
    for cfg in (None, dict(bits1=6, bits2=5, lds_log=12, probe_unit=4096, grid_mul=1)):
        assert ctx.join_partitioned_reduce(rk, rv, sk, sv, cfg=cfg) == want
    _free(rk, rv, sk, sv)
    # a build side with one hot key (duplicates overflow any LDS table)
    rs = np.random.default_rng(5)
    bk = rs.integers(0, 1000, 30000).astype(np.int32)
    bk[:20000] = 7
    pk = rs.integers(0, 1000, 5000).astype(np.int32)
    bv = rs.integers(0, 100, 30000).astype(np.int32)
    pv = rs.integers(0, 100, 5000).astype(np.int32)
    rc, want = O.join_chained(bk, bv, pk, pv, 12)
    d = _dev(ctx, bk, bv, pk, pv)
    got = ctx.join_partitioned_reduce(d[0], d[1], d[2], d[3])
    assert got == want
    _free(*d)


def test_join_partitioned_c2_scaled(ctx):
    """The radix join at 1/8 of the benchmark size: R = 2^24 permuted keys, S = 2^27 with ~25%
    misses, uniform and Zipf(1.0); dense oracle, bit-exact."""
    nr, ns, seed = 1 << 24, 1 << 27, 99
    rng = int(1.33 * nr)
    rk = ctx.gen_perm_i32(nr, seed)
    rv = ctx.gen_uniform_i32(nr, seed, 1, 0, 1 << 31)
    sv = ctx.gen_uniform_i32(ns, seed, 3, 0, 1 << 31)
    rkh, rvh, svh = rk.numpy(), rv.numpy(), sv.numpy()
    for variant in ("uniform", "zipf"):
        if variant == "uniform":
            sk = ctx.gen_uniform_i32(ns, seed, 2, 0, rng)
        else:
            sk = ctx.gen_zipf_fk_i32(ns, seed, 2, nr, rng, 1.0, seed)
        got = ctx.join_partitioned_reduce(rk, rv, sk, sv)
        rc, want = O.join_dense(rkh, rvh, sk.numpy(), svh)
        assert rc == O.OR_OK and got == want
        assert 0.74 < got["count"] / ns < 0.76
        sk.free()
    _free(rk, rv, sv)


@pytest.mark.parametrize("variant", ["uniform", "zipf"])
def test_join_partitioned_c2_full_size(ctx, variant):
    """The benchmark's join at its full size, 2^27 x 2^30 (fk over
    1.33 x 2^27, uniform or Zipf(1.0) over R's keys with the same miss
    rate), against the dense oracle on the host: bit-exact."""
    nr, ns, seed = 1 << 27, 1 << 30, 20251203
    rk = ctx.gen_perm_i32(nr, seed)
    rv = ctx.gen_uniform_i32(nr, seed, 1, 0, 1 << 31)
    if variant == "uniform":
        sk = ctx.gen_uniform_i32(ns, seed, 2, 0, int(1.33 * nr))
    else:
        sk = ctx.gen_zipf_fk_i32(ns, seed, 2, nr, int(1.33 * nr), 1.0, seed)
    sv = ctx.gen_uniform_i32(ns, seed, 3, 0, 1 << 31)
    got = ctx.join_partitioned_reduce(rk, rv, sk, sv)
    rc, want = O.join_dense(rk.numpy(), rv.numpy(), sk.numpy(), sv.numpy())
    assert rc == O.OR_OK and got == want
    assert 0.74 < got["count"] / ns < 0.76
    _free(rk, rv, sk, sv)


def test_join_partitioned_capacity_and_empty(ctx):
    k = ctx.to_device(np.arange(5000, dtype=np.int32))
    e = ctx.to_device(np.zeros(0, np.int32))
    with pytest.raises(pa.PjError) as ex:
        ctx.join_partitioned_reduce(k, None, k, None, max_build=4999)
    assert ex.value.code == pa.PJ_E_CAPACITY
    with pytest.raises(pa.PjError) as ex:
        ctx.join_partitioned_reduce(k, None, k, None, max_probe=10)
    assert ex.value.code == pa.PJ_E_CAPACITY
    assert ctx.join_partitioned_reduce(e, None, k, None)["count"] == 0
    assert ctx.join_partitioned_reduce(k, None, e, None)["count"] == 0
    _free(k, e)


def test_join_partitioned_equals_chained(ctx):
    """Two independent GPU algorithms agree on the same inputs."""
    nr, ns, seed = 1 << 20, 1 << 23, 5
    rk = ctx.gen_perm_i32(nr, seed)
    rv = ctx.gen_uniform_i32(nr, seed, 1, 0, 1 << 31)
    sk = ctx.gen_uniform_i32(ns, seed, 2, 0, int(1.33 * nr))
    sv = ctx.gen_uniform_i32(ns, seed, 3, 0, 1 << 31)
    a = ctx.join_partitioned_reduce(rk, rv, sk, sv)
    # the allocation-free entry point always builds the chained table
    b = ctx.join_chained_reduce_noalloc(rk, rv, sk, sv, 21)
    # build/probe swapped: COUNT and the checksum are symmetric
    c = ctx.join_partitioned_reduce(sk, sv, rk, rv)
    assert a == b
    assert c["count"] == a["count"] and c["sum_b"] == a["sum_p"] and c["sum_p"] == a["sum_b"]
    _free(rk, rv, sk, sv)


# ---- group-by ------------------------------------------------------------------------
def _check_groups(got, k, v, w):
    rc, want = O.groupby_chained(k, v, w, 16, max(len(np.unique(k)), 1))
    assert rc == O.OR_OK
    order = np.argsort(got["keys"], kind="stable")
    worder = np.argsort(want["keys"], kind="stable")
    assert np.array_equal(got["keys"][order], want["keys"][worder])
    for name in ("cnt", "sum", "min", "max"):
        assert np.array_equal(got[name][order], want[name][worder]), name
    np.testing.assert_allclose(got["fsum"][order], want["fsum"][worder], rtol=1e-9, atol=0)


@pytest.mark.parametrize("case", [
    dict(n=50000, g=1 << 6, zipf=0.0, maxg=1 << 6),
    dict(n=200000, g=1 << 10, zipf=0.8, maxg=1 << 10),
    dict(n=300000, g=1 << 16, zipf=0.0, maxg=1 << 16),
    dict(n=300000, g=1 << 16, zipf=0.8, maxg=1 << 17),
    dict(n=400000, g=1 << 19, zipf=0.0, maxg=1 << 19),
])
def test_groupby_parity(ctx, case):
    n, seed = case["n"], 31
    if case["zipf"]:
        k = O.gen_zipf_i32(n, seed, 1, case["g"], case["zipf"], 3)
    else:
        k = O.gen_uniform_i32(n, seed, 1, -case["g"] // 2, case["g"])
    v = O.gen_uniform_i64(n, seed, 2, -(1 << 40), 1 << 41)
    w = O.gen_unit_f32(n, seed, 3)
    d = _dev(ctx, k, v, w)
    got = ctx.groupby_i32(d[0], d[1], d[2], case["maxg"])
    _check_groups(got, k, v, w)
    _free(*d)


def _key_set(kind, n, seed):
    """Group keys whose min / max steer pj_groupby_i32 to its direct tables
    (dense key ranges) or to its hashed tables (sparse keys)."""
    rs = np.random.default_rng(seed)
    lo, hi = np.iinfo(np.int32).min, np.iinfo(np.int32).max
    if kind == "sparse_full_range":  # 2^16 keys spread over all of int32: hashed
        pool = rs.integers(lo, hi, 1 << 16, endpoint=True).astype(np.int32)
        pool[:2] = (lo, hi)
        return pool[rs.integers(0, len(pool), n)], 1 << 16
    if kind == "sparse_strided":  # multiples of 2^20: span beyond 2^14 ranges, hashed
        return (rs.integers(-1000, 1000, n) * (1 << 20)).astype(np.int32), 2048
    if kind == "near_max":  # [INT32_MAX - 3000, INT32_MAX]: direct, two ranges
        return (hi - rs.integers(0, 3001, n)).astype(np.int32), 4096
    if kind == "near_min":  # [INT32_MIN, INT32_MIN + 5000]: direct, three ranges
        return (lo + rs.integers(0, 5001, n)).astype(np.int32), 8192
    if kind == "span_2048":  # exactly one range: direct, no partitioning
        k = rs.integers(100, 100 + 2048, n).astype(np.int32)
        k[:2] = (100, 100 + 2047)
        return k, 2048
    if kind == "span_2049":  # one key past one range: direct, two ranges
        k = rs.integers(-7, -7 + 2049, n).astype(np.int32)
        k[:2] = (-7, -7 + 2048)
        return k, 4096
    if kind == "holes":  # multiples of 7 over 11 ranges: direct with empty slots
        return (rs.integers(0, 3000, n) * 7).astype(np.int32), 1 << 15
    if kind == "small_span_large_bound":  # maxInputSize 2^22, 1000 keys: direct, one range
        return rs.integers(0, 1000, n).astype(np.int32), 1 << 22
    if kind == "narrow_span_large_bound":  # 2^15 keys under a 2^22 bound: fewer digits, one pass
        return rs.integers(0, 1 << 15, n).astype(np.int32), 1 << 22
    if kind == "cross_aligned":  # [1000, 66536) crosses 2^16: direct, keys wrap in the window
        return rs.integers(1000, 66536, n).astype(np.int32), 1 << 16
    if kind == "sorted_spaced":  # sorted multiples of 10: each unit fits a table, the whole does not
        return np.sort(rs.integers(0, 1000, n) * 10).astype(np.int32), 1024
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["sparse_full_range", "sparse_strided", "near_max", "near_min",
                                  "span_2048", "span_2049", "holes", "small_span_large_bound",
                                  "narrow_span_large_bound", "cross_aligned", "sorted_spaced"])
def test_groupby_direct_and_hashed_tables(ctx, kind):
    """Every key layout against the chained-table oracle: dense key ranges
    take the direct tables (including ranges at both ends of int32 and a
    range that crosses an aligned boundary), sparse keys the hashed tables,
    and sorted sparse keys direct tables per unit merged through hashed
    ones."""
    n, seed = 150000, 41
    k, maxg = _key_set(kind, n, seed)
    v = O.gen_uniform_i64(n, seed, 2, -(1 << 40), 1 << 41)
    w = O.gen_unit_f32(n, seed, 3)
    d = _dev(ctx, k, v, w)
    got = ctx.groupby_i32(d[0], d[1], d[2], maxg)
    _check_groups(got, k, v, w)
    _free(*d)


def test_groupby_optional_columns_and_capacity(ctx):
    n = 20000
    k = O.gen_uniform_i32(n, 1, 1, 0, 500)
    d = ctx.to_device(k)
    got = ctx.groupby_i32(d, None, None, 1000)
    u, c = np.unique(k, return_counts=True)
    order = np.argsort(got["keys"])
    assert np.array_equal(got["keys"][order], u) and np.array_equal(got["cnt"][order], c)
    with pytest.raises(pa.PjError) as ex:
        ctx.groupby_i32(d, None, None, 100)
    assert ex.value.code == pa.PJ_E_CAPACITY
    e = ctx.to_device(np.zeros(0, np.int32))
    assert len(ctx.groupby_i32(e, None, None, 10)["keys"]) == 0
    _free(d, e)


def test_groupby_c3_exact_fsum(ctx):
    """The group-by workload at 1/16 scale: FP64 sums of u/2^24 floats are exact."""
    n, G, seed = 1 << 25, 1 << 16, 8
    k = ctx.gen_zipf_i32(n, seed, 1, G, 0.8, seed)
    v = ctx.gen_uniform_i64(n, seed, 2, 0, 1 << 20)
    w = ctx.gen_unit_f32(n, seed, 3)
    got = ctx.groupby_i32(k, v, w, G)
    rc, want = O.groupby_dense(k.numpy(), v.numpy(), w.numpy(), G)
    keys = got["keys"]
    assert len(keys) == int((want["cnt"] > 0).sum())
    for name in ("cnt", "sum", "min", "max", "fsum"):
        assert np.array_equal(got[name], want[name][keys]), name
    _free(k, v, w)


# ---- reduce / hash-rearrange ----------------------------------------------------------
def test_reduce_parity(ctx):
    for n in (0, 1, 1000, 1 << 20):
        v = O.gen_uniform_i64(n, 4, 1, -(1 << 62), 1 << 63) if n else np.zeros(0, np.int64)
        d = ctx.to_device(v)
        assert ctx.reduce_i64(d) == O.reduce_i64(v)
        d.free()


@pytest.mark.parametrize("buckets", [1, 2, 7, 31, 32, 200])
def test_hash_rearrange_parity(ctx, buckets):
    n = 100000
    k = O.gen_uniform_i32(n, 9, 1, -2**31, 1 << 32)
    v = np.arange(n, dtype=np.int32)
    d = _dev(ctx, k, v)
    ok, ov, offs = ctx.hash_rearrange_i32(d[0], d[1], buckets)
    okh, ovh = ok.numpy(), ov.numpy()
    want_b = np.array([O.rearrange_bucket(int(x), buckets) for x in k])
    assert offs[0] == 0 and offs[-1] == n
    for b in range(buckets):
        lo, hi = int(offs[b]), int(offs[b + 1])
        idx = np.sort(ovh[lo:hi])
        assert np.array_equal(idx, np.nonzero(want_b == b)[0])
        assert np.array_equal(np.sort(okh[lo:hi]), np.sort(k[want_b == b]))
    _free(*d, ok, ov)
