"""The allocation-free chained join (pj_join_chained_reduce_noalloc) against
pj_join_chained_reduce, through the C ABI: two calls in a row, the second on a
smaller input (the static scratch keeps the first call's table and must be
cleared), bit-exact, with no pool allocation during either call."""
import numpy as np
import pytest

import proteus_amd as pa
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _inputs(seed, nr, ns, key_range, dup):
    rs = np.random.default_rng(seed)
    if dup:
        rk = rs.integers(-key_range, key_range, nr).astype(np.int32)
    else:
        rk = (rs.permutation(nr).astype(np.int64) - nr // 2).astype(np.int32) * 3
    rv = rs.integers(-2**31, 2**31, nr).astype(np.int32)
    sk = rs.integers(-key_range, key_range, ns).astype(np.int32)
    sv = rs.integers(-2**31, 2**31, ns).astype(np.int32)
    return rk, rv, sk, sv


def _run_pair(ctx, arrs, hb):
    d = [ctx.to_device(a) for a in arrs]
    want = ctx.join_chained_reduce(d[0], d[1], d[2], d[3], hb)
    before = ctx.mem_stats()
    got = ctx.join_chained_reduce_noalloc(d[0], d[1], d[2], d[3], hb)
    after = ctx.mem_stats()
    for a in d:
        a.free()
    return want, got, before, after


def test_noalloc_matches_pool_entry_point_twice():
    with pa.Context(0) as ctx:
        # first call: 2^20 build rows at hash_bits 21 (2^21 slots, 16 MiB)
        big = _inputs(11, 1 << 20, 1 << 22, 1 << 21, False)
        want, got, before, after = _run_pair(ctx, big, 21)
        assert got == want
        assert after == before, (before, after)  # no pool allocation, no release
        rc, oracle = O.join_chained(*big, 21, threads=8)
        assert rc == O.OR_OK and got == oracle
        # second call, smaller, with duplicate build keys and a smaller table:
        # stale slots of the first call would add matches if left behind
        small = _inputs(12, 3000, 20000, 500, True)
        want2, got2, before2, after2 = _run_pair(ctx, small, 6)
        assert got2 == want2
        assert after2 == before2
        rc, oracle2 = O.join_chained(*small, 6, threads=8)
        assert rc == O.OR_OK and got2 == oracle2


def test_noalloc_limits_and_empty_sides():
    with pa.Context(0) as ctx:
        k = ctx.to_device(np.arange(100, dtype=np.int32))
        e = ctx.to_device(np.zeros(0, np.int32))
        assert ctx.join_chained_reduce_noalloc(e, None, k, None, 4)["count"] == 0
        assert ctx.join_chained_reduce_noalloc(k, None, e, None, 4)["count"] == 0
        assert ctx.join_chained_reduce_noalloc(k, None, k, None, 4)["count"] == 100
        with pytest.raises(pa.PjError) as ex:
            ctx.join_chained_reduce_noalloc(k, None, k, None, 23)  # 2^23 slots > 2^22
        assert ex.value.code == pa.PJ_E_CAPACITY
        with pytest.raises(pa.PjError) as ex:
            ctx.join_chained_reduce_noalloc(k, None, k, None, 4, max_build=99)
        assert ex.value.code == pa.PJ_E_CAPACITY
        k.free()
        e.free()
