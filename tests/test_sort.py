"""The device radix sort (pj_sort_records, and the plan operator sort built on
the same passes) against numpy's stable sort of the same records: int32 and
int64 fields in C-struct layout, duplicates, negatives and the extreme
values, narrow key ranges (digit positions skipped), partial tiles and a
2^25-record case."""
import numpy as np
import pytest

import proteus_amd as pa

I32 = np.iinfo(np.int32)
I64 = np.iinfo(np.int64)


def _dtype(fields):
    return np.dtype([("f%d" % i, "<i4" if f == "i" else "<i8") for i, f in enumerate(fields)],
                    align=True)


def _records(fields, n, rs, narrow=False):
    r = np.zeros(n, _dtype(fields))
    for i, f in enumerate(fields):
        info = I32 if f == "i" else I64
        if narrow:
            v = rs.integers(-3, 1 << 12, n)
        else:
            v = rs.integers(info.min, info.max, n, endpoint=True)
        if n >= 4:
            v[:4] = [info.min, info.max, -1, 0]
        r["f%d" % i] = v
    return r


def _sorted(r, fields):
    keys = tuple(r["f%d" % i] for i in reversed(range(len(fields))))
    return r[np.lexsort(keys)]


def test_record_layout_is_a_c_struct():
    assert _dtype("il").itemsize == 16 and _dtype("il").fields["f1"][1] == 8
    assert _dtype("li").itemsize == 16 and _dtype("iil").itemsize == 16
    assert _dtype("iii").itemsize == 12


@pytest.mark.gpu
@pytest.mark.parametrize("fields", ["i", "l", "il", "li", "iil", "lil", "iiii"])
@pytest.mark.parametrize("n", [0, 1, 2, 1000, 2048, 2049, 4095, 4096, 4097, 100_003])
@pytest.mark.parametrize("narrow", [False, True])
def test_sort_records_matches_numpy(ctx, fields, n, narrow):
    rs = np.random.default_rng(n * 31 + len(fields) + narrow)
    r = _records(fields, n, rs, narrow)
    d = ctx.to_device(r.view(np.uint8)) if n else ctx.empty(1, np.uint8)
    try:
        ctx.sort_records(d, n, fields)
        if n:
            got = d.numpy().view(r.dtype)
            assert np.array_equal(got, _sorted(r, fields)), (fields, n, narrow)
    finally:
        d.free()


@pytest.mark.gpu
def test_sort_records_large_and_skewed(ctx):
    """2^25 int32 keys over the full range with an int64 tie-breaker, and 2^22
    records whose first field is one value for 90% of the rows (one digit per
    pass holds most keys)."""
    rs = np.random.default_rng(5)
    n = 1 << 25
    r = _records("il", n, rs)
    r["f1"] = rs.integers(0, 1 << 20, n)
    d = ctx.to_device(r.view(np.uint8))
    try:
        print("sort 2^25 il", flush=True)
        ctx.sort_records(d, n, "il")
        assert np.array_equal(d.numpy().view(r.dtype), _sorted(r, "il"))
    finally:
        d.free()
    n = 1 << 22
    r = _records("li", n, rs)
    r["f0"][rs.random(n) < 0.9] = 12345
    d = ctx.to_device(r.view(np.uint8))
    try:
        ctx.sort_records(d, n, "li")
        assert np.array_equal(d.numpy().view(r.dtype), _sorted(r, "li"))
    finally:
        d.free()


@pytest.mark.gpu
def test_sort_records_errors(ctx):
    d = ctx.empty(64, np.uint8)
    try:
        for bad in ("", "x", "iiiiiiiii"):
            with pytest.raises(pa.PjError) as e:
                ctx.sort_records(d, 2, bad)
            assert e.value.code == pa.PJ_E_ARG
    finally:
        d.free()
