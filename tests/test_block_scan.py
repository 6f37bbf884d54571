"""Prepared ("block") relations and host -> HBM staging.

The layout is the reference's tools/prepare_data.py:374-443 (one headerless
file per column, string columns as int32 codes with a sorted "value:code"
dictionary).  The CPU tests pin the test-side plan evaluator (tests/
plan_eval.py) with hand-written SSB formulas; the GPU tests check the staging
ring byte for byte against the files and host arrays it copies, and run the
reference's thirteen SSB bare plans (tests/inputs/plans/proteus_bare_plans)
over synthetic SSB column files against the evaluator."""
import glob
import json
import os

import numpy as np
import pytest

import proteus_amd as pa
from tests import helpers as H
from tests import plan_eval, ssb

PLANS = sorted(glob.glob(os.path.join(H.GOLDEN, "inputs", "plans", "ssb", "*.json")))
CHUNK = 8 << 20


@pytest.fixture(scope="module")
def ssb_dir(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("ssb"))
    tabs = ssb.tables(300_000, seed=3)
    ssb.write(root, "ssbm100", tabs)
    return root, tabs


def _by_key(keys, values):
    return dict(zip(keys.tolist(), values.tolist()))


# ---- CPU: the evaluator and the layout ---------------------------------------
def test_evaluator_matches_hand_written_ssb_queries(ssb_dir):
    """SSB Q1.1 (sum of extendedprice * discount for 1993, discount 1..3,
    quantity < 25) and Q2.1's bare form (sums of revenue, year and size for
    category MFGR#12 and suppliers in AMERICA), written directly in numpy."""
    root, t = ssb_dir
    ev = plan_eval.Evaluator(root)
    lo, d, p, s = t["lineorder"], t["date"], t["part"], t["supplier"]
    year_of = _by_key(d["d_datekey"], d["d_year"])
    year = np.array([year_of[x] for x in lo["lo_orderdate"].tolist()])
    m = (year == 1993) & (lo["lo_discount"] >= 1) & (lo["lo_discount"] <= 3) & (lo["lo_quantity"] < 25)
    q11 = int((lo["lo_extendedprice"][m].astype(np.int64) * lo["lo_discount"][m]).sum())
    assert ev.run(json.load(open(PLANS[0]))) == [(q11,)]
    cat = p["p_category"][lo["lo_partkey"] - 1]
    reg = s["s_region"][lo["lo_suppkey"] - 1]
    m = (cat == "MFGR#12") & (reg == "AMERICA")
    size = p["p_size"][lo["lo_partkey"] - 1]
    q21 = (int(lo["lo_revenue"][m].astype(np.int64).sum()), int(year[m].sum()), int(size[m].sum()))
    assert os.path.basename(PLANS[3]) == "q2_1.json"
    assert ev.run(json.load(open(PLANS[3]))) == [q21]


def test_layout_of_written_relations(ssb_dir):
    root, t = ssb_dir
    base = os.path.join(root, "inputs", "ssbm100", "part.csv.p_brand1")
    lines = open(base + ".dict").read().splitlines()
    values = [ln.rpartition(":")[0] for ln in lines]
    assert values == sorted(values) and [int(ln.rpartition(":")[2]) for ln in lines] == list(range(len(lines)))
    codes = np.fromfile(base, np.int32)
    assert [values[c] for c in codes[:50]] == t["part"]["p_brand1"][:50].tolist()
    assert os.path.getsize(os.path.join(root, "inputs", "ssbm100", "lineorder.csv.lo_quantity")) == 4 * 300_000


# ---- GPU: the staging ring -----------------------------------------------------
@pytest.mark.gpu
def test_stage_file_bytes(ctx, tmp_path):
    rs = np.random.default_rng(11)
    data = rs.integers(0, 256, 3 * CHUNK + 4096 + 5, dtype=np.uint8)
    f = str(tmp_path / "bytes.bin")
    data.tofile(f)
    cases = [(0, 0), (0, 1), (3, 4093), (0, CHUNK - 4), (0, CHUNK), (7, CHUNK + 12),
             (4096, 3 * CHUNK + 5), (0, len(data))]
    for off, n in cases:
        d = ctx.empty(max(n, 1), np.uint8)
        try:
            ctx.stage_file(f, d, n, offset=off)
            if n:
                assert np.array_equal(d.numpy()[:n], data[off:off + n]), (off, n)
        finally:
            d.free()


@pytest.mark.gpu
def test_stage_host_arrays_many_in_flight(ctx):
    """Six copies queued at once (host arrays and a file), waited for in
    reverse order; each lands intact."""
    rs = np.random.default_rng(12)
    srcs = [rs.integers(0, 256, n, dtype=np.uint8) for n in (1, 5000, CHUNK, 2 * CHUNK + 3, 123457, 0)]
    devs = [ctx.empty(max(len(a), 1), np.uint8) for a in srcs]
    try:
        tickets = [ctx.stage_begin(d, len(a), host=a) for d, a in zip(devs, srcs)]
        for t in reversed(tickets):
            ctx.stage_wait(t)
        for d, a in zip(devs, srcs):
            if len(a):
                assert np.array_equal(d.numpy()[:len(a)], a)
    finally:
        for d in devs:
            d.free()


@pytest.mark.gpu
def test_stage_errors(ctx, tmp_path):
    f = str(tmp_path / "short.bin")
    np.zeros(100, np.uint8).tofile(f)
    d = ctx.empty(256, np.uint8)
    try:
        with pytest.raises(pa.PjError) as e:
            ctx.stage_file(f, d, 101)
        assert e.value.code == pa.PJ_E_ARG and "holds 100 bytes" in str(e.value)
        with pytest.raises(pa.PjError) as e:
            ctx.stage_file(f, d, 10, offset=95)
        assert e.value.code == pa.PJ_E_ARG
        with pytest.raises(pa.PjError) as e:
            ctx.stage_file(str(tmp_path / "missing.bin"), d, 4)
        assert e.value.code == pa.PJ_E_ARG
        t = ctx.stage_begin(d, 100, path=f)
        ctx.stage_wait(t)
        with pytest.raises(pa.PjError):
            ctx.stage_wait(t)  # a ticket is waited for once
        ctx.stage_file(f, d, 100)
        assert not d.numpy()[:100].any()
    finally:
        d.free()


# ---- GPU: SSB bare plans over column files --------------------------------------
def _rows(text):
    return [tuple(int(x) for x in ln.split(",")) for ln in text.strip().splitlines()]


@pytest.mark.gpu
@pytest.mark.parametrize("plan", PLANS, ids=[os.path.basename(p)[:-5] for p in PLANS])
def test_ssb_bare_plans(ctx, ssb_dir, plan):
    root, _ = ssb_dir
    print("ssb plan", os.path.basename(plan), flush=True)
    doc = json.load(open(plan))
    want = plan_eval.Evaluator(root).run(doc)
    assert _rows(ctx.execute_plan(doc, base_dir=root)) == want


@pytest.mark.gpu
def test_block_columns_stay_resident_until_the_file_changes(ctx, tmp_path):
    """A second run reuses the staged columns; rewriting a column file
    (new mtime) restages it; clear_catalog drops them."""
    root = str(tmp_path)
    tabs = ssb.tables(100_000, seed=5)
    ssb.write(root, "ssbm100", tabs)
    doc = json.load(open(PLANS[0]))
    ev = plan_eval.Evaluator(root)
    ctx.clear_catalog()
    base = ctx.mem_in_use()
    first = _rows(ctx.execute_plan(doc, base_dir=root))
    resident = ctx.mem_in_use()
    assert resident > base
    assert _rows(ctx.execute_plan(doc, base_dir=root)) == first == ev.run(doc)
    assert ctx.mem_in_use() == resident
    f = os.path.join(root, "inputs", "ssbm100", "lineorder.csv.lo_discount")
    disc = np.fromfile(f, np.int32)
    np.where(disc == 2, 3, disc).astype(np.int32).tofile(f)
    st = os.stat(f)
    os.utime(f, ns=(st.st_atime_ns, st.st_mtime_ns + 1_000_000_000))
    second = _rows(ctx.execute_plan(doc, base_dir=root))
    assert second == ev.run(doc) and second != first
    ctx.clear_catalog()
    assert ctx.mem_in_use() == base


@pytest.mark.gpu
def test_block_scan_errors(ctx, tmp_path):
    root = str(tmp_path)
    ssb.write(root, "ssbm100", ssb.tables(1000, seed=2))
    doc = json.load(open(PLANS[0]))
    f = os.path.join(root, "inputs", "ssbm100", "lineorder.csv.lo_quantity")
    with open(f, "ab") as fh:
        fh.write(b"\x01\x02")  # not a whole number of int32 values
    with pytest.raises(pa.PjError) as e:
        ctx.execute_plan(doc, base_dir=root)
    assert "whole number" in str(e.value)
    os.remove(f)
    with pytest.raises(pa.PjError) as e:
        ctx.execute_plan(doc, base_dir=root)
    assert "no column file" in str(e.value)
    ctx.clear_catalog()
