"""Pins the CPU oracle (oracle/oracle.c) before results are checked against it:
the reference's own golden outputs, closed forms, brute force on tiny inputs
and a library routine.  CPU only."""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import helpers as H

GOLD = json.load(open(os.path.join(H.GOLDEN, "reference_outputs.json")))


def _triple(name):
    return [float(x) for x in GOLD[name]["expected"].replace("[", "").replace("]", "").split(",")]


# ---- reference goldens (RelBuilder operator tests, plan-parsing tests) ------
def test_sailors_join_left3_golden():
    """tests-join-sailors.cpp:100-145: sailors.join(reserves on sid)
    .join(boats on bid).reduce(MAX sid, SUM sid, SUM 1) -> 74,383,10."""
    s, r, b = H.sailors(), H.reserves(), H.boats()
    rc, ob, op = O.join_pairs(r["sid"], s["sid"], 16, 128)
    assert rc == O.OR_OK
    bid = r["bid"][ob]
    sid = s["sid"][op]
    rc, ob2, op2 = O.join_pairs(b["bid"], bid, 16, 128)
    assert rc == O.OR_OK
    sid2 = sid[op2].astype(np.int64)
    got = [sid2.max(), sid2.sum(), len(sid2)]
    assert got == _triple("sailorsJoinLeft3")
    j = json.loads(GOLD["3wayJoin"]["expected"])[0]
    assert got == [j["max_sid"], j["sum_sid"], j["cnt"]]


def test_plan_join_count_golden():
    """PlanTest.Join (tests-plan-parsing.cpp:96-101): COUNT(*) of sailors
    JOIN reserves on sid = 10."""
    s, r = H.sailors(), H.reserves()
    rc, out = O.join_chained(r["sid"], None, s["sid"], None, 16)
    assert rc == O.OR_OK
    assert out["count"] == int(GOLD["reduce-join"]["expected"])


def test_sailors_scan_select_goldens():
    s, b = H.sailors(), H.boats()
    sm, mn, mx = O.reduce_i64(s["sid"].astype(np.int64))
    assert [mx, sm, len(s["sid"])] == _triple("sailorsScan")
    sel = s["sid"][s["age"] > 40.0].astype(np.int64)
    sm, mn, mx = O.reduce_i64(sel)
    assert [mx, sm, len(sel)] == _triple("sailorsSel")
    sm, mn, mx = O.reduce_i64(b["bid"].astype(np.int64))
    assert [mx, sm, len(b["bid"])] == _triple("sailorsScanBoats")


# ---- join restatement vs brute force / dense / closed form -------------------
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_join_chained_matches_brute_force(seed):
    rs = np.random.default_rng(seed)
    nr, ns = 300, 900
    rk = rs.integers(-40, 40, nr).astype(np.int32)  # duplicates and negatives
    rk[:3] = [np.iinfo(np.int32).min, np.iinfo(np.int32).max, -1]
    rv = rs.integers(-2**31, 2**31, nr).astype(np.int32)
    sk = rs.integers(-50, 50, ns).astype(np.int32)
    sk[:2] = [np.iinfo(np.int32).min, -1]
    sv = rs.integers(-2**31, 2**31, ns).astype(np.int32)
    for hb in (0, 3, 10):
        rc, out = O.join_chained(rk, rv, sk, sv, hb)
        assert rc == O.OR_OK
        assert out == H.brute_join(rk, rv, sk, sv)
    rc, out = O.join_chained(rk, rv, sk, sv, 8, threads=4)
    assert out == H.brute_join(rk, rv, sk, sv)


def test_join_pairs_match_brute_force():
    rs = np.random.default_rng(7)
    rk = rs.integers(0, 20, 50).astype(np.int32)
    sk = rs.integers(0, 25, 80).astype(np.int32)
    rc, ob, op = O.join_pairs(rk, sk, 5)
    got = sorted(zip(ob.tolist(), op.tolist()))
    want = sorted((i, j) for i in range(len(rk)) for j in range(len(sk)) if rk[i] == sk[j])
    assert rc == O.OR_OK and got == want


def test_join_c1_shape_closed_form():
    """R = permutation of 0..n-1, S uniform over R's keys: COUNT = |S|."""
    nr, ns, seed = 1 << 16, 1 << 20, 11
    rk = O.gen_perm_i32(nr, seed)
    assert np.array_equal(np.sort(rk), np.arange(nr, dtype=np.int32))
    rv = O.gen_uniform_i32(nr, seed, 1, 0, 1 << 31)
    sk = O.gen_uniform_i32(ns, seed, 2, 0, nr)
    sv = O.gen_uniform_i32(ns, seed, 3, 0, 1 << 31)
    rc, chained = O.join_chained(rk, rv, sk, sv, 17)
    rc2, dense = O.join_dense(rk, rv, sk, sv)
    rc3, mt = O.join_chained(rk, rv, sk, sv, 17, threads=8)
    assert rc == rc2 == rc3 == O.OR_OK
    assert chained == dense == mt
    assert chained["count"] == ns
    # closed form of the checksum: sum over S of payR[fk] + payS
    pay = np.zeros(nr, np.int64)
    pay[rk] = rv
    assert chained["sum_b"] == int(pay[sk].sum())
    assert chained["sum_p"] == int(sv.astype(np.int64).sum())


def test_join_capacity_error():
    rk = np.arange(10, dtype=np.int32)
    rc, _ = O.join_chained(rk, None, rk, None, 4, max_build=9)
    assert rc == O.OR_E_CAPACITY


# ---- group-by restatement ------------------------------------------------------
@pytest.mark.parametrize("seed", [1, 2])
def test_groupby_chained_matches_brute_force(seed):
    rs = np.random.default_rng(seed)
    n = 2000
    k = rs.integers(-30, 30, n).astype(np.int32)
    v = rs.integers(-10**12, 10**12, n).astype(np.int64)
    w = (rs.integers(0, 1 << 24, n) / float(1 << 24)).astype(np.float32)
    rc, g = O.groupby_chained(k, v, w, 4, 64)
    assert rc == O.OR_OK
    want = H.brute_groupby(k, v, w)
    assert len(g["keys"]) == len(want)
    # first-appearance order, as the reference's sequential scan emits it
    first = list(dict.fromkeys(k.tolist()))
    assert g["keys"].tolist() == first
    for i, key in enumerate(g["keys"].tolist()):
        e = want[key]
        assert [g["cnt"][i], g["sum"][i], g["min"][i], g["max"][i]] == e[:4]
        assert g["fsum"][i] == pytest.approx(e[4], rel=1e-12)


def test_groupby_dense_exact_fsum_and_capacity():
    n, G = 1 << 16, 1 << 8
    k = O.gen_zipf_i32(n, 5, 1, G, 0.8, 9)
    v = O.gen_uniform_i64(n, 5, 2, 0, 1 << 20)
    w = O.gen_unit_f32(n, 5, 3)
    rc, d = O.groupby_dense(k, v, w, G)
    rc2, c = O.groupby_chained(k, v, w, 10, G)
    assert rc == rc2 == O.OR_OK
    order = np.argsort(c["keys"])
    for name in ("cnt", "sum", "min", "max"):
        live = d["cnt"] > 0
        assert np.array_equal(d[name][live], c[name][order])
    # exact: every w is u / 2^24, so the FP64 sum is an int64 count of units
    units = np.zeros(G, np.int64)
    np.add.at(units, k, (w.astype(np.float64) * (1 << 24)).astype(np.int64))
    assert np.array_equal(d["fsum"], units / float(1 << 24))
    assert np.allclose(c["fsum"], d["fsum"][c["keys"]], rtol=1e-12)
    rc3, _ = O.groupby_chained(k, v, w, 10, G // 2)
    assert rc3 == O.OR_E_CAPACITY


# ---- hashes -------------------------------------------------------------------------
def _murmur3_body(key, seed=0):
    """MurmurHash3_x86_32 (Appleby) of one 4-byte block, up to (not
    including) the finaliser."""
    c1, c2, M = 0xcc9e2d51, 0x1b873593, 0xffffffff
    k1 = (key * c1) & M
    k1 = ((k1 << 15) | (k1 >> 17)) & M
    k1 = (k1 * c2) & M
    h = seed ^ k1
    h = ((h << 13) | (h >> 19)) & M
    h = (h * 5 + 0xe6546b64) & M
    return h ^ 4


def test_fmix32_matches_library_murmur3():
    """The partition hash (fmix32 of gpu-partitioned-hash-join-chained.cpp:
    595-610, = or_phjc_bucket with shift 0 and 32 bits) is MurmurHash3's
    finaliser: pinned against scikit-learn's murmurhash3_32."""
    from sklearn.utils import murmurhash3_32
    rs = np.random.default_rng(3)
    for key in [0, 1, 2, 0x7fffffff, 0x80000000, 0xffffffff] + rs.integers(0, 2**32, 200).tolist():
        pre = _murmur3_body(int(key))
        fm = O.phjc_bucket(np.uint32(pre).view(np.int32), 0, 32)
        assert fm == murmurhash3_32(int(np.uint32(key).view(np.int32)), seed=0, positive=True)


def _unxorshift64(y, s):
    """Inverse of x -> x ^ (x >> s) on 64 bits (fixed-point iteration: each
    round fixes s more high bits)."""
    x = y
    for _ in range(64 // s + 1):
        x = y ^ (x >> s)
    return x


def test_hash_i32_inverts_to_sign_extended_key():
    """hashInt32 (or_hash_i32, expressions-hasher.cpp:39-58) is MurmurHash3's
    fmix32 step sequence (shifts 16, 13, 16; multipliers 0x85ebca6b,
    0xc2b2ae35, as published with MurmurHash3) evaluated in 64-bit arithmetic
    on the sign-extended key.  Each step is a bijection of 64-bit words, so
    undoing them (xorshift inverses, multiplicative inverses mod 2^64 computed
    here) must give back the sign-extended key exactly.  A wrong shift, a
    wrong constant, a 32-bit evaluation or a zero extension all fail this."""
    M = (1 << 64) - 1
    inv1 = pow(0x85ebca6b, -1, 1 << 64)
    inv2 = pow(0xc2b2ae35, -1, 1 << 64)
    rs = np.random.default_rng(11)
    keys = [0, 1, -1, 2, -2, 0x7fffffff, -0x80000000] + rs.integers(-2**31, 2**31, 300).tolist()
    for k in keys:
        h = int(O.hash_i32(int(k))) & M
        x = _unxorshift64(h, 16)
        x = (x * inv2) & M
        x = _unxorshift64(x, 13)
        x = (x * inv1) & M
        x = _unxorshift64(x, 16)
        assert x == (int(k) & M), k


def test_hash_special_cases():
    """Special cases of hashInt32 (or_hash_i32) and the rearrange bucket."""
    assert O.hash_i32(0) == 0
    # bucket of a 1-bucket rearrange is 0; buckets cover [0, b)
    assert all(O.rearrange_bucket(k, 1) == 0 for k in range(-5, 5))
    assert all(0 <= O.rearrange_bucket(k, 7) < 7 for k in range(-50, 50))
    # hashInt32 sign-extends: a negative key does not hash like its low 31 bits
    for k in (-1, -7, np.iinfo(np.int32).min):
        assert O.hash_i32(k) != O.hash_i32(k & 0x7fffffff)


# ---- generators -------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 1000, 2556, 1 << 12])
def test_perm_is_bijection(n):
    p = np.array([O.perm(42, i, n) for i in range(n)])
    assert np.array_equal(np.sort(p), np.arange(n))


def test_uniform_and_unit_ranges():
    u = O.gen_uniform_i32(1 << 16, 3, 1, 90000, 20001)
    assert u.min() >= 90000 and u.max() <= 110000
    assert abs(u.mean() - 100000) < 100
    w = O.gen_unit_f32(1 << 16, 3, 2)
    assert (w >= 0).all() and (w < 1).all()
    assert np.all((w.astype(np.float64) * (1 << 24)) % 1 == 0)


@pytest.mark.parametrize("theta", [0.8, 1.0])
def test_zipf_matches_closed_form(theta):
    """P(rank 0) = 1 / H(n, theta) for Zipf(theta) over n ranks."""
    n, m = 1 << 10, 1 << 18
    z = O.gen_zipf_i32(m, 7, 1, n, theta, 0)
    # perm_seed 0 still permutes: compare rank frequencies via the inverse
    inv = {O.perm(0, r, n): r for r in range(n)}
    ranks = np.array([inv[int(x)] for x in z[:m]])
    hn = np.sum(1.0 / np.arange(1, n + 1) ** theta)
    for r in (0, 1, 9):
        p = (1.0 / (r + 1) ** theta) / hn
        f = np.mean(ranks == r)
        assert abs(f - p) < 5 * np.sqrt(p * (1 - p) / m) + 1e-4


def test_zipf_fk_miss_rate():
    nb, rng = 1 << 14, int(1.33 * (1 << 14))
    fk = O.gen_zipf_fk_i32(1 << 18, 1, 2, nb, rng, 1.0, 5)
    miss = np.mean(fk >= nb)
    assert abs(miss - (rng - nb) / rng) < 0.01
    assert fk.min() >= 0 and fk.max() < rng


def test_histogram_conservation():
    k = O.gen_uniform_i32(1 << 15, 1, 1, -2**31, 1 << 32)
    h = O.histogram(k, 32 - 7, 7)
    assert h.sum() == len(k)


def test_histogram_matches_bucket_by_bucket_count():
    """or_histogram's bins equal a count of each key's partition digit, the
    digit taken from or_phjc_bucket (fmix32, pinned against scikit-learn's
    murmur3 above), for the top bits and for a middle bit field."""
    k = O.gen_uniform_i32(1 << 12, 2, 1, -2**31, 1 << 32)
    for shift, bits in ((32 - 7, 7), (9, 5)):
        want = np.zeros(1 << bits, np.uint64)
        for x in k:
            want[O.phjc_bucket(int(x), shift, bits)] += 1
        assert np.array_equal(O.histogram(k, shift, bits), want), (shift, bits)


def test_groupby_chained_mt_equals_single_thread():
    """The per-worker chained group-by with its merge equals the
    single-threaded one (as a set of groups) and the dense oracle."""
    n, g = 200003, 1 << 12
    k = O.gen_zipf_i32(n, 5, 1, g, 0.8, 2)
    v = O.gen_uniform_i64(n, 5, 2, -(1 << 40), 1 << 41)
    w = O.gen_unit_f32(n, 5, 3)
    rc1, one = O.groupby_chained(k, v, w, 12, g)
    rc8, mt = O.groupby_chained(k, v, w, 12, g, threads=8)
    rcd, dense = O.groupby_dense(k, v, w, g)
    assert rc1 == rc8 == rcd == O.OR_OK
    o1, o8 = np.argsort(one["keys"]), np.argsort(mt["keys"])
    for name in ("keys", "cnt", "sum", "min", "max"):
        assert np.array_equal(one[name][o1], mt[name][o8]), name
    keys = mt["keys"][o8]
    for name in ("cnt", "sum", "min", "max", "fsum"):
        assert np.array_equal(mt[name][o8], dense[name][keys]), name
    rc, _ = O.groupby_chained(k, v, w, 12, 100, threads=4)
    assert rc == O.OR_E_CAPACITY
