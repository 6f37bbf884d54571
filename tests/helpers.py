"""Shared test data: seeded inputs (via the oracle's generators) and the
reference's sailors fixtures."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_csv(name, cols, dtype=np.int64):
    rows = []
    with open(os.path.join(GOLDEN, "inputs", name)) as f:
        for line in f:
            parts = line.strip().split(";")
            if parts and parts[0]:
                rows.append([parts[c] for c in cols])
    return [np.array([dtype(float(r[i])) if dtype is not str else r[i] for r in rows])
            for i in range(len(cols))]


def sailors():
    sid, rating, age = load_csv("sailors.csv", [0, 2, 3], float)
    return {"sid": sid.astype(np.int32), "rating": rating.astype(np.int32), "age": age}


def reserves():
    sid, bid = load_csv("reserves.csv", [0, 1], float)
    return {"sid": sid.astype(np.int32), "bid": bid.astype(np.int32)}


def boats():
    (bid,) = load_csv("boats.csv", [0], float)
    return {"bid": bid.astype(np.int32)}


def brute_join(rk, rv, sk, sv):
    """Nested-loop equi-join -> (count, sum_b, sum_p, min_b, max_b, min_p, max_p)."""
    rk = np.asarray(rk, np.int64)
    sk = np.asarray(sk, np.int64)
    rv = np.zeros(len(rk), np.int64) if rv is None else np.asarray(rv, np.int64)
    sv = np.zeros(len(sk), np.int64) if sv is None else np.asarray(sv, np.int64)
    eq = rk[:, None] == sk[None, :]
    bi, pi = np.nonzero(eq)
    if len(bi) == 0:
        return dict(count=0, sum_b=0, sum_p=0, min_b=2**63 - 1, max_b=-2**63,
                    min_p=2**63 - 1, max_p=-2**63)
    b, p = rv[bi], sv[pi]
    return dict(count=len(bi), sum_b=int(b.sum()), sum_p=int(p.sum()), min_b=int(b.min()),
                max_b=int(b.max()), min_p=int(p.min()), max_p=int(p.max()))


def brute_groupby(k, v, w):
    out = {}
    for i in range(len(k)):
        key = int(k[i])
        vi = int(v[i]) if v is not None else 0
        wi = float(w[i]) if w is not None else 0.0
        if key not in out:
            out[key] = [1, vi, vi, vi, wi]
        else:
            e = out[key]
            e[0] += 1
            e[1] += vi
            e[2] = min(e[2], vi)
            e[3] = max(e[3], vi)
            e[4] += wi
    return out


def star_expected(star_obj, slice_rows=1 << 26):
    """The star-schema query's result (rows [s_nation, c_city, revenue]) from
    the fact columns of a proteus_amd.star.Star, computed with numpy in
    slices of slice_rows rows copied from the device, so 6e8-row facts fit
    in host memory.  Dimension attributes are the closed forms of their row
    ids (star.dims())."""
    import proteus_amd as pa
    from proteus_amd import star as st
    tot = np.zeros(25 * 160, np.int64)
    cnt = np.zeros(25 * 160, np.int64)
    n = star_obj.n
    for lo in range(0, n, slice_rows):
        m = min(slice_rows, n - lo)
        f = {name: pa.DeviceArray(star_obj.ctx, d.ptr + lo * 4, m, np.int32).numpy()
             for name, d in star_obj.fact.items()}
        keep = ((1992 + f["fk_date"] // 365) == st.YEAR) & (f["fk_supp"] % 5 == st.REGION) & \
               (f["fk_part"] % 25 == st.CATEGORY)
        g = {name: c[keep].astype(np.int64) for name, c in f.items()}
        key = (g["fk_supp"] % 25) * 160 + g["fk_cust"] % 160
        rev = g["price"] * (100 - g["discount"])
        np.add.at(tot, key, rev)
        np.add.at(cnt, key, 1)
    return [[int(k // 160), int(k % 160), int(tot[k])] for k in np.nonzero(cnt)[0]]
