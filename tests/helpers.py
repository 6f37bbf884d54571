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
