"""The batched / served query mix: a seeded stream of join -> group-by
queries over resident dimensions.

Query i: dimension size 2^U with U uniform in [u_lo, u_hi] (log-uniform
build sizes), probe rows m * 2^U with m uniform in [8, 64], group
cardinality 2^V with V uniform in [4, 16], fk uniform over [0, 1.33 * 2^U)
(~25% misses), arrivals a Poisson process at `qps`.  Dimension U holds
keys = seeded permutation of 0..2^U-1 and values uniform [0, 2^31).
Everything is a function of the master seed.
"""
import numpy as np

from . import Query

M_LO, M_HI = 8, 64
V_LO, V_HI = 4, 16


def _rng(seed, tag):
    return np.random.default_rng([int(seed), int(tag)])


class Mix:
    def __init__(self, ctx, nq, seed, u_lo=10, u_hi=24, qps=None):
        self.ctx, self.nq, self.seed = ctx, nq, seed
        r = _rng(seed, 1)
        self.u = r.integers(u_lo, u_hi + 1, nq)
        self.m = r.integers(M_LO, M_HI + 1, nq)
        self.v = r.integers(V_LO, V_HI + 1, nq)
        self.n = self.m.astype(np.int64) << self.u
        self.groups = (1 << self.v).astype(np.int64)
        self.gen_seed = (np.arange(nq, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
                         + np.uint64(seed) * np.uint64(1000003) + np.uint64(1))
        self.fk_range = (1.33 * (1 << self.u.astype(np.int64))).astype(np.int64)
        if qps:
            gaps = _rng(seed, 2).exponential(1.0 / qps, nq)
            self.arrivals = np.cumsum(gaps) - gaps[0]
        else:
            self.arrivals = np.zeros(nq)
        self.dims = {}
        self.dim_cols = {}
        for u in sorted(set(self.u.tolist())):
            nb = 1 << u
            k = ctx.gen_perm_i32(nb, seed * 131 + u)
            v = ctx.gen_uniform_i32(nb, seed * 131 + u, 7, 0, 1 << 31)
            self.dims[u] = ctx.dim_create(k, v)
            self.dim_cols[u] = (k, v)
        self.out_cnt = [ctx.empty(int(g), np.int64) for g in self.groups]
        self.out_sum = [ctx.empty(int(g), np.int64) for g in self.groups]

    def queries(self):
        qs = []
        for i in range(self.nq):
            qs.append(Query(self.dims[int(self.u[i])], None, None, None, int(self.n[i]),
                            int(self.groups[i]), 0, int(self.gen_seed[i]), int(self.fk_range[i]),
                            self.out_cnt[i].ptr, self.out_sum[i].ptr))
        return qs

    def autotune(self, k=64, max_rows=1 << 26):
        """Tunes the context's batch defaults (tile rows, workgroups per CU,
        LDS table threshold) on a seeded sample of k queries of at most
        max_rows probe rows, with their probe columns allocated for the run.
        Returns the tuned configuration."""
        idx = np.flatnonzero(self.n <= max_rows)
        pick = _rng(self.seed, 3).choice(idx, size=min(k, len(idx)), replace=False)
        qs, bufs = [], []
        for i in sorted(pick.tolist()):
            n = int(self.n[i])
            cols = [self.ctx.empty(n, np.int32) for _ in range(3)]
            bufs += cols
            qs.append(Query(self.dims[int(self.u[i])], cols[0].ptr, cols[1].ptr, cols[2].ptr, n,
                            int(self.groups[i]), 0, int(self.gen_seed[i]),
                            int(self.fk_range[i]), self.out_cnt[i].ptr, self.out_sum[i].ptr))
        try:
            return self.ctx.batch_autotune(qs)
        finally:
            for b in bufs:
                b.free()

    def result(self, i):
        return self.out_cnt[i].numpy(), self.out_sum[i].numpy()

    def results(self):
        return [self.result(i) for i in range(self.nq)]

    def probe_rows(self):
        return int(self.n.sum())

    def close(self):
        for d in self.dims.values():
            self.ctx.dim_destroy(d)
        for k, v in self.dim_cols.values():
            k.free()
            v.free()
        for a in self.out_cnt + self.out_sum:
            a.free()
        self.dims, self.dim_cols, self.out_cnt, self.out_sum = {}, {}, [], []
