"""Synthetic star schema (Star Schema Benchmark Q4.x shape) for the plan
executor: a fact table joined to four dimensions, three of them filtered,
grouped by two dimension attributes.

fact (n rows, int32): fk_date in [0, 2556), fk_supp in [0, 2e5),
fk_part in [0, 3e6), fk_cust in [0, 8e5) uniform; quantity [1, 50],
price [90000, 110000], discount [0, 10].
date (2556): d_datekey = row, d_year = 1992 + row // 365  (filter = 1997: 1/7)
supplier (2e5): s_suppkey = row, s_region = row % 5 (filter = 1: 1/5),
                s_nation = row % 25
part (3e6): p_partkey = row, p_category = row % 25 (filter = 3: 1/25)
customer (8e5): c_custkey = row, c_city = row % 160
Query: SELECT s_nation, c_city, SUM(price * (100 - discount)) (int64)
       GROUP BY s_nation, c_city -> 5 x 160 = 800 groups.
Fact columns are generated on the device from (seed, column); the
dimensions are closed forms of the row id.

With bloom_size > 0 the plan is the shape of the reference's ssb100_bloom
queries (apps/benchmarks/olap/prepared-queries/ssb100_bloom/q4_1.sql.json.cpp):
the filtered supplier and part dimensions build bloom filters on their keys
(bloomfilter_build, ids 2 and 3) and the fact side is probed against both
before the join chain (bloomfilter_probe).  The joins still decide every
match, so the result is the same for any filter size.
"""
import numpy as np

from . import relbuilder as rb

N_DATE, N_SUPP, N_PART, N_CUST = 2556, 200000, 3000000, 800000
FACT = "star/lineorder"
FACT_COLS = [("fk_date", 0, N_DATE), ("fk_supp", 0, N_SUPP), ("fk_part", 0, N_PART),
             ("fk_cust", 0, N_CUST), ("quantity", 1, 50), ("price", 90000, 20001),
             ("discount", 0, 11)]
YEAR, REGION, CATEGORY = 1997, 1, 3


def dims():
    d = np.arange(N_DATE, dtype=np.int32)
    s = np.arange(N_SUPP, dtype=np.int32)
    p = np.arange(N_PART, dtype=np.int32)
    c = np.arange(N_CUST, dtype=np.int32)
    return {
        "star/date": {"d_datekey": d, "d_year": (1992 + d // 365).astype(np.int32)},
        "star/supplier": {"s_suppkey": s, "s_region": s % 5, "s_nation": s % 25},
        "star/part": {"p_partkey": p, "p_category": p % 25},
        "star/customer": {"c_custkey": c, "c_city": c % 160},
    }


def plan(sort=True, bloom_size=0):
    """The query as a JSON plan (dict); bloom_size > 0 adds the bloom
    filters (a power of two entries each)."""
    b = rb.RelBuilder()
    col = rb.col

    def dim(rel, attrs, pred=None):
        r = b.scan(rel, attrs).to_gpu().unpack()
        return r.filter(pred) if pred is not None else r

    date = dim("star/date", ["d_datekey", "d_year"],
               rb.eq(col("star/date", "d_year"), rb.lit(YEAR)))
    supp = dim("star/supplier", ["s_suppkey", "s_region", "s_nation"],
               rb.eq(col("star/supplier", "s_region"), rb.lit(REGION)))
    part = dim("star/part", ["p_partkey", "p_category"],
               rb.eq(col("star/part", "p_category"), rb.lit(CATEGORY)))
    cust = dim("star/customer", ["c_custkey", "c_city"])
    f = b.scan(FACT, [c[0] for c in FACT_COLS]).to_gpu().unpack()
    if bloom_size:
        supp = supp.bloomfilter_build(col("star/supplier", "s_suppkey"), bloom_size, 2)
        part = part.bloomfilter_build(col("star/part", "p_partkey"), bloom_size, 3)
        f = f.bloomfilter_probe(col(FACT, "fk_part"), bloom_size, 3) \
             .bloomfilter_probe(col(FACT, "fk_supp"), bloom_size, 2)
    j = f.join(date, col("star/date", "d_datekey"), col(FACT, "fk_date"), 13, N_DATE)
    j = j.join(supp, col("star/supplier", "s_suppkey"), col(FACT, "fk_supp"), 19, N_SUPP)
    j = j.join(part, col("star/part", "p_partkey"), col(FACT, "fk_part"), 23, N_PART)
    j = j.join(cust, col("star/customer", "c_custkey"), col(FACT, "fk_cust"), 21, N_CUST)
    rev = rb.cast(rb.mul(col(FACT, "price"), rb.sub(rb.lit(100), col(FACT, "discount"))),
                  "int64")
    g = j.groupby([col("star/supplier", "s_nation"), col("star/customer", "c_city")],
                  [(rb.as_(rev, "agg", "revenue"), "sum")], 10, 4096)
    if sort:
        g = g.sort([col("star/supplier", "s_nation"), col("star/customer", "c_city"),
                    col("agg", "revenue")], ["ASC", "ASC", "NONE"])
    return g.to_cpu().print().plan()


class Star:
    def __init__(self, ctx, n, seed):
        self.ctx, self.n, self.seed = ctx, n, seed
        self.fact = {}
        for i, (name, lo, rng) in enumerate(FACT_COLS):
            self.fact[name] = ctx.gen_uniform_i32(n, seed, 10 + i, lo, rng)
            ctx.register_column(FACT, name, "int", self.fact[name])
        self.dim_dev = []
        for rel, cols in dims().items():
            for name, v in cols.items():
                d = ctx.to_device(v.astype(np.int32))
                self.dim_dev.append(d)
                ctx.register_column(rel, name, "int", d)

    def plan(self, sort=True, bloom_size=0):
        return plan(sort, bloom_size)

    def close(self):
        for v in list(self.fact.values()) + self.dim_dev:
            v.free()
        self.fact, self.dim_dev = {}, []
        self.ctx.clear_catalog()
