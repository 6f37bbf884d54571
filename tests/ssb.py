"""Seeded synthetic Star Schema Benchmark relations written as prepared
("block") relations, the layout of the reference's tools/prepare_data.py:
374-443 and 494-526:

  <dir>/<rel>.csv.<attr>        headerless native little-endian values:
                                int32 for "int", int64 for "int64"/"date",
                                FP64 for "float", int32 codes for strings
  <dir>/<rel>.csv.<attr>.dict   for strings: the distinct values sorted, one
                                "value:code" line each, code = rank from 0
  <dir>/catalog.json            {"<schema>_<rel>": {"path": "inputs/<schema>/
                                <rel>.csv", "type": bag of record attributes
                                (type, relName, attrName, attrNo), "plugin":
                                {"type": "block", "linehint": rows}}}

The attribute value domains follow the SSB specification closely enough for
every filter of the SSB queries to select rows (regions, nations, cities as
nation[:9] + digit, MFGR#m / MFGR#mc / MFGR#mcbb, dates 1992-1998); value
ranges are kept small so every aggregate of the bare queries fits in int32."""
import datetime
import json
import os

import numpy as np

REGIONS = {
    "AFRICA": ["ALGERIA", "ETHIOPIA", "KENYA", "MOROCCO", "MOZAMBIQUE"],
    "AMERICA": ["ARGENTINA", "BRAZIL", "CANADA", "PERU", "UNITED STATES"],
    "ASIA": ["CHINA", "INDIA", "INDONESIA", "JAPAN", "VIETNAM"],
    "EUROPE": ["FRANCE", "GERMANY", "ROMANIA", "RUSSIA", "UNITED KINGDOM"],
    "MIDDLE EAST": ["EGYPT", "IRAN", "IRAQ", "JORDAN", "SAUDI ARABIA"],
}
MONTHS = ["Jan", "Feb", "Mar", "Apr", "May", "Jun", "Jul", "Aug", "Sep", "Oct", "Nov", "Dec"]


def _places(rs, n):
    nations = [(r, x) for r in REGIONS for x in REGIONS[r]]
    pick = rs.integers(0, len(nations), n)
    digit = rs.integers(0, 10, n)
    region = np.array([nations[i][0] for i in pick], dtype=object)
    nation = np.array([nations[i][1] for i in pick], dtype=object)
    city = np.array([nations[i][1][:9].ljust(9) + str(d) for i, d in zip(pick, digit)], dtype=object)
    return region, nation, city


def tables(lineorder_rows, seed=1, parts=20000, suppliers=2000, customers=3000):
    """Columns of date, part, supplier, customer and lineorder: numpy arrays
    (int32) or object arrays of strings."""
    rs = np.random.default_rng(seed)
    day0 = datetime.date(1992, 1, 1)
    days = [day0 + datetime.timedelta(d) for d in range((datetime.date(1998, 12, 31) - day0).days + 1)]
    date = {
        "d_datekey": np.array([d.year * 10000 + d.month * 100 + d.day for d in days], np.int32),
        "d_year": np.array([d.year for d in days], np.int32),
        "d_yearmonthnum": np.array([d.year * 100 + d.month for d in days], np.int32),
        "d_yearmonth": np.array([MONTHS[d.month - 1] + str(d.year) for d in days], dtype=object),
        "d_weeknuminyear": np.array([(d.timetuple().tm_yday - 1) // 7 + 1 for d in days], np.int32),
    }
    m = rs.integers(1, 6, parts)
    c = rs.integers(1, 6, parts)
    b = rs.integers(1, 41, parts)
    part = {
        "p_partkey": np.arange(1, parts + 1, dtype=np.int32),
        "p_mfgr": np.array(["MFGR#%d" % x for x in m], dtype=object),
        "p_category": np.array(["MFGR#%d%d" % (x, y) for x, y in zip(m, c)], dtype=object),
        "p_brand1": np.array(["MFGR#%d%d%02d" % (x, y, z) for x, y, z in zip(m, c, b)], dtype=object),
        "p_size": rs.integers(1, 51, parts).astype(np.int32),
    }
    sr, sn, sc = _places(rs, suppliers)
    supplier = {"s_suppkey": np.arange(1, suppliers + 1, dtype=np.int32),
                "s_city": sc, "s_nation": sn, "s_region": sr}
    cr, cn, cc = _places(rs, customers)
    customer = {"c_custkey": np.arange(1, customers + 1, dtype=np.int32),
                "c_city": cc, "c_nation": cn, "c_region": cr}
    n = lineorder_rows
    lineorder = {
        "lo_custkey": rs.integers(1, customers + 1, n).astype(np.int32),
        "lo_partkey": rs.integers(1, parts + 1, n).astype(np.int32),
        "lo_suppkey": rs.integers(1, suppliers + 1, n).astype(np.int32),
        "lo_orderdate": date["d_datekey"][rs.integers(0, len(days), n)],
        "lo_quantity": rs.integers(1, 51, n).astype(np.int32),
        "lo_extendedprice": rs.integers(100, 2000, n).astype(np.int32),
        "lo_discount": rs.integers(0, 11, n).astype(np.int32),
        "lo_revenue": rs.integers(100, 2000, n).astype(np.int32),
        "lo_supplycost": rs.integers(50, 1000, n).astype(np.int32),
    }
    return {"date": date, "part": part, "supplier": supplier, "customer": customer,
            "lineorder": lineorder}


def write(root, schema, tabs):
    """Writes the relations under root/inputs/<schema>/ and returns the
    catalog written there."""
    d = os.path.join(root, "inputs", schema)
    os.makedirs(d, exist_ok=True)
    catalog = {}
    for rel, cols in tabs.items():
        path = "inputs/%s/%s.csv" % (schema, rel)
        attrs = []
        n = None
        for no, (attr, v) in enumerate(cols.items(), 1):
            f = os.path.join(root, path + "." + attr)
            n = len(v)
            if v.dtype == object:
                values = sorted(set(v.tolist()))
                code = {s: i for i, s in enumerate(values)}
                with open(f + ".dict", "w") as fh:
                    fh.writelines("%s:%d\n" % (s, i) for i, s in enumerate(values))
                np.array([code[s] for s in v], np.int32).tofile(f)
                t = "dstring"
            else:
                t = {np.dtype(np.int32): "int", np.dtype(np.int64): "int64",
                     np.dtype(np.float64): "float"}[v.dtype]
                v.astype(v.dtype.newbyteorder("<")).tofile(f)
            attrs.append({"type": {"type": t}, "relName": path, "attrName": attr, "attrNo": no})
        catalog["%s_%s" % (schema, rel)] = {
            "path": path, "type": {"type": "bag", "inner": {"type": "record", "attributes": attrs}},
            "plugin": {"type": "block", "linehint": n}}
    with open(os.path.join(d, "catalog.json"), "w") as fh:
        json.dump(catalog, fh, indent=1)
    return catalog
