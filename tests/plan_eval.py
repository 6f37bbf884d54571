"""A numpy evaluator of the physical-plan subset the SSB bare plans use
(scan of prepared "block" relations, select, project, hashjoin-chained,
reduce, print and the placement operators, which pass their input through),
used as the expected side of the plan tests over column files.  It follows the
plan language's meaning (core/olap/lib/plan/plan-parser.cpp,
expression-parser.cpp): attributes resolve by relation and name, else by name;
a join emits every (build, probe) pair with equal keys; arithmetic and sums
are exact in int64.  Test infrastructure only."""
import json
import os

import numpy as np

PASS = {"unpack", "block-to-tuples", "pack", "tuples-to-block", "cpu-to-gpu", "gpu-to-cpu",
        "mem-move-device", "mem-move-local-to", "mem-broadcast-device", "router", "exchange",
        "split"}
DTYPES = {"int": np.int32, "dstring": np.int32, "int64": np.int64, "date": np.int64,
          "float": np.float64}


class Rel:
    def __init__(self, n, cols):
        self.n = n
        self.cols = cols  # list of ((rel, name), array)

    def get(self, rel, name):
        for (r, a), v in self.cols:
            if r == rel and a == name:
                return v
        for (r, a), v in self.cols:
            if a == name:
                return v
        raise KeyError("%s.%s" % (rel, name))

    def take(self, idx):
        return Rel(len(idx), [(k, v[idx]) for k, v in self.cols])


class Evaluator:
    def __init__(self, root):
        self.root = root

    def _type(self, rel, attr):
        cat = os.path.join(self.root, os.path.dirname(rel), "catalog.json")
        with open(cat) as f:
            for entry in json.load(f).values():
                if entry["path"] != rel:
                    continue
                for a in entry["type"]["inner"]["attributes"]:
                    if a["attrName"] == attr:
                        return a["type"]["type"]
        raise KeyError(rel + "." + attr)

    def _dict_code(self, value, path):
        with open(os.path.join(self.root, path)) as f:
            for line in f:
                v, _, code = line.rstrip("\n").rpartition(":")
                if v == value:
                    return int(code)
        raise KeyError(value)

    # ---- expressions ------------------------------------------------------
    def expr(self, e, rel):
        k = e["expression"]
        if k == "recordProjection":
            v = rel.get(e["attribute"]["relName"], e["attribute"]["attrName"])
            return v.astype(np.float64 if v.dtype == np.float64 else np.int64)
        if k in ("int", "int64"):
            return np.int64(e["v"])
        if k == "float":
            return np.float64(e["v"])
        if k == "bool":
            return np.bool_(e["v"])
        if k == "dstring":
            return np.int64(self._dict_code(e["v"], e["dict"]["path"]))
        if k == "cast":
            v = self.expr(e["e"], rel)
            return v.astype(np.float64) if e["type"]["type"] == "float" else v.astype(np.int64)
        if k == "neg":
            return -self.expr(e["e"], rel)
        a, b = self.expr(e["left"], rel), self.expr(e["right"], rel)
        ops = {"eq": np.equal, "neq": np.not_equal, "lt": np.less, "le": np.less_equal,
               "gt": np.greater, "ge": np.greater_equal, "and": np.logical_and,
               "or": np.logical_or, "add": np.add, "sub": np.subtract, "multiply": np.multiply}
        return ops[k](a, b)

    def _full(self, v, n):
        return np.broadcast_to(v, (n,)) if np.ndim(v) == 0 else v

    # ---- operators --------------------------------------------------------
    def run(self, op):
        k = op["operator"]
        if k in PASS:
            return self.run(op["input"])
        if k == "scan":
            pl = op["plugin"]
            name = pl["name"]
            cols = []
            for p in pl["projections"]:
                a = p["attrName"]
                v = np.fromfile(os.path.join(self.root, name + "." + a), DTYPES[self._type(name, a)])
                cols.append(((name, a), v))
            return Rel(len(cols[0][1]), cols)
        if k == "select":
            r = self.run(op["input"])
            return r.take(np.flatnonzero(self._full(self.expr(op["p"], r), r.n)))
        if k == "project":
            r = self.run(op["input"])
            cols = []
            for e in op["e"]:
                ra = e["register_as"]
                cols.append(((ra["relName"], ra["attrName"]), self._full(self.expr(e, r), r.n)))
            return Rel(r.n, cols)
        if k == "hashjoin-chained":
            b, p = self.run(op["build_input"]), self.run(op["probe_input"])
            bk = self._full(self.expr(op["build_k"], b), b.n)
            pk = self._full(self.expr(op["probe_k"], p), p.n)
            order = np.argsort(bk, kind="stable")
            sb = bk[order]
            lo = np.searchsorted(sb, pk, "left")
            hi = np.searchsorted(sb, pk, "right")
            cnt = hi - lo
            pi = np.repeat(np.arange(p.n), cnt)
            start = np.repeat(lo - np.concatenate(([0], np.cumsum(cnt)[:-1])), cnt)
            bi = order[start + np.arange(len(pi))]
            return Rel(len(pi), b.take(bi).cols + p.take(pi).cols)
        if k == "reduce":
            r = self.run(op["input"])
            keep = self._full(self.expr(op["p"], r), r.n).astype(bool)
            cols = []
            for e, m in zip(op["e"], op["accumulator"]):
                v = self._full(self.expr(e, r), r.n)[keep]
                if m == "sum":
                    x = v.sum()
                elif m == "max":
                    x = v.max() if len(v) else None
                else:
                    x = v.min() if len(v) else None
                ra = e.get("register_as", {"relName": "reduce", "attrName": "x"})
                cols.append(((ra["relName"], ra["attrName"]), np.array([x])))
            return Rel(1, cols)
        if k == "print":
            r = self.run(op["input"])
            if "e" in op:
                return [tuple(self._full(self.expr(e, r), r.n)[i].item() for e in op["e"])
                        for i in range(r.n)]
            return [tuple(v[i].item() for _, v in r.cols) for i in range(r.n)]
        raise NotImplementedError(k)
