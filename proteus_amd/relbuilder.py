"""RelBuilder: builds physical plans in the reference's JSON plan language and
runs them through pj_plan_execute.

Mirrors the call shapes of the reference's C++ RelBuilder
(core/olap/include/olap/operators/relbuilder.hpp:110-609): scan, to_gpu,
to_cpu, unpack, pack, filter, project, join, groupby, reduce, sort, print,
bloomfilter_build/_probe/_repack, prepare -> PreparedStatement.execute() ->
QueryResult.  Expressions are the
plan language's JSON objects (core/olap/lib/plan/expression-parser.cpp), made
with the helpers below.  `prepare` lowers to the C ABI instead of a JIT.
"""
import json


# ---- expressions -----------------------------------------------------------
def col(rel, attr):
    """recordProjection of rel.attr"""
    return {"expression": "recordProjection", "e": {"expression": "argument"},
            "attribute": {"relName": rel, "attrName": attr}}


def lit(v, type_=None):
    if type_ is None:
        type_ = "bool" if isinstance(v, bool) else "float" if isinstance(v, float) else "int"
    return {"expression": type_, "v": v}


def int64(v):
    return {"expression": "int64", "v": int(v)}


def cast(e, type_):
    return {"expression": "cast", "type": {"type": type_}, "e": e}


def _bin(op):
    def f(left, right):
        return {"expression": op, "left": left, "right": right}
    f.__name__ = op
    return f


add, sub, mul, div, mod = _bin("add"), _bin("sub"), _bin("multiply"), _bin("div"), _bin("mod")
eq, ne, lt, le, gt, ge = _bin("eq"), _bin("neq"), _bin("lt"), _bin("le"), _bin("gt"), _bin("ge")
and_, or_ = _bin("and"), _bin("or")


def neg(e):
    return {"expression": "neg", "e": e}


def if_(cond, then, else_):
    return {"expression": "if", "cond": cond, "then": then, "else": else_}


def as_(e, rel, attr):
    """register the expression's value as rel.attr (register_as)"""
    e = dict(e)
    e["register_as"] = {"relName": rel, "attrName": attr}
    return e


# ---- plans -----------------------------------------------------------------
class QueryResult:
    def __init__(self, text):
        self.text = text

    def rows(self):
        return [ln.split(",") for ln in self.text.splitlines() if ln]

    def __str__(self):
        return self.text


class PreparedStatement:
    def __init__(self, plan, ctx, base_dir):
        self.plan, self.ctx, self.base_dir = plan, ctx, base_dir
        self.text = json.dumps(plan)

    def execute(self):
        return QueryResult(self.ctx.execute_plan(self.text, self.base_dir))


class RelBuilder:
    def __init__(self, ctx=None, base_dir=".", node=None):
        self.ctx, self.base_dir, self.node = ctx, base_dir, node

    def _wrap(self, node):
        return RelBuilder(self.ctx, self.base_dir, node)

    def _on(self, op, **fields):
        n = {"operator": op}
        n.update(fields)
        n["input"] = self.node
        return self._wrap(n)

    # sources
    def scan(self, rel, attrs, plugin="block", **plugin_fields):
        pl = {"name": rel, "type": plugin,
              "projections": [{"relName": rel, "attrName": a} for a in attrs]}
        pl.update(plugin_fields)
        return self._wrap({"operator": "scan", "plugin": pl})

    def scan_csv(self, rel, attrs, delimiter=";"):
        """attrs: [(name, type, position)]"""
        pl = {"name": rel, "type": "csv", "delimiter": delimiter, "brackets": False,
              "projections": [{"relName": rel, "attrName": a, "type": {"type": t}, "attrNo": no}
                              for a, t, no in attrs]}
        return self._wrap({"operator": "scan", "plugin": pl})

    # placement (no work on one GPU with HBM-resident columns)
    def to_gpu(self):
        return self._on("cpu-to-gpu")

    def to_cpu(self):
        return self._on("gpu-to-cpu")

    def unpack(self):
        return self._on("unpack")

    def pack(self, projections=None, hash_expr=None, num_buckets=None):
        if hash_expr is None:
            return self._on("pack")
        return self._on("hash-rearrange", projections=projections, e=hash_expr,
                        buckets=num_buckets)

    # relational
    def filter(self, pred):
        return self._on("select", p=pred)

    def project(self, exprs, rel_name="project"):
        return self._on("project", e=exprs, relName=rel_name)

    def join(self, build, build_k, probe_k, hash_bits=None, max_build=None):
        """Chained hash join; self is the probe side."""
        n = {"operator": "hashjoin-chained", "build_input": build.node, "probe_input": self.node,
             "build_k": build_k, "probe_k": probe_k}
        if hash_bits is not None:
            n["hash_bits"] = hash_bits
        if max_build is not None:
            n["maxBuildInputSize"] = max_build
        return self._wrap(n)

    def partitioned_join(self, build, build_k, probe_k, max_build, max_probe):
        n = {"operator": "partitioned-hashjoin-chained", "build_input": build.node,
             "probe_input": self.node, "build_k": build_k, "probe_k": probe_k,
             "maxBuildInputSize": max_build, "maxProbeInputSize": max_probe}
        return self._wrap(n)

    def out_of_gpu_join(self, build, build_e, probe_e, num_buckets, max_build, max_probe):
        """out-of-gpu-join (plan-parser.cpp:531-850); self is the probe side.
        build_e / probe_e: the attributes carried, the first one the join key."""
        n = {"operator": "out-of-gpu-join", "build_input": build.node, "probe_input": self.node,
             "numOfBuckets": num_buckets, "numPartitioners": 1, "numConcurrent": 1,
             "maxBuildInputSize": max_build, "maxProbeInputSize": max_probe, "slack": 8,
             "build_e": [{"original": e} for e in build_e],
             "probe_e": [{"original": e} for e in probe_e]}
        return self._wrap(n)

    def groupby(self, k, e, hash_bits, max_input_size):
        """k: [expr]; e: [(expr, monoid)]"""
        return self._on("groupby", k=k, e=[{"e": x, "m": m, "packet": i + 1, "offset": 0}
                                           for i, (x, m) in enumerate(e)],
                        hash_bits=hash_bits, maxInputSize=max_input_size)

    def reduce(self, exprs, accs, pred=None):
        """Aggregates over the rows where pred holds (all rows by default)."""
        return self._on("reduce", e=exprs, accumulator=accs, p=lit(True) if pred is None else pred)

    # bloom filters: RelBuilder::bloomfilter_build / _probe / _repack
    # (relbuilder.hpp:427-462, relbuilder.cpp:957-975), as the operators
    # bloomfilter-build / -probe / -repack.  pred is the int32 key expression.
    def bloomfilter_build(self, pred, filter_size, bloom_id):
        """Sets the key's entry of filter bloom_id; passes every row on."""
        return self._on("bloomfilter-build", e=pred, filterSize=filter_size, bloomId=bloom_id)

    def bloomfilter_probe(self, pred, filter_size, bloom_id):
        """Keeps the rows whose key's entry is set."""
        return self._on("bloomfilter-probe", e=pred, filterSize=filter_size, bloomId=bloom_id)

    def bloomfilter_repack(self, pred, filter_size, bloom_id, attrs=None):
        """Keeps the rows whose key's entry is set, repacked to attrs (every
        input attribute when attrs is None)."""
        fields = dict(e=pred, filterSize=filter_size, bloomId=bloom_id)
        if attrs is not None:
            fields["projections"] = attrs
        return self._on("bloomfilter-repack", **fields)

    def sort(self, exprs, dirs):
        return self._on("sort", e=[{"expression": x, "direction": d} for x, d in zip(exprs, dirs)])

    def print(self, exprs=None, fmt="pm-csv"):
        fields = {"plugin": {"type": fmt}}
        if exprs is not None:
            fields["e"] = exprs
        return self._on("print", **fields)

    def plan(self):
        return self.node

    def prepare(self):
        return PreparedStatement(self.node, self.ctx, self.base_dir)
