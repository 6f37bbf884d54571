"""Benchmark of the MI355X join hot path.

Workload (one step): the radix-partitioned equi-join R(2^27 unique int32 keys,
int32 payload) JOIN S(2^30 int32 fk drawn uniform from [0, 1.33*2^27), int32
payload) with COUNT(*) and the int64 checksum SUM(R.payload + S.payload) --
partitioned-hashjoin-chained -> reduce.  Inputs are generated in HBM before
the timed region (seeded, synthetic).  value = probe tuples per second of the
whole join (both sides partitioned, joined and reduced).

Prints ONE JSON line.  Only one GPU is used; --gpus N > 1 exits with a message.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PCIE_PEAK_GBS = 64.0  # PCIe 5.0 x16, one direction
NR, NS = 1 << 27, 1 << 30
RANGE = int(1.33 * NR)  # 178,509,578: ~24.8% probe misses
SEED = 20251203


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--variant", choices=["uniform", "zipf"], default="uniform")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the secondary measurements (Zipf fk variant)")
    ap.add_argument("--blocks", default="c1,groupby,star,sort,oog,staging",
                    help="secondary measurements to run (comma-separated subset of "
                         "c1,groupby,star,sort,oog,staging)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched / served query mix")
    ap.add_argument("--batched-queries", type=int, default=4096)
    ap.add_argument("--qps", type=float, default=200.0,
                    help="offered load of the Poisson arrivals for the latency run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, copy 1 GiB once with the library's copy "
                         "kernel (the known-size launch that scales rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE readings)")
    ap.add_argument("--verify", action="store_true",
                    help="check COUNT/checksum against the dense oracle (host, full size)")
    return ap.parse_args()


def kernel_table(ctx, names):
    out = {}
    for n in names:
        t, c = ctx.prof_get(n)
        if c:
            out[n] = (t, c)
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported (the join "
              "would shard as a broadcast build side with 1/N of the probe side per GPU)"
              % a.gpus, file=sys.stderr)
        sys.exit(2)
    import proteus_amd as pa

    ctx = pa.Context(0)
    info = ctx.device_info()
    rk = ctx.gen_perm_i32(NR, SEED)
    rv = ctx.gen_uniform_i32(NR, SEED, 1, 0, 1 << 31)
    if a.variant == "uniform":
        sk = ctx.gen_uniform_i32(NS, SEED, 2, 0, RANGE)
    else:
        sk = ctx.gen_zipf_fk_i32(NS, SEED, 2, NR, RANGE, 1.0, SEED)
    sv = ctx.gen_uniform_i32(NS, SEED, 3, 0, 1 << 31)
    ctx.sync()
    cfg = ctx.join_partitioned_config(NR, NS)

    for _ in range(a.warmup):
        res = ctx.join_partitioned_reduce(rk, rv, sk, sv)
    sec, kt, res, per = timed(ctx, a.steps, lambda: ctx.join_partitioned_reduce(rk, rv, sk, sv))
    # per-launch algorithmic bytes: what an ideal kernel must move through HBM
    algo = {
        "radix_scatter1_probe": 16.0 * NS, "radix_scatter2_probe": 16.0 * NS,
        "radix_scatter1_build": 16.0 * NR, "radix_scatter2_build": 16.0 * NR,
        "pjoin_probe": 8.0 * (NR + NS), "radix_hist": 4.0 * (NR + NS) / 2,
    }
    # the dominant kernel is the longest of the probe side's, which run one
    # after another on the context stream and make up the step; the build
    # side's run beside them on a second stream, so their event times
    # include the time they share the device
    critical = ("radix_scatter1_probe", "radix_scatter2_probe", "pjoin_probe")
    dom = max((n for n in kt if n in critical), key=lambda n: kt[n][0] / kt[n][1])
    dom_ms = kt[dom][0] / kt[dom][1]
    achieved = algo[dom] / (dom_ms * 1e-3) / 1e9
    e2e_gbs = 40.0 * (NR + NS) / sec / 1e9

    line = {
        "metric": "join probe Gtuples/s (radix-partitioned 2^27 x 2^30 join -> COUNT, SUM)",
        "value": NS / sec / 1e9,
        "unit": "Gtuples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded in HBM: R keys = permutation of 0..2^27-1, S fk uniform "
                "over [0, 1.33*2^27), payloads uniform [0, 2^31))",
        "config": {"workload": "partitioned-hashjoin-chained -> reduce(COUNT, SUM), "
                               "|R|=2^27, |S|=2^30, %s fk, 8-byte columnar tuples, "
                               "int64 accumulators" % a.variant,
                   "radix": cfg},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": pmc_traffic(dom, a.variant),
                     "algorithmic_bytes": algo[dom], "ms_per_launch": dom_ms},
        "e2e_roofline": {"algorithmic_bytes": 40.0 * (NR + NS), "achieved_gbs": e2e_gbs,
                         "frac": e2e_gbs / HBM_PEAK_GBS},
        "input_gtuples_per_s": (NR + NS) / sec / 1e9,
        "step_ms": {"median": statistics.median(per) * 1e3, "min": min(per) * 1e3,
                    "max": max(per) * 1e3},
        "protocol": protocol(ctx, lambda: ctx.join_partitioned_reduce(rk, rv, sk, sv)),
        "result": res,
        "kernels_ms_per_step": {n: round(v[0] / a.steps, 4) for n, v in kt.items()},
        "device": info,
    }
    if a.verify:
        from oracle import pyoracle as O
        rc, want = O.join_dense(rk.numpy(), rv.numpy(), sk.numpy(), sv.numpy())
        line["verified"] = (rc == 0 and want == res)
    if a.copy_calibration:
        src = ctx.empty(1 << 28, "int32")
        dst = ctx.empty(1 << 28, "int32")
        ctx.copy(dst, src, 1 << 30)
        src.free()
        dst.free()
    if not a.no_extra and a.variant == "uniform":
        # the Zipf(theta=1) fk variant on the same R (same miss rate)
        sk.free()
        sk = ctx.gen_zipf_fk_i32(NS, SEED, 2, NR, RANGE, 1.0, SEED)
        ctx.join_partitioned_reduce(rk, rv, sk, sv)
        zsec, zkt, zres, _ = timed(ctx, a.steps,
                                   lambda: ctx.join_partitioned_reduce(rk, rv, sk, sv))
        line["variants"] = {"zipf_theta1": {
            "ms_per_step": zsec * 1e3, "value": NS / zsec / 1e9, "vs_uniform": zsec / sec,
            "protocol": protocol(ctx, lambda: ctx.join_partitioned_reduce(rk, rv, sk, sv)),
            "kernels_ms_per_step": {n: round(v[0] / a.steps, 4) for n, v in zkt.items()},
            "result": zres}}
    if not a.no_extra:
        for x in (rk, rv, sk, sv):
            x.free()
        rk = rv = sk = sv = None
        blocks = set(a.blocks.split(","))
        if "c1" in blocks:
            line["c1"] = chained_c1(ctx, a.steps)
        if "groupby" in blocks:
            line["groupby"] = groupby_c3(ctx, a.steps)
        if "star" in blocks:
            line["star"] = star_c4(ctx, a.steps)
        if "sort" in blocks:
            line["sort"] = sort_plan(ctx, a.steps)
        if "oog" in blocks:
            line["out_of_gpu_join"] = out_of_gpu(ctx, 2)
        if "staging" in blocks:
            line["staging"] = staging(ctx, 3)
    if not a.no_batched:
        for x in (rk, rv, sk, sv):
            if x is not None:
                x.free()
        line["batched"] = batched_mix(ctx, a.batched_queries, a.qps)
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(a.variant)
    print(json.dumps(line), flush=True)
    ctx.close()


def chained_c1(ctx, steps):
    """The plain equi-join hashjoin-chained -> reduce (pj_join_chained_reduce):
    R = 2^20 permuted int32 keys with int32 payload, S = 2^24 fk uniform over
    R's keys; COUNT = 2^24 (closed form) and the int64 checksum.  hash_bits =
    21 and maxBuildInputSize = 1.1 * |R| (the RelBuilder defaults,
    relbuilder.cpp:653-668).  R's keys are dense (a permutation of
    0..2^20-1), so the entry point joins through a direct table of R's values
    (4 MiB, indexed by key - min; the key range and the density test stay on
    the device); the chained table's own kernels are timed beside it through
    the allocation-free entry point, which always builds the table.
    Algorithmic bytes: 8 per input tuple (each column read once), 142.6 MB."""
    nr, ns = 1 << 20, 1 << 24
    rk = ctx.gen_perm_i32(nr, SEED)
    rv = ctx.gen_uniform_i32(nr, SEED, 1, 0, 1 << 31)
    sk = ctx.gen_uniform_i32(ns, SEED, 2, 0, nr)
    sv = ctx.gen_uniform_i32(ns, SEED, 3, 0, 1 << 31)
    maxb = int(1.1 * nr)

    def run():
        return ctx.join_chained_reduce(rk, rv, sk, sv, 21, maxb)

    def run_table():
        return ctx.join_chained_reduce_noalloc(rk, rv, sk, sv, 21, maxb)
    for _ in range(3):
        res = run()
        run_table()
    chained_names = ["join_chained_build", "join_chained_probe", "fill_u32", "join_acc_init"]
    direct_names = ["rel_minmax", "rel_minmax_fold", "fill_u32", "join_acc_init",
                    "join_direct_build", "join_direct_probe"]
    nsteps = max(steps, 10)
    # kernel times from a run with launch events, step times from one without
    # (at 0.3 ms a step, the events' host work shows in the step time)
    _, kt, _, _ = timed(ctx, nsteps, run, JOIN_KERNELS + chained_names + direct_names)
    _, kt_t, _, _ = timed(ctx, nsteps, run_table, chained_names)
    sec, res, per = untimed_kernels(ctx, nsteps, run)
    sec_t, res_t, _ = untimed_kernels(ctx, nsteps, run_table)
    algo = 8.0 * (nr + ns)
    path = ("partitioned" if "pjoin_probe" in kt else
            "chained" if "join_chained_probe" in kt else "direct")
    out = {"build_rows": nr, "probe_rows": ns, "hash_bits": 21, "path": path, "ms": sec * 1e3,
           "step_ms_median": statistics.median(per) * 1e3,
           "input_gtuples_per_s": (nr + ns) / sec / 1e9, "probe_gtuples_per_s": ns / sec / 1e9,
           "roofline": {"bound": "hbm", "algorithmic_bytes": algo, "achieved_gbs": algo / sec / 1e9,
                        "frac": algo / sec / 1e9 / HBM_PEAK_GBS},
           "kernels_ms_per_step": {n: round(v[0] / nsteps, 4) for n, v in kt.items()},
           "protocol": protocol(ctx, run), "result": res, "count_is_closed_form": res["count"] == ns,
           "chained_table": {"ms": sec_t * 1e3,
                             "probe_kernel_ms": kt_t["join_chained_probe"][0] / kt_t["join_chained_probe"][1],
                             "kernels_ms_per_step": {n: round(v[0] / nsteps, 4) for n, v in kt_t.items()},
                             "identical": res_t == res}}
    for x in (rk, rv, sk, sv):
        x.free()
    return out


def groupby_c3(ctx, steps):
    """Hash group-by of 2^29 rows (int32 key, int64 v uniform [0, 2^20),
    float32 w = u/2^24): COUNT, SUM, MIN, MAX of v and FP64 SUM of w per
    group, for 2^10 / 2^16 / 2^22 groups, uniform and Zipf(0.8) keys (dense:
    direct LDS tables) and the uniform keys spread over all of int32 by a
    bijection (sparse: hashed LDS tables); each timed through the C ABI
    (pj_groupby_i32), and the dense ones also as a plan through
    pj_plan_execute (scan -> groupby -> reduce(COUNT(*), SUM(cnt)), the
    group-by lowered to the same LDS kernels; the reduce keeps the printed
    result small).  Algorithmic bytes: 16 per row read + 44 per group
    written.  Checked by conservation: the counts sum to the row count and
    the SUMs to SUM(v); the plan's row and count totals must agree."""
    import ctypes as C
    import numpy as np
    import proteus_amd as pa
    from proteus_amd import relbuilder as rb
    lib = pa.lib()
    n = 1 << 29
    v = ctx.gen_uniform_i64(n, SEED, 5, 0, 1 << 20)
    w = ctx.gen_unit_f32(n, SEED, 6)
    vsum = ctx.reduce_i64(v)[0]
    out = {"rows": n, "configs": []}
    for g in (1 << 10, 1 << 16, 1 << 22):
        outs = [ctx.empty(g, t) for t in (np.int32, np.int64, np.int64, np.int64, np.int64,
                                          np.float64)]
        for dist in ("uniform", "zipf0.8", "sparse"):
            if dist == "uniform":
                k = ctx.gen_uniform_i32(n, SEED, 4, 0, g)
            elif dist == "zipf0.8":
                k = ctx.gen_zipf_i32(n, SEED, 4, g, 0.8, SEED)
            else:
                # the uniform keys times an odd constant (a bijection of
                # int32): the same groups spread over all of int32, so the
                # hashed tables run instead of the direct ones
                k0 = ctx.gen_uniform_i32(n, SEED, 4, 0, g)
                kh = (k0.numpy().view(np.uint32) * np.uint32(2654435761)).view(np.int32)
                k0.free()
                k = ctx.to_device(kh)
                del kh
            ng = C.c_uint64()
            hb = max(1, (g - 1).bit_length())

            def run():
                pa._check(lib.pj_groupby_i32(ctx.h, k.ptr, v.ptr, w.ptr, n, hb, g, outs[0].ptr,
                                             outs[1].ptr, outs[2].ptr, outs[3].ptr, outs[4].ptr,
                                             outs[5].ptr, C.byref(ng)))
            run()
            names = ["gb_key_minmax", "radix_hist", "radix_colsum", "radix_colscan",
                     "radix_scan", "radix_items", "gb_scatter1", "gb_scatter2", "gb_agg",
                     "gb_merge", "gb_compact"]
            sec, kt, _, _ = timed(ctx, steps, run, names)
            m = ng.value
            cnt = outs[1].numpy()[:m]
            sums = outs[2].numpy()[:m]
            ok = int(cnt.sum()) == n and int(sums.sum()) == vsum
            gbs = (16.0 * n + 44.0 * m) / sec / 1e9
            entry = {
                "groups": g, "keys": dist, "ms": sec * 1e3, "rows_per_s": n / sec,
                "groups_out": m, "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
                "tables": "hashed" if dist == "sparse" else "direct",
                "conservation_ok": ok,
                "kernels_ms_per_call": {x: round(t / steps, 4) for x, (t, c) in kt.items()}}
            out["configs"].append(entry)
            if dist == "sparse":
                k.free()
                continue
            # the same group-by as a plan
            ctx.clear_catalog()
            ctx.register_column("C3", "k", "int", k)
            ctx.register_column("C3", "v", "int64", v)
            ctx.register_column("C3", "w", "float32", w)
            q = rb.RelBuilder(ctx).scan("C3", ["k", "v", "w"]).to_gpu().unpack() \
                .groupby([rb.col("C3", "k")],
                         [(rb.as_(rb.int64(1), "g", "cnt"), "sum"),
                          (rb.as_(rb.col("C3", "v"), "g", "s"), "sum"),
                          (rb.as_(rb.col("C3", "v"), "g", "mn"), "min"),
                          (rb.as_(rb.col("C3", "v"), "g", "mx"), "max"),
                          (rb.as_(rb.col("C3", "w"), "g", "fs"), "sum")], hb, g) \
                .reduce([rb.int64(1), rb.col("g", "cnt"), rb.col("g", "s")],
                        ["sum", "sum", "sum"])
            text = json.dumps(q.plan())
            res = [int(x) for x in ctx.execute_plan(text).strip().split(",")]
            psec, pkt, _, _ = timed(ctx, steps, lambda: ctx.execute_plan(text),
                                    ["gb_agg", "rel_groupby"])
            ctx.clear_catalog()
            entry.update({"plan_ms": psec * 1e3, "plan_vs_direct": psec / sec,
                          "plan_kernels": {x: c for x, (t, c) in pkt.items()},
                          "plan_result_ok": res == [m, n, vsum]})
            k.free()
        for o in outs:
            o.free()
    v.free()
    w.free()
    return out


def out_of_gpu(ctx, steps):
    """out-of-gpu-join -> reduce over host-resident inputs, |R| = 2^26,
    |S| = 2^29 (the C2 distributions at half size, 4.5 GB in host memory),
    under an 8 GiB HBM budget: bucketed on the device through a pinned host
    arena, then joined bucket by bucket.  The rate includes every PCIe
    transfer (24 B per tuple).  Checked against the same join on device
    columns."""
    nr, ns = 1 << 26, 1 << 29
    rng = int(1.33 * nr)
    dev = [ctx.gen_perm_i32(nr, SEED), ctx.gen_uniform_i32(nr, SEED, 1, 0, 1 << 31),
           ctx.gen_uniform_i32(ns, SEED, 2, 0, rng), ctx.gen_uniform_i32(ns, SEED, 3, 0, 1 << 31)]
    ref = ctx.join_partitioned_reduce(*dev)
    host = [d.numpy() for d in dev]
    for d in dev:
        d.free()
    budget = 8 << 30
    got = ctx.join_out_of_gpu_reduce(*host, hbm_budget=budget)
    m0 = ctx.mem_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.join_out_of_gpu_reduce(*host, hbm_budget=budget)
    sec = (time.perf_counter() - t0) / steps
    m1 = ctx.mem_stats()
    return {"build_rows": nr, "probe_rows": ns, "hbm_budget_bytes": budget, "ms": sec * 1e3,
            "input_gtuples_per_s": (nr + ns) / sec / 1e9,
            "pcie_gbs": 24.0 * (nr + ns) / sec / 1e9, "identical": got == ref,
            "pool": {"mallocs": m1["mallocs"] - m0["mallocs"],
                     "evictions": m1["evictions"] - m0["evictions"], "cached": m1["cached"]}}


def staging(ctx, steps):
    """Host -> HBM staging through the pinned ring (pj_stage_file /
    pj_stage_begin), against one hipMemcpy from the same pageable array
    (pj_memcpy_h2d), over 2 GiB; then a prepared relation of 2^27 rows
    (three int32 column files, 1.5 GiB) scanned by the plan
    reduce(SUM(a * b)) where c < 25: the first run stages the files, later
    runs read the columns kept in HBM.  Files are written to a temporary
    directory (and so sit in the page cache); rates are bytes moved over
    PCIe per second."""
    import json as _json
    import shutil
    import tempfile

    import numpy as np

    from proteus_amd import relbuilder as rb
    nbytes = 1 << 31
    host = np.random.default_rng(SEED).integers(0, 1 << 31, nbytes // 4, dtype=np.int32)
    tmp = tempfile.mkdtemp(prefix="pj_stage_")
    out = {"bytes": nbytes, "pcie_peak_gbs": PCIE_PEAK_GBS}
    try:
        f = os.path.join(tmp, "blob.bin")
        host.tofile(f)
        d = ctx.empty(nbytes // 4, "int32")

        def rate(fn):
            fn()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                fn()
            ctx.sync()
            return nbytes / ((time.perf_counter() - t0) / steps) / 1e9

        out["file_gbs"] = rate(lambda: ctx.stage_file(f, d, nbytes))
        out["file_identical"] = bool(np.array_equal(d.numpy()[:1 << 20], host[:1 << 20]))
        out["host_gbs"] = rate(lambda: ctx.stage_wait(ctx.stage_begin(d, nbytes, host=host)))
        out["hipmemcpy_pageable_gbs"] = rate(lambda: ctx.copy_h2d(d, host))
        d.free()
        os.remove(f)
        # a prepared relation: <rel>.<attr> files and a catalog next to them
        n = 1 << 27
        rs = np.random.default_rng(SEED + 1)
        cols = {"a": rs.integers(0, 1000, n, dtype=np.int32),
                "b": rs.integers(0, 1000, n, dtype=np.int32),
                "c": rs.integers(0, 50, n, dtype=np.int32)}
        rel = "inputs/blk/fact.csv"
        os.makedirs(os.path.join(tmp, "inputs", "blk"))
        attrs = []
        for no, (name, v) in enumerate(cols.items(), 1):
            v.tofile(os.path.join(tmp, rel + "." + name))
            attrs.append({"type": {"type": "int"}, "relName": rel, "attrName": name, "attrNo": no})
        with open(os.path.join(tmp, "inputs", "blk", "catalog.json"), "w") as fh:
            _json.dump({"blk_fact": {"path": rel, "type": {"type": "bag", "inner": {
                "type": "record", "attributes": attrs}}, "plugin": {"type": "block"}}}, fh)
        m = cols["c"] < 25
        want = int((cols["a"][m].astype(np.int64) * cols["b"][m]).sum())
        col = lambda a: rb.col(rel, a)  # noqa: E731
        q = rb.RelBuilder(ctx, tmp).scan(rel, ["a", "b", "c"]).to_gpu().unpack() \
            .filter(rb.lt(col("c"), rb.lit(25))) \
            .reduce([rb.mul(rb.cast(col("a"), "int64"), rb.cast(col("b"), "int64"))], ["sum"])
        text = _json.dumps(q.plan())
        ctx.clear_catalog()
        ctx.sync()
        t0 = time.perf_counter()
        got = int(ctx.execute_plan(text, base_dir=tmp).strip())
        cold = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.execute_plan(text, base_dir=tmp)
        warm = (time.perf_counter() - t0) / steps
        ctx.clear_catalog()
        fbytes = 12 * n
        out["block_plan"] = {"rows": n, "file_bytes": fbytes, "first_run_ms": cold * 1e3,
                             "first_run_gbs": fbytes / cold / 1e9, "resident_run_ms": warm * 1e3,
                             "resident_hbm_gbs": fbytes / warm / 1e9, "identical": got == want}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    out["frac"] = max(out.get("file_gbs", 0), out.get("host_gbs", 0)) / PCIE_PEAK_GBS
    return out


def sort_plan(ctx, steps):
    """The plan operator sort on the device: 2^25 rows ordered by an int32
    key ascending (2^20 distinct values) then an FP64 key descending, through
    pj_plan_execute (scan -> sort -> reduce COUNT)."""
    import json as _json
    from proteus_amd import relbuilder as rb
    n = 1 << 25
    k = ctx.gen_uniform_i32(n, SEED, 20, 0, 1 << 20)
    w = ctx.gen_unit_f32(n, SEED, 21)
    wd = ctx.to_device(w.numpy().astype("float64"))
    ctx.clear_catalog()
    ctx.register_column("sortbench", "k", "int", k)
    ctx.register_column("sortbench", "w", "float", wd)
    r = rb.RelBuilder(ctx).scan("sortbench", ["k", "w"]).to_gpu().unpack()
    q = r.sort([rb.col("sortbench", "k"), rb.col("sortbench", "w")], ["ASC", "DESC"]) \
         .reduce([rb.int64(1)], ["sum"])
    text = _json.dumps(q.plan())
    assert int(ctx.execute_plan(text).strip()) == n
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.execute_plan(text)
    sec = (time.perf_counter() - t0) / steps
    # the radix passes, timed with HIP events in one more run
    ctx.prof_enable(True)
    ctx.prof_reset()
    ctx.execute_plan(text)
    names = ("sort_key", "sort_digits", "sort_hist", "sort_colsum", "sort_colscan", "sort_scan",
             "sort_scatter")
    kt = kernel_table(ctx, names)
    ctx.prof_enable(False)
    ctx.clear_catalog()
    for x in (k, w, wd):
        x.free()
    out = {"rows": n, "keys": "int32 ASC, FP64 DESC", "ms": sec * 1e3, "rows_per_s": n / sec,
           "kernels_ms": {nm: round(v[0], 4) for nm, v in kt.items()},
           "launches": {nm: v[1] for nm, v in kt.items()}}
    if "sort_scatter" in kt:
        # a pass of (key, row) pairs reads and writes both: the int32 key's
        # passes move 16 B per row, the FP64 key's 24 B per row
        t, c = kt["sort_scatter"]
        p32 = 3  # the int32 key in [0, 2^20): digits 0..2 vary
        p64 = c - p32
        algo = n * (16 * p32 + 24 * p64)
        out["scatter_roofline"] = {"bound": "hbm", "passes": c, "algorithmic_bytes": algo,
                                   "achieved_gbs": algo / (t * 1e-3) / 1e9,
                                   "frac": algo / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
    return out


def star_c4(ctx, steps):
    """Star-schema plan (fact 6e8 rows x 7 int32 columns, four dimensions,
    filters 1/7, 1/5, 1/25, ~800 groups of SUM(price*(100-discount)))
    through pj_plan_execute, fused probe pipeline; the operator-at-a-time
    execution of the same plan is run once as a cross-check, and the plan
    with bloom filters on the supplier and part keys is timed beside it.
    Algorithmic bytes: the 28 B of fact columns per row."""
    import json as _json
    from proteus_amd import star
    n = 600_000_000
    ctx.clear_catalog()
    s = star.Star(ctx, n, SEED)
    text = _json.dumps(s.plan())
    ref = ctx.execute_plan(text)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        got = ctx.execute_plan(text)
    sec = (time.perf_counter() - t0) / steps
    os.environ["PROTEUS_AMD_NO_PIPELINE"] = "1"
    try:
        t1 = time.perf_counter()
        mat = ctx.execute_plan(text)
        sec_mat = time.perf_counter() - t1
    finally:
        del os.environ["PROTEUS_AMD_NO_PIPELINE"]
    # the ssb100_bloom shape: bloom filters (2^22 entries) built on the
    # filtered supplier and part keys, the fact side probed before the joins
    btext = _json.dumps(s.plan(bloom_size=1 << 22))
    bref = ctx.execute_plan(btext)
    ctx.sync()
    t2 = time.perf_counter()
    for _ in range(steps):
        bgot = ctx.execute_plan(btext)
    sec_bloom = (time.perf_counter() - t2) / steps
    rows = [ln for ln in got.splitlines() if ln]
    s.close()
    gbs = 28.0 * n / sec / 1e9
    return {"fact_rows": n, "ms": sec * 1e3, "rows_per_s": n / sec, "groups": len(rows),
            "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS, "operator_at_a_time_ms": sec_mat * 1e3,
            "bloom_ms": sec_bloom * 1e3, "identical": got == ref == mat == bref == bgot}


def batched_mix(ctx, nq, qps):
    """The served mix: nq seeded join -> group-by queries over resident
    dimensions (build 2^10..2^24 log-uniform, probe 8-64x build, 2^4..2^16
    groups).  Closed loop (all queries waiting at t=0) for queries/s of the
    batched path and of the one-launch-per-query path, results compared
    bit for bit; then Poisson arrivals at `qps` for p50/p99 latency
    (arrival -> result) of both paths.  Probe inputs are generated on the
    device inside each query's service time."""
    import numpy as np
    from proteus_amd import workload
    budget = 48 << 30
    out = {"queries": nq, "offered_qps": qps, "hbm_budget_bytes": budget}
    mix = workload.Mix(ctx, nq, seed=SEED, qps=qps)
    out["probe_rows"] = mix.probe_rows()
    # tile rows, workgroups per CU and the LDS table threshold, tuned on a
    # sample of the mix before anything is timed, for the batched path; the
    # plain path keeps the untuned configuration, which is faster for one
    # query at a time than the batch-tuned one
    out["autotuned"] = mix.autotune()
    plain_cfg = dict(tile_rows=65536, wg_per_cu=2, lds_dim=2048)
    zeros = np.zeros(nq)
    res = {}
    for mode, name in ((0, "batched"), (1, "plain")):
        m0 = ctx.mem_stats()
        done, st = ctx.serve_run(mix.queries(), zeros, mode=mode, slots=3, hbm_budget=budget,
                                 **(plain_cfg if mode else {}))
        m1 = ctx.mem_stats()
        # device allocations and cache releases during the run (each release
        # is a hipFree, which waits for the whole device)
        out[name + "_pool"] = {k: m1[k] - m0[k] for k in ("mallocs", "evictions")}
        res[mode] = mix.results()
        out[name + "_qps"] = nq / st["wall_s"]
        out[name + "_probe_rows_per_s"] = mix.probe_rows() / st["wall_s"]
        # every probe row is generated in the query's service time (12 B
        # written) and read by the probe (12 B): 24 B of HBM traffic per row
        out[name + "_hbm_gbs"] = 24.0 * mix.probe_rows() / st["wall_s"] / 1e9
        if mode == 0:
            out["batches"] = st["batches"]
    out["speedup"] = out["batched_qps"] / out["plain_qps"]
    out["identical"] = all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
                           for x, y in zip(res[0], res[1]))
    arr = np.asarray(mix.arrivals)
    for mode, name in ((0, "batched"), (1, "plain")):
        done, st = ctx.serve_run(mix.queries(), arr, mode=mode, slots=3, hbm_budget=budget,
                                 **(plain_cfg if mode else {}))
        lat = (np.asarray(done) - arr) * 1e3
        out[name + "_p50_ms"] = float(np.percentile(lat, 50))
        out[name + "_p99_ms"] = float(np.percentile(lat, 99))
        got = mix.results()
        out["identical"] = out["identical"] and all(
            np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) for x, y in zip(got, res[1]))
    mix.close()
    # The same closed loop over small queries only (builds of 2^10..2^16
    # rows): per-query launches rather than bytes bound the plain path here.
    # Same tuned configuration as above.
    small = workload.Mix(ctx, nq, seed=SEED + 1, u_lo=10, u_hi=16)
    sres = {}
    sm = {"u_range": [10, 16], "probe_rows": small.probe_rows()}
    for mode, name in ((0, "batched"), (1, "plain")):
        done, st = ctx.serve_run(small.queries(), zeros, mode=mode, slots=3, hbm_budget=budget,
                                 **(plain_cfg if mode else {}))
        sres[mode] = small.results()
        sm[name + "_qps"] = nq / st["wall_s"]
    sm["speedup"] = sm["batched_qps"] / sm["plain_qps"]
    sm["identical"] = all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
                          for x, y in zip(sres[0], sres[1]))
    small.close()
    out["small_queries"] = sm
    return out


JOIN_KERNELS = ["radix_hist", "radix_colsum", "radix_colscan", "radix_scan", "radix_items",
                "radix_scatter1_build", "radix_scatter1_probe", "radix_scatter2_build",
                "radix_scatter2_probe", "pjoin_units", "pjoin_probe", "join_acc_init"]


def timed(ctx, steps, fn, names=JOIN_KERNELS):
    """Runs fn `steps` times between two device synchronisations (every call
    returns after its results are on the host); returns the mean seconds per
    step, the per-kernel (ms, launches) recorded with HIP events on the
    streams the kernels run on, the last result and the per-step seconds."""
    ctx.prof_enable(True)
    ctx.prof_reset()
    ctx.sync()
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        res = fn()
        per.append(time.perf_counter() - a)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.prof_enable(False)
    return (t1 - t0) / steps, kernel_table(ctx, names), res, per


def untimed_kernels(ctx, steps, fn):
    """Like timed() with the per-launch events off: mean seconds per step,
    the last result and the per-step seconds."""
    ctx.sync()
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        res = fn()
        per.append(time.perf_counter() - a)
    ctx.sync()
    return (time.perf_counter() - t0) / steps, res, per


def protocol(ctx, fn, runs=5):
    """The reference's perftest protocol (benchmark-aliases.hpp:60-67,
    query-perf.cpp:32-45): one warm-up run, then `runs` timed runs; the
    median milliseconds."""
    fn()
    ctx.sync()
    ts = []
    for _ in range(runs):
        a = time.perf_counter()
        fn()
        ctx.sync()
        ts.append((time.perf_counter() - a) * 1e3)
    return {"warmup": 1, "runs": runs, "median_ms": statistics.median(ts),
            "min_ms": min(ts), "max_ms": max(ts)}


def pmc_traffic(kernel, variant):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE passes, each scaled by the reading of a
    1 GiB copy), or None when that summary does not cover the kernel."""
    path = os.path.join(ROOT, "profiles", "join_2p27x2p30_pmc_traffic.txt")
    if variant != "uniform" or not os.path.exists(path):
        return None
    for ln in open(path):
        f = dict(kv.split("=", 1) for kv in ln.split() if "=" in kv)
        if f.get("kernel") == kernel and "traffic_bytes" in f:
            return float(f["traffic_bytes"])
    return None


def cpu_info():
    """CPU model and core counts of this host."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(),
            "affinity_cores": len(os.sched_getaffinity(0))}


def cpu_threads():
    """Worker threads for the CPU baseline: this process's CPU share on the
    GPU box (OMP_NUM_THREADS, set to the share there), else every core it
    may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_protocol(fn, runs=5):
    """1 warm-up + `runs` timed runs, median seconds (the reference's perftest
    protocol, benchmark-aliases.hpp:60-67)."""
    fn()
    ts = []
    for _ in range(runs):
        a = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - a)
    return statistics.median(ts)


def cpu_baseline(variant):
    """The oracle's restatement of the reference's CPU operators, timed on
    this host's cores with the same seeded inputs as the GPU runs:
      - C2 (the headline's workload) at 1/4 scale, |R| = 2^25, |S| = 2^28, on
        `threads` workers: the morsel-parallel chained hash join
        (HashJoinChainedMorsel: one shared table, 64-bit hashInt32,
        hash_bits = 26 as RelBuilder picks it) -> COUNT, SUM;
      - C1 at full size (2^20 x 2^24) on `threads` workers and on 1;
      - C3 at 1/16 scale (2^25 rows, 2^16 uniform groups) on `threads`
        workers: per-worker chained group-by tables merged per key.
    Each: 1 warm-up + 5 runs, median; input generation excluded."""
    from oracle import pyoracle as O
    threads = cpu_threads()
    out = {"kind": "port", "cores": threads, "unit": "Gtuples/s"}
    out.update(cpu_info())
    # C2 at 1/4 scale
    nr, ns = NR >> 2, NS >> 2
    rng = int(1.33 * nr)
    rk = O.gen_perm_i32(nr, SEED)
    rv = O.gen_uniform_i32(nr, SEED, 1, 0, 1 << 31)
    if variant == "uniform":
        sk = O.gen_uniform_i32(ns, SEED, 2, 0, rng)
    else:
        sk = O.gen_zipf_fk_i32(ns, SEED, 2, nr, rng, 1.0, SEED)
    sv = O.gen_uniform_i32(ns, SEED, 3, 0, 1 << 31)
    res = {}

    def c2():
        res["c2"] = O.join_chained(rk, rv, sk, sv, 26, max_build=nr, threads=threads)[1]
    t = cpu_protocol(c2)
    out["value"] = ns / t / 1e9
    out["sample"] = ("C2 at 1/4 scale: |R| = 2^25, |S| = 2^28 (same distributions), chained hash "
                     "join (morsel-parallel shared table, hash_bits = 26) -> COUNT, SUM on %d "
                     "threads; median of 5 runs after 1 warm-up, %.2f s per run" % (threads, t))
    out["c2_quarter"] = {"ms": t * 1e3, "probe_gtuples_per_s": ns / t / 1e9,
                         "input_gtuples_per_s": (nr + ns) / t / 1e9,
                         "count": res["c2"]["count"]}
    del rk, rv, sk, sv
    # C1 at full size, all threads and one
    nr, ns = 1 << 20, 1 << 24
    rk = O.gen_perm_i32(nr, SEED)
    rv = O.gen_uniform_i32(nr, SEED, 1, 0, 1 << 31)
    sk = O.gen_uniform_i32(ns, SEED, 2, 0, nr)
    sv = O.gen_uniform_i32(ns, SEED, 3, 0, 1 << 31)
    c1 = {}
    for th in (threads, 1):
        def run(th=th):
            res["c1"] = O.join_chained(rk, rv, sk, sv, 21, max_build=int(1.1 * nr),
                                       threads=th)[1]
        t = cpu_protocol(run)
        c1["threads_%d" % th] = {"ms": t * 1e3, "probe_gtuples_per_s": ns / t / 1e9,
                                 "input_gtuples_per_s": (nr + ns) / t / 1e9}
    c1["count"] = res["c1"]["count"]
    out["c1_full"] = c1
    out["threads_1"] = c1["threads_1"]
    del rk, rv, sk, sv
    # C3 at 1/16 scale, 2^16 uniform groups
    n, g = 1 << 25, 1 << 16
    k = O.gen_uniform_i32(n, SEED, 4, 0, g)
    v = O.gen_uniform_i64(n, SEED, 5, 0, 1 << 20)
    w = O.gen_unit_f32(n, SEED, 6)

    def c3():
        res["c3"] = O.groupby_chained(k, v, w, 17, g, threads=threads)[1]
    t = cpu_protocol(c3)
    out["c3_sixteenth"] = {"rows": n, "groups": g, "ms": t * 1e3, "rows_per_s": n / t,
                           "groups_out": len(res["c3"]["keys"])}
    return out


if __name__ == "__main__":
    main()
