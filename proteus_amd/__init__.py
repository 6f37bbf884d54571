"""proteus_amd -- MI355X (gfx950) join / aggregation hot path of Proteus.

A thin ctypes binding of the C ABI in include/proteus_amd.h, implemented by
libproteus_amd.so (hand-written HIP kernels + the C++ plan engine).  There is
no CPU fallback: importing this package raises if the library is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libproteus_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "proteus_amd: %s is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `make -C proteus_amd`" % LIB_PATH)

_lib = C.CDLL(LIB_PATH)

PJ_OK = 0
PJ_E_CAPACITY = 1
PJ_E_OOM = 2
PJ_E_UNSUPPORTED = 3
PJ_E_HIP = 4
PJ_E_ARG = 5
PJ_E_PARSE = 6
PJ_E_CANCELLED = 7

_STATUS = {1: "CAPACITY", 2: "OOM", 3: "UNSUPPORTED", 4: "HIP", 5: "ARG", 6: "PARSE",
           7: "CANCELLED"}


class PjError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("PJ_E_%s: %s" % (_STATUS.get(code, str(code)), msg))
        self.code = code


class JoinAgg(C.Structure):
    _fields_ = [(n, C.c_int64) for n in
                ("count", "sum_b", "sum_p", "min_b", "max_b", "min_p", "max_p")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class RadixCfg(C.Structure):
    _fields_ = [("bits1", C.c_int), ("bits2", C.c_int), ("lds_log", C.c_int),
                ("probe_unit", C.c_uint32), ("grid_mul", C.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Query(C.Structure):
    """pj_query: one join -> group-by over a resident dimension."""
    _fields_ = [("dim", C.c_void_p), ("d_fk", C.c_void_p), ("d_val", C.c_void_p),
                ("d_grp", C.c_void_p), ("n", C.c_uint64), ("groups", C.c_uint32),
                ("flags", C.c_uint32), ("gen_seed", C.c_uint64), ("fk_range", C.c_uint64),
                ("d_cnt", C.c_void_p), ("d_sum", C.c_void_p)]


class BatchCfg(C.Structure):
    _fields_ = [("tile_rows", C.c_uint32), ("wg_per_cu", C.c_int), ("lds_dim", C.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ServeCfg(C.Structure):
    _fields_ = [("mode", C.c_int), ("slots", C.c_int), ("max_batch", C.c_uint32),
                ("hbm_budget", C.c_uint64), ("heavy_rows", C.c_uint64), ("batch", BatchCfg),
                ("max_batch_rows", C.c_uint64), ("cancel", C.c_void_p)]


class ServeStats(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("max_in_flight_bytes", C.c_uint64),
                ("wall_s", C.c_double)]


vp, u64, i64, i32, u32 = C.c_void_p, C.c_uint64, C.c_int64, C.c_int, C.c_uint32
_SIGS = {
    "pj_last_error": ([], C.c_char_p),
    "pj_ctx_create": ([i32, u64, C.POINTER(vp)], i32),
    "pj_ctx_destroy": ([vp], i32),
    "pj_ctx_sync": ([vp], i32),
    "pj_ctx_device_info": ([vp, C.POINTER(C.c_int), C.POINTER(u64), C.POINTER(u64)], i32),
    "pj_malloc": ([vp, u64, C.POINTER(vp)], i32),
    "pj_free": ([vp, vp], i32),
    "pj_malloc_pinned": ([vp, u64, C.POINTER(vp)], i32),
    "pj_free_pinned": ([vp, vp], i32),
    "pj_memcpy_h2d": ([vp, vp, vp, u64], i32),
    "pj_memcpy_d2h": ([vp, vp, vp, u64], i32),
    "pj_memset": ([vp, vp, i32, u64], i32),
    "pj_mem_in_use": ([vp, C.POINTER(u64)], i32),
    "pj_mem_stats": ([vp, C.POINTER(u64)], i32),
    "pj_copy": ([vp, vp, vp, u64], i32),
    "pj_prof_enable": ([vp, i32], i32),
    "pj_prof_reset": ([vp], i32),
    "pj_prof_get": ([vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(u64)], i32),
    "pj_gen_perm_i32": ([vp, vp, u64, u64], i32),
    "pj_gen_uniform_i32": ([vp, vp, u64, u64, u64, i64, u64], i32),
    "pj_gen_uniform_i64": ([vp, vp, u64, u64, u64, i64, u64], i32),
    "pj_gen_unit_f32": ([vp, vp, u64, u64, u64], i32),
    "pj_gen_zipf_fk_i32": ([vp, vp, u64, u64, u64, u64, u64, C.c_double, u64], i32),
    "pj_gen_zipf_i32": ([vp, vp, u64, u64, u64, u64, C.c_double, u64], i32),
    "pj_join_chained_reduce": ([vp, vp, vp, u64, vp, vp, u64, i32, u64, C.POINTER(JoinAgg)],
                               i32),
    "pj_join_chained_reduce_noalloc": ([vp, vp, vp, u64, vp, vp, u64, i32, u64,
                                        C.POINTER(JoinAgg)], i32),
    "pj_join_chained_pairs": ([vp, vp, u64, vp, u64, i32, u64, vp, vp, u64, C.POINTER(u64)],
                              i32),
    "pj_join_partitioned_reduce": ([vp, vp, vp, u64, vp, vp, u64, u64, u64, C.POINTER(RadixCfg),
                                    C.POINTER(JoinAgg)], i32),
    "pj_join_out_of_gpu_reduce": ([vp, vp, vp, u64, vp, vp, u64, C.c_uint32, u64,
                                   C.POINTER(JoinAgg)], i32),
    "pj_join_partitioned_config": ([vp, u64, u64, C.POINTER(RadixCfg)], i32),
    "pj_autotune_enable": ([vp, i32], i32),
    "pj_groupby_i32": ([vp, vp, vp, vp, u64, i32, u64, vp, vp, vp, vp, vp, vp, C.POINTER(u64)],
                       i32),
    "pj_reduce_i64": ([vp, vp, u64, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)], i32),
    "pj_hash_rearrange_i32": ([vp, vp, vp, u64, u32, vp, vp, C.POINTER(u64)], i32),
    "pj_dim_create": ([vp, vp, vp, u64, C.POINTER(vp)], i32),
    "pj_dim_destroy": ([vp, vp], i32),
    "pj_dim_info": ([vp, vp, C.POINTER(u64), C.POINTER(u64)], i32),
    "pj_bloom_build_i32": ([vp, u64, u64, vp, u64], i32),
    "pj_bloom_probe_i32": ([vp, u64, vp, u64, vp, C.POINTER(u64)], i32),
    "pj_bloom_repack": ([vp, u64, vp, u64, i32, C.POINTER(vp), C.POINTER(i32), C.POINTER(vp),
                         C.POINTER(u64)], i32),
    "pj_bloom_release": ([vp, u64], i32),
    "pj_query_run": ([vp, C.POINTER(Query), C.POINTER(BatchCfg)], i32),
    "pj_batch_slots": ([vp, i32], i32),
    "pj_batch_launch": ([vp, i32, C.POINTER(Query), u32, C.POINTER(BatchCfg)], i32),
    "pj_batch_poll": ([vp, i32, C.POINTER(C.c_int)], i32),
    "pj_batch_wait": ([vp, i32], i32),
    "pj_batch_autotune": ([vp, C.POINTER(Query), u32, C.POINTER(BatchCfg)], i32),
    "pj_serve_run": ([vp, C.POINTER(Query), C.POINTER(C.c_double), u32, C.POINTER(ServeCfg),
                      C.POINTER(C.c_double), C.POINTER(ServeStats)], i32),
    "pj_sort_records": ([vp, vp, u64, C.c_char_p], i32),
    "pj_stage_begin": ([vp, C.c_char_p, vp, u64, u64, vp, C.POINTER(u64)], i32),
    "pj_stage_wait": ([vp, u64], i32),
    "pj_stage_file": ([vp, C.c_char_p, u64, u64, vp], i32),
    "pj_register_column": ([vp, C.c_char_p, C.c_char_p, C.c_char_p, vp, u64, i32], i32),
    "pj_clear_catalog": ([vp], i32),
    "pj_plan_execute": ([vp, C.c_char_p, C.c_char_p, C.c_char_p, u64], i32),
}
for _name, (_args, _res) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.argtypes = _args
    _f.restype = _res


def lib():
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def _check(rc):
    if rc != PJ_OK:
        raise PjError(rc, _lib.pj_last_error().decode(errors="replace"))


_DT = {np.dtype(np.int32): "int32", np.dtype(np.int64): "int64",
       np.dtype(np.float32): "float32", np.dtype(np.float64): "float64",
       np.dtype(np.uint32): "uint32", np.dtype(np.uint8): "uint8"}


class DeviceArray:
    """A column in HBM owned by a Context (freed with free() or the context)."""

    def __init__(self, ctx, ptr, n, dtype):
        self.ctx, self.ptr, self.n, self.dtype = ctx, ptr, n, np.dtype(dtype)

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        if self.n:
            _check(_lib.pj_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            _check(_lib.pj_free(self.ctx.h, self.ptr))
            self.ptr = None

    def __len__(self):
        return self.n


def _same_len(*pairs):
    """Key and value columns of one side must have the same row count."""
    for k, v in pairs:
        if v is not None and len(v) != len(k):
            raise PjError(PJ_E_ARG, "value column has %d rows, key column %d" % (len(v), len(k)))


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    return int(a)


class Context:
    """One device context: stream, HBM pool (optional byte budget), profiler."""

    def __init__(self, device=0, hbm_budget=0):
        h = vp()
        _check(_lib.pj_ctx_create(device, hbm_budget, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _check(_lib.pj_ctx_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_info(self):
        n, hbm, lds = C.c_int(), u64(), u64()
        _check(_lib.pj_ctx_device_info(self.h, C.byref(n), C.byref(hbm), C.byref(lds)))
        return {"num_cus": n.value, "hbm_bytes": hbm.value, "lds_per_cu": lds.value}

    def sync(self):
        _check(_lib.pj_ctx_sync(self.h))

    # ---- memory ----------------------------------------------------------
    def empty(self, n, dtype):
        dtype = np.dtype(dtype)
        p = vp()
        _check(_lib.pj_malloc(self.h, max(n * dtype.itemsize, 1), C.byref(p)))
        return DeviceArray(self, p.value, n, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = self.empty(len(a), a.dtype)
        if len(a):
            _check(_lib.pj_memcpy_h2d(self.h, d.ptr, a.ctypes.data, a.nbytes))
        return d

    def sort_records(self, rows, n, fields):
        """Sort n device records (fields 'i' int32 / 'l' int64, C-struct
        layout) in place, ascending, first field most significant, stable."""
        _check(_lib.pj_sort_records(self.h, _ptr(rows), n, fields.encode()))

    def copy_h2d(self, dst, a):
        """One hipMemcpy of host array a into device memory dst."""
        a = np.ascontiguousarray(a)
        _check(_lib.pj_memcpy_h2d(self.h, _ptr(dst), a.ctypes.data, a.nbytes))

    # ---- host -> HBM staging through the pinned ring ------------------------
    def stage_file(self, path, dst, nbytes, offset=0):
        """Copy nbytes of file `path` from `offset` into device memory dst."""
        _check(_lib.pj_stage_file(self.h, os.fsencode(path), offset, nbytes, _ptr(dst)))

    def stage_begin(self, dst, nbytes, path=None, host=None, offset=0):
        """Queue a copy from a file or a host numpy array; returns the ticket.
        A host array must stay alive until stage_wait returns."""
        t = u64()
        hp = None if host is None else host.ctypes.data
        _check(_lib.pj_stage_begin(self.h, None if path is None else os.fsencode(path), hp,
                                   offset, nbytes, _ptr(dst), C.byref(t)))
        return t.value

    def stage_wait(self, ticket):
        _check(_lib.pj_stage_wait(self.h, ticket))

    def copy(self, dst, src, nbytes):
        _check(_lib.pj_copy(self.h, _ptr(dst), _ptr(src), nbytes))

    def mem_in_use(self):
        b = u64()
        _check(_lib.pj_mem_in_use(self.h, C.byref(b)))
        return b.value

    def mem_stats(self):
        """Pool counters: bytes in use and cached, hipMalloc calls, cached
        blocks released."""
        b = (u64 * 4)()
        _check(_lib.pj_mem_stats(self.h, b))
        return {"in_use": b[0], "cached": b[1], "mallocs": b[2], "evictions": b[3]}

    # ---- profiling -------------------------------------------------------
    def prof_enable(self, on=True):
        _check(_lib.pj_prof_enable(self.h, 1 if on else 0))

    def prof_reset(self):
        _check(_lib.pj_prof_reset(self.h))

    def prof_get(self, prefix):
        t, n = C.c_double(), u64()
        _check(_lib.pj_prof_get(self.h, prefix.encode(), C.byref(t), C.byref(n)))
        return t.value, n.value

    # ---- generators ------------------------------------------------------
    def gen_perm_i32(self, n, seed):
        d = self.empty(n, np.int32)
        _check(_lib.pj_gen_perm_i32(self.h, d.ptr, n, seed))
        return d

    def gen_uniform_i32(self, n, seed, stream, lo, rng):
        d = self.empty(n, np.int32)
        _check(_lib.pj_gen_uniform_i32(self.h, d.ptr, n, seed, stream, lo, rng))
        return d

    def gen_uniform_i64(self, n, seed, stream, lo, rng):
        d = self.empty(n, np.int64)
        _check(_lib.pj_gen_uniform_i64(self.h, d.ptr, n, seed, stream, lo, rng))
        return d

    def gen_unit_f32(self, n, seed, stream):
        d = self.empty(n, np.float32)
        _check(_lib.pj_gen_unit_f32(self.h, d.ptr, n, seed, stream))
        return d

    def gen_zipf_fk_i32(self, n, seed, stream, nbuild, rng, theta, perm_seed):
        d = self.empty(n, np.int32)
        _check(_lib.pj_gen_zipf_fk_i32(self.h, d.ptr, n, seed, stream, nbuild, rng, theta,
                                       perm_seed))
        return d

    def gen_zipf_i32(self, n, seed, stream, g, theta, perm_seed):
        d = self.empty(n, np.int32)
        _check(_lib.pj_gen_zipf_i32(self.h, d.ptr, n, seed, stream, g, theta, perm_seed))
        return d

    # ---- operators -------------------------------------------------------
    def join_chained_reduce(self, bk, bv, pk, pv, hash_bits, max_build=None):
        _same_len((bk, bv), (pk, pv))
        out = JoinAgg()
        nb, np_ = len(bk), len(pk)
        _check(_lib.pj_join_chained_reduce(self.h, _ptr(bk), _ptr(bv), nb, _ptr(pk), _ptr(pv),
                                           np_, hash_bits,
                                           nb if max_build is None else max_build,
                                           C.byref(out)))
        return out.as_dict()

    def join_chained_reduce_noalloc(self, bk, bv, pk, pv, hash_bits, max_build=None):
        _same_len((bk, bv), (pk, pv))
        out = JoinAgg()
        nb, np_ = len(bk), len(pk)
        _check(_lib.pj_join_chained_reduce_noalloc(self.h, _ptr(bk), _ptr(bv), nb, _ptr(pk),
                                                   _ptr(pv), np_, hash_bits,
                                                   nb if max_build is None else max_build,
                                                   C.byref(out)))
        return out.as_dict()

    def join_chained_pairs(self, bk, pk, hash_bits, cap, max_build=None):
        ob = self.empty(max(cap, 1), np.uint32)
        op = self.empty(max(cap, 1), np.uint32)
        n = u64()
        rc = _lib.pj_join_chained_pairs(self.h, _ptr(bk), len(bk), _ptr(pk), len(pk), hash_bits,
                                        len(bk) if max_build is None else max_build, ob.ptr,
                                        op.ptr, cap, C.byref(n))
        if rc not in (PJ_OK, PJ_E_CAPACITY):
            ob.free()
            op.free()
            _check(rc)
        return rc, n.value, ob, op

    def join_partitioned_reduce(self, bk, bv, pk, pv, max_build=None, max_probe=None, cfg=None):
        _same_len((bk, bv), (pk, pv))
        out = JoinAgg()
        c = None
        if cfg is not None:
            c = RadixCfg(**cfg) if isinstance(cfg, dict) else cfg
        nb, np_ = len(bk), len(pk)
        _check(_lib.pj_join_partitioned_reduce(
            self.h, _ptr(bk), _ptr(bv), nb, _ptr(pk), _ptr(pv), np_,
            nb if max_build is None else max_build, np_ if max_probe is None else max_probe,
            C.byref(c) if c is not None else None, C.byref(out)))
        return out.as_dict()

    def join_out_of_gpu_reduce(self, bk, bv, pk, pv, num_buckets=0, hbm_budget=0):
        """out-of-gpu-join -> reduce over host (numpy int32) columns."""
        _same_len((bk, bv), (pk, pv))
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.int32)
                for a in (bk, bv, pk, pv)]
        hp = [None if a is None else a.ctypes.data for a in arrs]
        out = JoinAgg()
        _check(_lib.pj_join_out_of_gpu_reduce(self.h, hp[0], hp[1], len(arrs[0]), hp[2], hp[3],
                                              len(arrs[2]), num_buckets, hbm_budget,
                                              C.byref(out)))
        return out.as_dict()

    def join_partitioned_config(self, nb, np_):
        c = RadixCfg()
        _check(_lib.pj_join_partitioned_config(self.h, nb, np_, C.byref(c)))
        return c.as_dict()

    def autotune_enable(self, on=True):
        _check(_lib.pj_autotune_enable(self.h, 1 if on else 0))

    def groupby_i32(self, k, v, w, max_groups, hash_bits=10):
        ok = self.empty(max(max_groups, 1), np.int32)
        cols = [self.empty(max(max_groups, 1), t) for t in
                (np.int64, np.int64, np.int64, np.int64, np.float64)]
        ng = u64()
        try:
            _check(_lib.pj_groupby_i32(self.h, _ptr(k), _ptr(v), _ptr(w), len(k), hash_bits,
                                       max_groups, ok.ptr, *[c.ptr for c in cols],
                                       C.byref(ng)))
            n = ng.value
            res = {"keys": ok.numpy()[:n]}
            for name, c in zip(("cnt", "sum", "min", "max", "fsum"), cols):
                res[name] = c.numpy()[:n]
            return res
        finally:
            ok.free()
            for c in cols:
                c.free()

    def reduce_i64(self, v):
        s, mn, mx = i64(), i64(), i64()
        _check(_lib.pj_reduce_i64(self.h, _ptr(v), len(v), C.byref(s), C.byref(mn),
                                  C.byref(mx)))
        return s.value, mn.value, mx.value

    def hash_rearrange_i32(self, k, v, buckets):
        ok = self.empty(len(k), np.int32)
        ov = self.empty(len(k), np.int32) if v is not None else None
        offs = (u64 * (buckets + 1))()
        _check(_lib.pj_hash_rearrange_i32(self.h, _ptr(k), _ptr(v), len(k), buckets, ok.ptr,
                                          None if ov is None else ov.ptr, offs))
        return ok, ov, np.array(list(offs), np.uint64)

    # ---- plan API --------------------------------------------------------
    def register_column(self, rel, attr, type_, data):
        if isinstance(data, DeviceArray):
            _check(_lib.pj_register_column(self.h, rel.encode(), attr.encode(), type_.encode(),
                                           data.ptr, data.n, 1))
        else:
            a = np.ascontiguousarray(data)
            self._keep = getattr(self, "_keep", [])
            self._keep.append(a)
            _check(_lib.pj_register_column(self.h, rel.encode(), attr.encode(), type_.encode(),
                                           a.ctypes.data, len(a), 0))

    # ---- bloom filters (bloomfilter_build / _probe / _repack) ---------------
    def bloom_build(self, bloom_id, filter_size, keys):
        """A fresh filter of filter_size entries (a power of two) with the
        entries of the int32 device column `keys` set; replaces the filter of
        the same id."""
        _check(_lib.pj_bloom_build_i32(self.h, bloom_id, filter_size, _ptr(keys), len(keys)))

    def bloom_probe(self, bloom_id, keys):
        """Ascending row ids (numpy uint32) of the keys whose entry is set."""
        n = len(keys)
        idx = self.empty(max(n, 1), np.uint32)
        m = u64()
        try:
            _check(_lib.pj_bloom_probe_i32(self.h, bloom_id, _ptr(keys), n, idx.ptr, C.byref(m)))
            return idx.numpy()[:m.value].copy()
        finally:
            idx.free()

    def bloom_repack(self, bloom_id, keys, cols):
        """The rows of the device columns `cols` whose key's entry is set, in
        row order (numpy arrays)."""
        n = len(keys)
        outs = [self.empty(max(n, 1), c.dtype) for c in cols]
        m = u64()
        src = (vp * max(len(cols), 1))(*[c.ptr for c in cols])
        dst = (vp * max(len(cols), 1))(*[o.ptr for o in outs])
        es = (i32 * max(len(cols), 1))(*[c.dtype.itemsize for c in cols])
        try:
            _check(_lib.pj_bloom_repack(self.h, bloom_id, _ptr(keys), n, len(cols), src, es, dst,
                                        C.byref(m)))
            return [o.numpy()[:m.value].copy() for o in outs]
        finally:
            for o in outs:
                o.free()

    def bloom_release(self, bloom_id):
        _check(_lib.pj_bloom_release(self.h, bloom_id))

    # ---- resident dimensions, batched queries, serving --------------------
    def dim_create(self, keys, vals=None):
        h = vp()
        _check(_lib.pj_dim_create(self.h, _ptr(keys), _ptr(vals), len(keys), C.byref(h)))
        return h.value

    def dim_destroy(self, dim):
        _check(_lib.pj_dim_destroy(self.h, dim))

    def dim_info(self, dim):
        rows, b = u64(), u64()
        _check(_lib.pj_dim_info(self.h, dim, C.byref(rows), C.byref(b)))
        return rows.value, b.value

    def query_run(self, q, cfg=None):
        _check(_lib.pj_query_run(self.h, C.byref(q), C.byref(cfg) if cfg else None))

    def batch_slots(self, n):
        _check(_lib.pj_batch_slots(self.h, n))

    def batch_launch(self, slot, queries, cfg=None):
        arr = (Query * len(queries))(*queries)
        _check(_lib.pj_batch_launch(self.h, slot, arr, len(queries), C.byref(cfg) if cfg else None))

    def batch_wait(self, slot):
        _check(_lib.pj_batch_wait(self.h, slot))

    def batch_autotune(self, queries):
        """Tunes tile rows, workgroups per CU and the LDS table threshold on
        these queries (outputs overwritten; probe columns allocated) and makes
        the result the context's default."""
        arr = (Query * len(queries))(*queries)
        best = BatchCfg()
        _check(_lib.pj_batch_autotune(self.h, arr, len(queries), C.byref(best)))
        return best.as_dict()

    def serve_run(self, queries, arrivals, mode=0, slots=3, max_batch=4096, hbm_budget=0,
                  tile_rows=0, wg_per_cu=0, heavy_rows=0, lds_dim=0, max_batch_rows=0,
                  cancel=None):
        """Runs the serving loop; returns (completion seconds per query, -1 for a
        query not run, and stats).  cancel: a ctypes.c_int32 the caller may set
        from another thread (the call then raises PJ_E_CANCELLED after the
        batches on the device finish)."""
        n = len(queries)
        arr = (Query * n)(*queries)
        arrv = (C.c_double * n)(*arrivals)
        done = (C.c_double * n)()
        cfg = ServeCfg(mode, slots, max_batch, hbm_budget, heavy_rows,
                       BatchCfg(tile_rows, wg_per_cu, lds_dim), max_batch_rows,
                       C.addressof(cancel) if cancel is not None else None)
        st = ServeStats()
        self.last_done, self.last_queries = done, arr
        _check(_lib.pj_serve_run(self.h, arr, arrv, n, C.byref(cfg), done, C.byref(st)))
        return list(done), {"batches": st.batches, "max_in_flight_bytes": st.max_in_flight_bytes,
                            "wall_s": st.wall_s}

    def clear_catalog(self):
        _check(_lib.pj_clear_catalog(self.h))
        self._keep = []

    def execute_plan(self, plan_json, base_dir=".", out_cap=1 << 20):
        """Runs a JSON physical plan; returns what its print operator wrote."""
        if not isinstance(plan_json, str):
            import json
            plan_json = json.dumps(plan_json)
        while True:
            buf = C.create_string_buffer(out_cap)
            rc = _lib.pj_plan_execute(self.h, plan_json.encode(), base_dir.encode(), buf, out_cap)
            if rc == PJ_E_CAPACITY:
                msg = _lib.pj_last_error().decode(errors="replace")
                if msg.startswith("plan output needs "):
                    out_cap = int(msg.split()[3]) + 1
                    continue
            _check(rc)
            return buf.value.decode()
