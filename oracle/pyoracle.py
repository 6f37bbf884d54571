"""ctypes binding of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results.  The product package never
imports this module.  Each wrapped function restates the reference file:line
cited in oracle.c.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OR_OK, OR_E_CAPACITY, OR_E_ARG, OR_E_OOM = 0, 1, 2, 3


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _setup(_lib)
    return _lib


class JoinOut(C.Structure):
    _fields_ = [(n, C.c_int64) for n in
                ("count", "sum_b", "sum_p", "min_b", "max_b", "min_p", "max_p")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Groups(C.Structure):
    _fields_ = [("keys", C.c_void_p), ("cnt", C.c_void_p), ("sum", C.c_void_p),
                ("min", C.c_void_p), ("max", C.c_void_p), ("fsum", C.c_void_p)]


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _setup(L):
    u64, i32, i64, vp = C.c_uint64, C.c_int32, C.c_int64, C.c_void_p
    L.or_hash_i32.restype = u64
    L.or_hash_i32.argtypes = [i32]
    L.or_bloom_build.restype = None
    L.or_bloom_build.argtypes = [vp, u64, vp, u64]
    L.or_bloom_probe.restype = u64
    L.or_bloom_probe.argtypes = [vp, u64, vp, u64, vp]
    L.or_phjc_bucket.restype = C.c_uint32
    L.or_phjc_bucket.argtypes = [i32, C.c_int, C.c_int]
    L.or_rearrange_bucket.restype = C.c_uint32
    L.or_rearrange_bucket.argtypes = [i32, C.c_uint32]
    L.or_mix64.restype = u64
    L.or_mix64.argtypes = [u64]
    L.or_rng.restype = u64
    L.or_rng.argtypes = [u64, u64, u64]
    L.or_uniform.restype = u64
    L.or_uniform.argtypes = [u64, u64, u64, u64]
    L.or_perm.restype = u64
    L.or_perm.argtypes = [u64, u64, u64]
    L.or_zipf.restype = u64
    L.or_zipf.argtypes = [u64, u64, u64, u64, C.c_double]
    L.or_gen_perm_i32.argtypes = [vp, u64, u64, C.c_int]
    L.or_gen_uniform_i32.argtypes = [vp, u64, u64, u64, i64, u64, C.c_int]
    L.or_gen_uniform_i64.argtypes = [vp, u64, u64, u64, i64, u64, C.c_int]
    L.or_gen_unit_f32.argtypes = [vp, u64, u64, u64, C.c_int]
    L.or_gen_zipf_fk_i32.argtypes = [vp, u64, u64, u64, u64, u64, C.c_double, u64, C.c_int]
    L.or_gen_zipf_i32.argtypes = [vp, u64, u64, u64, u64, C.c_double, u64, C.c_int]
    L.or_join_chained.argtypes = [vp, vp, u64, vp, vp, u64, C.c_int, u64, C.POINTER(JoinOut)]
    L.or_join_chained_mt.argtypes = [vp, vp, u64, vp, vp, u64, C.c_int, u64, C.c_int,
                                     C.POINTER(JoinOut)]
    L.or_join_dense.argtypes = [vp, vp, u64, vp, vp, u64, C.c_int, C.POINTER(JoinOut)]
    L.or_join_pairs.argtypes = [vp, u64, vp, u64, C.c_int, u64, vp, vp, u64,
                                C.POINTER(u64)]
    L.or_groupby_chained.argtypes = [vp, vp, vp, u64, C.c_int, u64, C.POINTER(Groups),
                                     C.POINTER(u64)]
    L.or_groupby_chained_mt.argtypes = [vp, vp, vp, u64, C.c_int, u64, C.c_int,
                                        C.POINTER(Groups), C.POINTER(u64)]
    L.or_groupby_dense.argtypes = [vp, vp, vp, u64, u64, C.c_int, C.POINTER(Groups)]
    L.or_reduce_i64.argtypes = [vp, u64, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
    L.or_histogram.argtypes = [vp, u64, C.c_int, C.c_int, vp]


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


# ---- hashing ---------------------------------------------------------------
def hash_i32(k):
    return lib().or_hash_i32(int(np.int32(k)))


def phjc_bucket(k, log_parts, hash_bits):
    return lib().or_phjc_bucket(int(np.int32(k)), log_parts, hash_bits)


def rearrange_bucket(k, buckets):
    return lib().or_rearrange_bucket(int(np.int32(k)), buckets)


# ---- generators --------------------------------------------------------------
def rng(seed, stream, i):
    return lib().or_rng(seed, stream, i)


def uniform(seed, stream, i, n):
    return lib().or_uniform(seed, stream, i, n)


def perm(seed, i, n):
    return lib().or_perm(seed, i, n)


def zipf(seed, stream, i, n, theta):
    return lib().or_zipf(seed, stream, i, n, theta)


def gen_perm_i32(n, seed):
    out = np.empty(n, np.int32)
    lib().or_gen_perm_i32(_p(out), n, seed, _threads())
    return out


def gen_uniform_i32(n, seed, stream, lo, rng_):
    out = np.empty(n, np.int32)
    lib().or_gen_uniform_i32(_p(out), n, seed, stream, lo, rng_, _threads())
    return out


def gen_uniform_i64(n, seed, stream, lo, rng_):
    out = np.empty(n, np.int64)
    lib().or_gen_uniform_i64(_p(out), n, seed, stream, lo, rng_, _threads())
    return out


def gen_unit_f32(n, seed, stream):
    out = np.empty(n, np.float32)
    lib().or_gen_unit_f32(_p(out), n, seed, stream, _threads())
    return out


def gen_zipf_fk_i32(n, seed, stream, nbuild, rng_, theta, perm_seed):
    out = np.empty(n, np.int32)
    lib().or_gen_zipf_fk_i32(_p(out), n, seed, stream, nbuild, rng_, theta, perm_seed,
                             _threads())
    return out


def gen_zipf_i32(n, seed, stream, g, theta, perm_seed):
    out = np.empty(n, np.int32)
    lib().or_gen_zipf_i32(_p(out), n, seed, stream, g, theta, perm_seed, _threads())
    return out


# ---- join ------------------------------------------------------------------
def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def join_chained(rk, rv, sk, sv, hash_bits, max_build=None, threads=1):
    rk, rv, sk, sv = _i32(rk), _i32(rv), _i32(sk), _i32(sv)
    for k, v in ((rk, rv), (sk, sv)):
        if v is not None and len(v) != len(k):
            raise ValueError("value column has %d rows, key column %d" % (len(v), len(k)))
    if max_build is None:
        max_build = len(rk)
    out = JoinOut()
    if threads == 1:
        rc = lib().or_join_chained(_p(rk), _p(rv), len(rk), _p(sk), _p(sv), len(sk),
                                   hash_bits, max_build, C.byref(out))
    else:
        rc = lib().or_join_chained_mt(_p(rk), _p(rv), len(rk), _p(sk), _p(sv), len(sk),
                                      hash_bits, max_build, threads, C.byref(out))
    return rc, out.as_dict()


def join_dense(rk, rv, sk, sv, threads=None):
    rk, rv, sk, sv = _i32(rk), _i32(rv), _i32(sk), _i32(sv)
    out = JoinOut()
    rc = lib().or_join_dense(_p(rk), _p(rv), len(rk), _p(sk), _p(sv), len(sk),
                             threads or _threads(), C.byref(out))
    return rc, out.as_dict()


def join_pairs(rk, sk, hash_bits, max_build=None, cap=None):
    rk, sk = _i32(rk), _i32(sk)
    if max_build is None:
        max_build = len(rk)
    if cap is None:
        cap = len(rk) * len(sk) if len(rk) * len(sk) < (1 << 22) else (1 << 22)
    ob = np.empty(max(cap, 1), np.uint32)
    op = np.empty(max(cap, 1), np.uint32)
    n = C.c_uint64(0)
    rc = lib().or_join_pairs(_p(rk), len(rk), _p(sk), len(sk), hash_bits, max_build,
                             _p(ob), _p(op), cap, C.byref(n))
    n = min(n.value, cap)
    return rc, ob[:n], op[:n]


# ---- group-by --------------------------------------------------------------
def _alloc_groups(g):
    arrs = dict(keys=np.zeros(g, np.int32), cnt=np.zeros(g, np.int64),
                sum=np.zeros(g, np.int64), min=np.zeros(g, np.int64),
                max=np.zeros(g, np.int64), fsum=np.zeros(g, np.float64))
    st = Groups(*[arrs[n].ctypes.data for n in ("keys", "cnt", "sum", "min", "max", "fsum")])
    return arrs, st


def groupby_chained(k, v, w, hash_bits, max_groups, threads=1):
    k = _i32(k)
    v = None if v is None else np.ascontiguousarray(v, np.int64)
    w = None if w is None else np.ascontiguousarray(w, np.float32)
    arrs, st = _alloc_groups(max(max_groups, 1))
    ng = C.c_uint64(0)
    if threads == 1:
        rc = lib().or_groupby_chained(_p(k), _p(v), _p(w), len(k), hash_bits, max_groups,
                                      C.byref(st), C.byref(ng))
    else:
        rc = lib().or_groupby_chained_mt(_p(k), _p(v), _p(w), len(k), hash_bits, max_groups,
                                         threads, C.byref(st), C.byref(ng))
    n = ng.value
    return rc, {name: a[:n] for name, a in arrs.items()}


def groupby_dense(k, v, w, g):
    k = _i32(k)
    v = None if v is None else np.ascontiguousarray(v, np.int64)
    w = None if w is None else np.ascontiguousarray(w, np.float32)
    arrs, st = _alloc_groups(g)
    rc = lib().or_groupby_dense(_p(k), _p(v), _p(w), len(k), g, _threads(), C.byref(st))
    return rc, arrs


def reduce_i64(v):
    v = np.ascontiguousarray(v, np.int64)
    s, mn, mx = C.c_int64(), C.c_int64(), C.c_int64()
    lib().or_reduce_i64(_p(v), len(v), C.byref(s), C.byref(mn), C.byref(mx))
    return s.value, mn.value, mx.value


def histogram(k, shift, bits):
    k = _i32(k)
    h = np.zeros(1 << bits, np.uint64)
    lib().or_histogram(_p(k), len(k), shift, bits, _p(h))
    return h


def bloom_build(keys, size):
    """or_bloom_build: the reference's one-byte-per-entry filter
    (bloom-filter-build.cpp:29-58) of `size` entries over the int32 keys."""
    keys = _i32(keys)
    f = np.zeros(int(size), np.uint8)
    lib().or_bloom_build(f.ctypes.data, int(size), keys.ctypes.data, len(keys))
    return f


def bloom_probe(filt, keys):
    """or_bloom_probe: ascending row ids of the keys whose flag is set
    (bloom-filter-probe.cpp:43-52)."""
    keys = _i32(keys)
    out = np.empty(max(len(keys), 1), np.uint32)
    m = lib().or_bloom_probe(filt.ctypes.data, len(filt), keys.ctypes.data, len(keys),
                             out.ctypes.data)
    return out[:m].copy()
