"""Python access to the CPU checker (TEST INFRASTRUCTURE ONLY).

Loads oracle/libs512oracle.so (the C restatement in s512_oracle.c) and, when
it has been built, oracle/_ref/libs512ref.so (the reference headers
instantiated by ref_shim.cpp).  Only tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py use this module, and only as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libs512oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libs512ref.so")

_vp, _sz, _d, _i64, _u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_long, ctypes.c_uint64


def build():
    """Compile the oracle (and the reference shim when the headers are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(ORACLE_SO):
        build()
    lib = ctypes.CDLL(ORACLE_SO)
    sigs = {
        "orc_partition_i32": ([_vp, _sz, _sz, ctypes.c_int32], _sz),
        "orc_partition_f64": ([_vp, _sz, _sz, _d], _sz),
        "orc_partition_kv": ([_vp, _vp, _sz, _sz, ctypes.c_int32], _sz),
        "orc_smallsort_i32": ([_vp, _sz], ctypes.c_int),
        "orc_smallsort_f64": ([_vp, _sz], ctypes.c_int),
        "orc_smallsort_kv": ([_vp, _vp, _sz], ctypes.c_int),
        "orc_sort_i32": ([_vp, _sz], ctypes.c_int),
        "orc_sort_f64": ([_vp, _sz], ctypes.c_int),
        "orc_sort_kv": ([_vp, _vp, _sz], ctypes.c_int),
        "orc_segsort_i32": ([_vp, _vp, _sz], ctypes.c_int),
        "orc_segsort_f64": ([_vp, _vp, _sz], ctypes.c_int),
        "orc_qsort_i32": ([_vp, _sz], None),
        "orc_qsort_f64": ([_vp, _sz], None),
        "orc_pair_multiset_kv": ([_vp, _vp, _sz, _vp], None),
        "orc_findmedian_i32": ([_vp, _i64, _i64, _vp, _vp], None),
        "orc_findmedian_f64": ([_vp, _i64, _i64, _vp, _vp], None),
        "orc_merge_i32": ([_vp, _sz, _sz], ctypes.c_int),
        "orc_merge_f64": ([_vp, _sz, _sz], ctypes.c_int),
        "orc_corank_i32": ([_vp, _sz, _vp, _sz, _sz], _sz),
        "orc_corank_f64": ([_vp, _sz, _vp, _sz, _sz], _sz),
        "orc_splitmix64": ([_u64, _u64], _u64),
        "orc_gen_i32_full": ([_vp, _sz, _u64, _u64], None),
        "orc_gen_f64_uniform": ([_vp, _sz, _u64, _u64], None),
        "orc_gen_f64_normal": ([_vp, _sz, _u64, _u64], None),
        "orc_gen_f64_distinct": ([_vp, _sz, _u64, _u64, _u64], None),
        "orc_srand48": ([ctypes.c_long], None),
        "orc_drand48": ([], _d),
        "orc_rand_vec_i32": ([_vp, _sz], None),
        "orc_rand_vec_f64": ([_vp, _sz], None),
        "orc_mlp_nparams": ([ctypes.c_int], _sz),
        "orc_mlp_forward": ([_vp, ctypes.c_int, _d], _d),
        "orc_mlp_loss_grad": ([_vp, ctypes.c_int, _vp, _vp, _sz, _vp], _d),
    }
    for name, (args, res) in sigs.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


lib = _load()


def ref_lib():
    """The compiled reference (None when it has not been built here)."""
    if not os.path.exists(REF_SO):
        return None
    r = ctypes.CDLL(REF_SO)
    for name in ("ref_sort_i32", "ref_sort_omp_partition_i32", "ref_sort_omp_merge_i32",
                 "ref_sort_omp_mergedeps_i32", "ref_sort_omp_parmerge_i32", "ref_smallsort_i32"):
        getattr(r, name).argtypes = [_vp, _sz]
    for name in ("ref_sort_f64", "ref_sort_omp_partition_f64", "ref_sort_omp_merge_f64",
                 "ref_sort_omp_mergedeps_f64", "ref_sort_omp_parmerge_f64", "ref_smallsort_f64"):
        getattr(r, name).argtypes = [_vp, _sz]
    r.ref_partition_i32.argtypes = [_vp, _sz, _sz, ctypes.c_int]
    r.ref_partition_i32.restype = _sz
    r.ref_partition_f64.argtypes = [_vp, _sz, _sz, _d]
    r.ref_partition_f64.restype = _sz
    r.ref_kv_sort.argtypes = [_vp, _vp, _sz]
    r.ref_kv_sort_omp_partition.argtypes = [_vp, _vp, _sz]
    r.ref_kv_partition.argtypes = [_vp, _vp, _sz, _sz, ctypes.c_int]
    r.ref_kv_partition.restype = _sz
    r.ref_kv_smallsort.argtypes = [_vp, _vp, _sz]
    r.ref_findmedian_i32.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]
    r.ref_findmedian_f64.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]
    for name in ("ref_segsort_loop_i32", "ref_segsort_loop_f64"):
        getattr(r, name).argtypes = [_vp, _vp, _sz, ctypes.c_int]
    return r


def cpu_has_avx512():
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return all(f" {f}" in flags for f in ("avx512f", "avx512cd", "avx512vl", "avx512bw", "avx512dq"))


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _suffix(a):
    return {np.dtype(np.int32): "i32", np.dtype(np.float64): "f64"}[a.dtype]


# ---------------------------------------------------------------- restated API
def sort(a):
    """Sort512::Sort restated (quicksort + bitonic leaves), in place."""
    rc = getattr(lib, "orc_sort_" + _suffix(a))(_p(a), a.size)
    assert rc == 0
    return a


def kv_sort(k, v):
    rc = lib.orc_sort_kv(_p(k), _p(v), k.size)
    assert rc == 0
    return k, v


def partition(a, left, right, pivot):
    """Partition512 restated: returns left + #{a[i] <= pivot, i in [left,right]}."""
    f = getattr(lib, "orc_partition_" + _suffix(a))
    return f(_p(a), left, right, pivot)


def kv_partition(k, v, left, right, pivot):
    return lib.orc_partition_kv(_p(k), _p(v), left, right, int(pivot))


def smallsort(a):
    rc = getattr(lib, "orc_smallsort_" + _suffix(a))(_p(a), a.size)
    assert rc == 0
    return a


def kv_smallsort(k, v):
    rc = lib.orc_smallsort_kv(_p(k), _p(v), k.size)
    assert rc == 0
    return k, v


def segsort(a, off):
    """Loop of Sort512::Sort over the segments a[off[i]:off[i+1]], in place."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    rc = getattr(lib, "orc_segsort_" + _suffix(a))(_p(a), _p(off), off.size - 1)
    assert rc == 0
    return a


def closed_sort(a):
    """Closed form: ascending order is unique (C qsort, never quadratic)."""
    getattr(lib, "orc_qsort_" + _suffix(a))(_p(a), a.size)
    return a


def pair_multiset(k, v):
    out = np.empty(k.size, dtype=np.uint64)
    lib.orc_pair_multiset_kv(_p(k), _p(v), k.size, _p(out))
    return out


def findmedian(a, center):
    mA, mB = ctypes.c_long(), ctypes.c_long()
    getattr(lib, "orc_findmedian_" + _suffix(a))(_p(a), center, a.size, ctypes.byref(mA), ctypes.byref(mB))
    return mA.value, mB.value


def merge(a, center):
    rc = getattr(lib, "orc_merge_" + _suffix(a))(_p(a), center, a.size)
    assert rc == 0
    return a


def corank(A, B, diag):
    return getattr(lib, "orc_corank_" + _suffix(A))(_p(A), A.size, _p(B), B.size, diag)


# ---------------------------------------------------------------- generators
def splitmix64(seed, i):
    return lib.orc_splitmix64(seed, i)


def gen_i32_full(n, seed=0, start=0):
    a = np.empty(n, dtype=np.int32)
    lib.orc_gen_i32_full(_p(a), n, seed, start)
    return a


def gen_f64(kind, n, seed, start=0, ndistinct=1000):
    a = np.empty(n, dtype=np.float64)
    if kind == "uniform":
        lib.orc_gen_f64_uniform(_p(a), n, seed, start)
    elif kind == "normal":
        lib.orc_gen_f64_normal(_p(a), n, seed, start)
    elif kind == "distinct":
        lib.orc_gen_f64_distinct(_p(a), n, seed, start, ndistinct)
    elif kind == "sorted":
        a[:] = (np.arange(start, start + n, dtype=np.float64)) / float(n)
    elif kind == "reversed":
        a[:] = (n - 1 - np.arange(start, start + n, dtype=np.float64)) / float(n)
    else:
        raise ValueError(kind)
    return a


def srand48(s):
    lib.orc_srand48(s)


def drand48():
    return lib.orc_drand48()


def rand_vec(n, dtype):
    """createRandVec (sort512test.cpp:77-82) on the current drand48 state."""
    a = np.empty(n, dtype=dtype)
    getattr(lib, "orc_rand_vec_" + _suffix(a))(_p(a), n)
    return a


# ---------------------------------------------------------------- MLP
def mlp_nparams(H):
    return lib.orc_mlp_nparams(H)


def mlp_forward(p, H, x):
    p = np.ascontiguousarray(p, dtype=np.float64)
    return lib.orc_mlp_forward(_p(p), H, float(x))


def mlp_loss_grad(p, H, x, t):
    p = np.ascontiguousarray(p, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    g = np.empty_like(p)
    loss = lib.orc_mlp_loss_grad(_p(p), H, _p(x), _p(t), x.size, _p(g))
    return loss, g
