"""Writes the golden fixtures used by tests/test_oracle.py.

Sources (read-only, never copied here):
  - the known-answer vectors written out in the reference's tests
    (sort512test.cpp:169-176, 195-202, 224-231, 290-315, 334-375), transcribed
    as data into kat_reference_tests.json by hand;
  - outputs of the reference itself, compiled from /root/reference by
    oracle/Makefile into oracle/_ref/libs512ref.so (this script needs it).

Inputs follow the reference tests: srand48(0) then createRandVec
(NumType(drand48()*size), sort512test.cpp:77-82), which the oracle restates.

Run:  python tests/golden/make_golden.py      (writes ref_vectors.npz)
"""
import ctypes
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402


def p(a):
    return ctypes.c_void_p(a.ctypes.data)


def main():
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libs512ref.so is not built (make -C oracle, needs /root/reference)")
    out = {}
    # SmallSort16V: one array per length, every length the reference tests
    # (sort512test.cpp:1550-1566): 1..256 int, 1..128 double.
    oracle.srand48(0)
    ins, outs = [], []
    for n in range(1, 257):
        a = oracle.rand_vec(n, np.int32)
        b = a.copy()
        ref.ref_smallsort_i32(p(b), n)
        ins.append(a)
        outs.append(b)
    out["small_i32_in"] = np.concatenate(ins)
    out["small_i32_out"] = np.concatenate(outs)
    oracle.srand48(0)
    ins, outs = [], []
    for n in range(1, 129):
        a = oracle.rand_vec(n, np.float64)
        b = a.copy()
        ref.ref_smallsort_f64(p(b), n)
        ins.append(a)
        outs.append(b)
    out["small_f64_in"] = np.concatenate(ins)
    out["small_f64_out"] = np.concatenate(outs)
    # kv SmallSort16V (sort512test.cpp:1568-1599): keys out + the value rule
    oracle.srand48(0)
    ks, kouts, vouts = [], [], []
    for n in range(1, 257):
        k = oracle.rand_vec(n, np.int32)
        v = (k.astype(np.int64) * 100 + 1).astype(np.int32)
        k2, v2 = k.copy(), v.copy()
        ref.ref_kv_smallsort(p(k2), p(v2), n)
        ks.append(k)
        kouts.append(k2)
        vouts.append(v2)
    out["kvsmall_in"] = np.concatenate(ks)
    out["kvsmall_kout"] = np.concatenate(kouts)
    out["kvsmall_vout"] = np.concatenate(vouts)
    # Partition512 return value for every n in 1..1000, pivot n/2
    # (sort512test.cpp:1462-1471)
    for name, dt, fn in (("i32", np.int32, ref.ref_partition_i32), ("f64", np.float64, ref.ref_partition_f64)):
        oracle.srand48(0)
        bounds = []
        for n in range(1, 1001):
            a = oracle.rand_vec(n, dt)
            pivot = dt(n // 2)
            bounds.append(fn(p(a), 0, n - 1, int(pivot) if dt == np.int32 else float(pivot)))
        out[f"part_{name}_bound"] = np.array(bounds, dtype=np.int64)
    # Sort at every power of two up to 1024 (sort512test.cpp:1367-1377)
    for name, dt, fn in (("i32", np.int32, ref.ref_sort_i32), ("f64", np.float64, ref.ref_sort_f64)):
        oracle.srand48(0)
        ins, outs = [], []
        n = 1
        while n <= 1024:
            a = oracle.rand_vec(n, dt)
            b = a.copy()
            fn(p(b), n)
            ins.append(a)
            outs.append(b)
            n *= 2
        out[f"sort_{name}_in"] = np.concatenate(ins)
        out[f"sort_{name}_out"] = np.concatenate(outs)
    # FindMedian on two sorted halves of assorted sizes
    rng = np.random.default_rng(12345)
    fm = []
    for trial in range(300):
        na, nb = int(rng.integers(0, 40)), int(rng.integers(0, 40))
        if na + nb == 0:
            continue
        a = np.concatenate([np.sort(rng.integers(0, 30, na)), np.sort(rng.integers(0, 30, nb))]).astype(np.int32)
        mA, mB = ctypes.c_int(), ctypes.c_int()
        ref.ref_findmedian_i32(p(a), na, na + nb, ctypes.byref(mA), ctypes.byref(mB))
        fm.append((trial, na, nb, mA.value, mB.value))
        out[f"fm_{trial}"] = a
    out["findmedian"] = np.array(fm, dtype=np.int64)
    # Reference quirk Q1: +inf in a short leaf comes back as DBL_MAX
    q = np.array([np.inf, 1.0, 2.0, 3.0, 4.0])
    ref.ref_smallsort_f64(p(q), q.size)
    out["quirk_inf_out"] = q
    # Config 1 (1,048,576 int32 full range, seed 0): digest of the reference
    # Sort512::Sort output
    a = oracle.gen_i32_full(1 << 20, 0)
    ref.ref_sort_i32(p(a), a.size)
    out["cfg1_sha256"] = np.frombuffer(hashlib.sha256(a.tobytes()).digest(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"))


if __name__ == "__main__":
    main()
