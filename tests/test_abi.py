"""CPU checks of the drop-in boundary: the library loads, exports every symbol
the headers declare, validates arguments on the host, and the C++ drop-in
headers compile against it.  No GPU compute is issued here."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_declared_symbols(s512):
    declared = s512.declared_symbols()
    out = subprocess.run(["nm", "-D", "--defined-only", s512.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert len(declared) >= 30
    missing = sorted(declared - exported)
    assert not missing, f"declared but not exported: {missing}"
    for name in declared:
        assert getattr(s512.lib, name) is not None


def test_version(s512):
    assert b"gfx950" in s512.lib.s512_version()


def test_host_validation_without_gpu(s512):
    lib = s512.lib
    # n <= 1 is a no-op, null arrays are rejected before any device work
    assert lib.s512_sort_i32(None, 0) == 0
    assert lib.s512_sort_f64(None, 1) == 0
    assert lib.s512_sort_i32(None, 10) == 1
    assert b"null" in lib.s512_last_error()
    assert lib.s512kv_sort_i32(None, None, 5) == 1
    assert lib.s512_sort_f64_dev(None, 7, None) == 1
    # beyond the supported size (key/value sorts: fewer than 2^32 pairs)
    assert lib.s512kv_sort_i32_dev(ctypes.c_void_p(16), ctypes.c_void_p(16), 1 << 33, None) == 5
    # partition: null bound pointer, empty range
    assert lib.s512_partition_i32(None, 0, 5, 0, None) == 1
    b = ctypes.c_size_t(99)
    assert lib.s512_partition_i32(None, 5, 4, 0, ctypes.byref(b)) == 0 and b.value == 5
    # SmallSort16V length limits (16 vectors)
    assert lib.s512_smallsort16v_i32(None, 257) == 1
    assert lib.s512_smallsort16v_f64(None, 129) == 1
    assert lib.s512kv_smallsort16v_i32(None, None, 257) == 1
    # merge: degenerate centers are no-ops
    assert lib.s512_merge_inplace_i32(None, 10, 0) == 0
    assert lib.s512_merge_inplace_f64(None, 10, 10) == 0
    # options validation
    opt = s512.Options()
    opt.oversample = 100000
    assert lib.s512_sort_i32_dev_ex(ctypes.c_void_p(16), 100, ctypes.byref(opt), None, None) == 1
    for field, bad in (("scatter_variant", 12), ("count_variant", 6), ("tile_variant", 4)):
        opt = s512.Options()
        setattr(opt, field, bad)
        assert lib.s512_sort_f64_dev_ex(ctypes.c_void_p(16), 100, ctypes.byref(opt), None, None) == 1
        assert field.split("_")[0] in lib.s512_last_error().decode()
    # synthetic generator: unknown distribution
    assert lib.s512x_fill_i32(ctypes.c_void_p(16), 10, 1, 0, 0, None) == 1


def test_binding_type_errors(s512):
    with pytest.raises(TypeError):
        s512.sort(np.zeros(4, dtype=np.float32))
    with pytest.raises(ValueError):
        s512.sort(np.zeros(8, dtype=np.int32)[::2])
    with pytest.raises(ValueError):
        s512.kv_sort(np.zeros(3, dtype=np.int32), np.zeros(4, dtype=np.int32))


def test_dropin_headers_compile(tmp_path):
    """A translation unit written like the reference's users compiles and
    links against libs512.so with only the drop-in headers."""
    src = tmp_path / "user.cpp"
    src.write_text(
        '#include "sort512.hpp"\n#include "sort512kv.hpp"\n#include "sort512kv.hpp"\n'
        "int main(){ int a[3]={3,1,2}; double d[2]={2.,1.}; int v[3]={0,1,2};\n"
        "Sort512::Sort<int,size_t>(a,3); Sort512::SortOmp<double,int>(d,2);\n"
        "Sort512::SortOmpPartition<int,size_t>(a,3); Sort512::SortOmpMerge<double,size_t>(d,2);\n"
        "Sort512::SortOmpMergeDeps<int,size_t>(a,3); Sort512::SortOmpParMerge<double,size_t>(d,2);\n"
        "size_t b = Sort512::Partition512<size_t>(a,0,2,2); (void)b; long c = Sort512::Partition512<long>(d,0L,1L,1.5); (void)c;\n"
        "Sort512::SmallSort16V(a,3); Sort512::SmallSort16V(d,2);\n"
        "Sort512kv::Sort<int,size_t>(a,v,3); Sort512kv::SortOmpPartition<int,size_t>(a,v,3);\n"
        "Sort512kv::Partition512<size_t>(a,v,0,2,1); Sort512kv::SmallSort16V(a,v,3);\n"
        "ParallelInplace::parallelMergeInPlace(d,2,1); return 0; }\n")
    lib = os.path.join(ROOT, "avx-512-sort_amd")
    r = subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                        "-o", str(tmp_path / "user"), "-L", lib, "-ls512", "-Wl,-rpath," + lib],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_header_rejects_other_types(tmp_path):
    src = tmp_path / "bad.cpp"
    src.write_text('#include "sort512.hpp"\nint main(){ float f[2]={1,0}; Sort512::Sort<float,size_t>(f,2); }\n')
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "sorts int and double" in r.stderr


def test_argsort_rejects_more_than_int32_positions(s512):
    """Positions are int32: n >= 2^31 is refused before any memory is touched."""
    import ctypes
    a = (ctypes.c_double * 1)()
    i = (ctypes.c_int32 * 1)()
    rc = s512.lib.s512_argsort_f64(ctypes.cast(a, ctypes.c_void_p), ctypes.cast(i, ctypes.c_void_p), 1 << 31, None)
    assert rc == 5  # S512_ERR_TOO_LARGE


def test_topk_select_argument_checks_on_the_host(s512):
    """k > n (top-k) and k >= n (selection) are refused before any GPU call."""
    import ctypes
    a = (ctypes.c_double * 4)()
    o = (ctypes.c_double * 8)()
    pa, po = ctypes.cast(a, ctypes.c_void_p), ctypes.cast(o, ctypes.c_void_p)
    assert s512.lib.s512_topk_f64(pa, 4, 5, po, None) == 1  # S512_ERR_INVALID
    assert s512.lib.s512_select_f64(pa, 4, 4, po, None) == 1
    assert s512.lib.s512_topk_f64(pa, 4, 0, po, None) == 0  # k = 0: nothing to do
