"""ctypes binding of libs512.so (the C ABI declared in include/s512.h).

Host-side mirror of the reference API for Python callers, tests and the
benchmark.  Names follow the reference (Sort512::Sort, Sort512kv::Sort, ...):

    sort(a)                    Sort512::Sort            sort512.hpp:5960
    kv_sort(keys, values)      Sort512kv::Sort          sort512kv.hpp:6132

Arguments may be numpy arrays (host memory: the library stages them through
the GPU) or torch tensors on a HIP device (sorted in place on the tensor's
current stream).  Errors raise S512Error with the library's message.  There is
no CPU fallback: if libs512.so is missing or fails to load, importing this
module raises.
"""
import ctypes
import os

import numpy as np

# torch ships its own libamdhip64.so (soname libamdhip64.so.7).  Loading it
# first lets libs512.so bind to that same runtime; loaded the other way round
# the process holds two HIP runtimes and the second cannot open the GPU.
try:
    import torch  # noqa: F401
except ImportError:  # C-only use: libs512.so binds to /opt/rocm's runtime
    torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libs512.so")


class S512Error(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"s512 status {status}: {msg}")
        self.status = status


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C avx-512-sort_amd")
lib = ctypes.CDLL(LIB_PATH)

_c_size = ctypes.c_size_t
_vp = ctypes.c_void_p

_SIGS = {
    "s512_last_error": ([], ctypes.c_char_p),
    "s512_version": ([], ctypes.c_char_p),
    "s512_sort_i32": ([_vp, _c_size], ctypes.c_int),
    "s512_sort_f64": ([_vp, _c_size], ctypes.c_int),
    "s512_sort_i32_dev": ([_vp, _c_size, _vp], ctypes.c_int),
    "s512_sort_f64_dev": ([_vp, _c_size, _vp], ctypes.c_int),
    "s512kv_sort_i32": ([_vp, _vp, _c_size], ctypes.c_int),
    "s512kv_sort_i32_dev": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
}
for _name, (_args, _res) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res


def declared_symbols():
    """Names declared in include/*.h (parsed from the headers)."""
    import glob
    import re
    names = set()
    for h in glob.glob(os.path.join(_HERE, "..", "include", "*.h")):
        for m in re.finditer(r"^\s*(?:s512_status|const char\*)\s+(s512\w+)\s*\(", open(h).read(), re.M):
            names.add(m.group(1))
    return names


def _check(status):
    if status != 0:
        raise S512Error(status, lib.s512_last_error().decode(errors="replace"))


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _stream_of(t):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


_DT = {np.dtype(np.int32): "i32", np.dtype(np.float64): "f64"}


def _kind(x):
    if _is_torch(x):
        import torch
        k = {torch.int32: "i32", torch.float64: "f64"}.get(x.dtype)
    else:
        k = _DT.get(np.dtype(x.dtype))
    if k is None:
        raise TypeError(f"unsupported dtype {x.dtype}: the reference API covers int32 and float64")
    return k


def _ptr(x):
    if _is_torch(x):
        if not x.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return ctypes.c_void_p(x.data_ptr())
    if not x.flags["C_CONTIGUOUS"] or not x.flags["WRITEABLE"]:
        raise ValueError("array must be C-contiguous and writeable")
    return ctypes.c_void_p(x.ctypes.data)


def sort(a):
    """Sort512::Sort: in-place ascending sort of int32 / float64 keys."""
    k = _kind(a)
    n = a.numel() if _is_torch(a) else a.size
    if _is_torch(a) and a.is_cuda:
        _check(getattr(lib, f"s512_sort_{k}_dev")(_ptr(a), n, _stream_of(a)))
    else:
        _check(getattr(lib, f"s512_sort_{k}")(_ptr(a), n))
    return a


def kv_sort(keys, values):
    """Sort512kv::Sort: sort int32 keys ascending, values follow their keys."""
    if _kind(keys) != "i32" or _kind(values) != "i32":
        raise TypeError("Sort512kv covers int32 keys and int32 values")
    n = keys.numel() if _is_torch(keys) else keys.size
    m = values.numel() if _is_torch(values) else values.size
    if n != m:
        raise ValueError("keys and values differ in length")
    if _is_torch(keys) and keys.is_cuda:
        _check(lib.s512kv_sort_i32_dev(_ptr(keys), _ptr(values), n, _stream_of(keys)))
    else:
        _check(lib.s512kv_sort_i32(_ptr(keys), _ptr(values), n))
    return keys, values
