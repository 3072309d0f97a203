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
    "s512_release": ([], ctypes.c_int),
    "s512_sort_i32": ([_vp, _c_size], ctypes.c_int),
    "s512_sort_f64": ([_vp, _c_size], ctypes.c_int),
    "s512_sort_i32_dev": ([_vp, _c_size, _vp], ctypes.c_int),
    "s512_sort_f64_dev": ([_vp, _c_size, _vp], ctypes.c_int),
    "s512kv_sort_i32": ([_vp, _vp, _c_size], ctypes.c_int),
    "s512kv_sort_i32_dev": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512_sort_i32_dev_ex": ([_vp, _c_size, _vp, _vp, _vp], ctypes.c_int),
    "s512kv_sort_f64": ([_vp, _vp, _c_size], ctypes.c_int),
    "s512kv_sort_i64": ([_vp, _vp, _c_size], ctypes.c_int),
    "s512kv_sort_u64": ([_vp, _vp, _c_size], ctypes.c_int),
    "s512kv_sort_f64_dev": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512kv_sort_i64_dev": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512kv_sort_u64_dev": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512kv_sort_i32_v64": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512kv_sort_f64_v64": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512kv_sort_i64_v64": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512kv_sort_u64_v64": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512_argsort_i32": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512_argsort_f64": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512_argsort_i64": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512_argsort_u64": ([_vp, _vp, _c_size, _vp], ctypes.c_int),
    "s512_topk_i32": ([_vp, _c_size, _c_size, _vp, _vp], ctypes.c_int),
    "s512_topk_f64": ([_vp, _c_size, _c_size, _vp, _vp], ctypes.c_int),
    "s512_topk_i64": ([_vp, _c_size, _c_size, _vp, _vp], ctypes.c_int),
    "s512_topk_u64": ([_vp, _c_size, _c_size, _vp, _vp], ctypes.c_int),
    "s512_select_i32": ([_vp, _c_size, _c_size, _vp, _vp], ctypes.c_int),
    "s512_select_f64": ([_vp, _c_size, _c_size, _vp, _vp], ctypes.c_int),
    "s512_select_i64": ([_vp, _c_size, _c_size, _vp, _vp], ctypes.c_int),
    "s512_select_u64": ([_vp, _c_size, _c_size, _vp, _vp], ctypes.c_int),
    "s512_sort_i64": ([_vp, _c_size], ctypes.c_int),
    "s512_sort_u64": ([_vp, _c_size], ctypes.c_int),
    "s512_sort_u32": ([_vp, _c_size], ctypes.c_int),
    "s512_sort_i64_dev": ([_vp, _c_size, _vp], ctypes.c_int),
    "s512_sort_u64_dev": ([_vp, _c_size, _vp], ctypes.c_int),
    "s512_sort_u32_dev": ([_vp, _c_size, _vp], ctypes.c_int),
    "s512_sort_i64_dev_ex": ([_vp, _c_size, _vp, _vp, _vp], ctypes.c_int),
    "s512_sort_u64_dev_ex": ([_vp, _c_size, _vp, _vp, _vp], ctypes.c_int),
    "s512_sort_u32_dev_ex": ([_vp, _c_size, _vp, _vp, _vp], ctypes.c_int),
    "s512_sort_f64_dev_ex": ([_vp, _c_size, _vp, _vp, _vp], ctypes.c_int),
    "s512kv_sort_i32_dev_ex": ([_vp, _vp, _c_size, _vp, _vp, _vp], ctypes.c_int),
    "s512_sort_merge_i32": ([_vp, _c_size], ctypes.c_int),
    "s512_sort_merge_f64": ([_vp, _c_size], ctypes.c_int),
    "s512_sort_merge_i32_dev": ([_vp, _c_size, _vp], ctypes.c_int),
    "s512_sort_merge_f64_dev": ([_vp, _c_size, _vp], ctypes.c_int),
    "s512_partition_i32": ([_vp, _c_size, _c_size, ctypes.c_int32, ctypes.POINTER(_c_size)], ctypes.c_int),
    "s512_partition_f64": ([_vp, _c_size, _c_size, ctypes.c_double, ctypes.POINTER(_c_size)], ctypes.c_int),
    "s512_partition_i32_dev": ([_vp, _c_size, _c_size, ctypes.c_int32, ctypes.POINTER(_c_size), _vp], ctypes.c_int),
    "s512_partition_f64_dev": ([_vp, _c_size, _c_size, ctypes.c_double, ctypes.POINTER(_c_size), _vp], ctypes.c_int),
    "s512_partition_f64_dev_noalloc": ([_vp, _c_size, _c_size, ctypes.c_double, ctypes.POINTER(_c_size), _vp],
                                       ctypes.c_int),
    "s512kv_partition_i32": ([_vp, _vp, _c_size, _c_size, ctypes.c_int32, ctypes.POINTER(_c_size)], ctypes.c_int),
    "s512kv_partition_i32_dev": ([_vp, _vp, _c_size, _c_size, ctypes.c_int32, ctypes.POINTER(_c_size), _vp],
                                 ctypes.c_int),
    "s512_smallsort16v_i32": ([_vp, _c_size], ctypes.c_int),
    "s512_smallsort16v_f64": ([_vp, _c_size], ctypes.c_int),
    "s512kv_smallsort16v_i32": ([_vp, _vp, _c_size], ctypes.c_int),
    "s512_merge_inplace_i32": ([_vp, _c_size, _c_size], ctypes.c_int),
    "s512_merge_inplace_f64": ([_vp, _c_size, _c_size], ctypes.c_int),
    "s512_merge_inplace_i32_dev": ([_vp, _c_size, _c_size, _vp], ctypes.c_int),
    "s512_merge_inplace_f64_dev": ([_vp, _c_size, _c_size, _vp], ctypes.c_int),
    "s512_mlp_nparams": ([], ctypes.c_uint32),
    "s512_mlp_loss_grad_f32": ([_vp, _vp, _vp, ctypes.c_uint32, _vp, _vp, _vp], ctypes.c_int),
    "s512_segsort_i32": ([_vp, _vp, _c_size], ctypes.c_int),
    "s512_segsort_f64": ([_vp, _vp, _c_size], ctypes.c_int),
    "s512_segsort_i32_dev": ([_vp, _vp, _c_size, _vp, _vp], ctypes.c_int),
    "s512_segsort_f64_dev": ([_vp, _vp, _c_size, _vp, _vp], ctypes.c_int),
    "s512_segsort_i32_dev_ex": ([_vp, _vp, _c_size, _vp, _vp, _vp], ctypes.c_int),
    "s512_segsort_f64_dev_ex": ([_vp, _vp, _c_size, _vp, _vp, _vp], ctypes.c_int),
    "s512_segsort_plan_create": ([ctypes.c_int, _vp, _vp, _c_size, _c_size, ctypes.POINTER(_vp)], ctypes.c_int),
    "s512_segsort_plan_launch": ([_vp, _vp, _vp], ctypes.c_int),
    "s512_segsort_plan_destroy": ([_vp], ctypes.c_int),
    "s512x_fill_f64": ([_vp, _c_size, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, _vp], ctypes.c_int),
    "s512x_fill_i32": ([_vp, _c_size, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, _vp], ctypes.c_int),
}
for _name, (_args, _res) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res


NPHASE = 10
PHASES = ["sample", "sample_sort", "tree", "count", "scan", "scatter", "plan", "copy", "tile", "train"]


class Options(ctypes.Structure):
    """s512_options (include/s512.h)."""
    _fields_ = [("target_leaf", ctypes.c_uint32), ("oversample", ctypes.c_uint32),
                ("max_fanout_log2", ctypes.c_uint32), ("profile", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("classifier", ctypes.c_uint32), ("sample_permille", ctypes.c_uint32),
                ("train_steps", ctypes.c_uint32), ("imbalance_pct", ctypes.c_uint32),
                ("scatter_variant", ctypes.c_uint32), ("count_variant", ctypes.c_uint32),
                ("tile_variant", ctypes.c_uint32), ("engine", ctypes.c_uint32),
                ("region_slack", ctypes.c_uint32)]


ENGINE_AUTO, ENGINE_TWO_PASS, ENGINE_ONE_PASS = 0, 1, 2


SPLITTERS, LEARNED = 0, 1


class Stats(ctypes.Structure):
    """s512_stats (include/s512.h)."""
    _fields_ = [("levels", ctypes.c_int32), ("tiles", ctypes.c_uint64), ("kernel_ms", ctypes.c_double * NPHASE),
                ("launches", ctypes.c_uint32 * NPHASE), ("keys", ctypes.c_uint64 * NPHASE),
                ("classifier", ctypes.c_int32), ("learned_fallback", ctypes.c_int32),
                ("learned_levels", ctypes.c_int32), ("learned_max_bucket_ratio", ctypes.c_float),
                ("model_loss", ctypes.c_float), ("engine", ctypes.c_int32), ("region_overflow", ctypes.c_int32),
                ("learned_leaf_max_ratio", ctypes.c_float)]

    def as_dict(self):
        return {"levels": self.levels, "tiles": self.tiles, "classifier": self.classifier,
                "engine": self.engine, "region_overflow": self.region_overflow,
                "learned_fallback": self.learned_fallback, "learned_levels": self.learned_levels,
                "learned_max_bucket_ratio": self.learned_max_bucket_ratio, "model_loss": self.model_loss,
                "learned_leaf_max_ratio": self.learned_leaf_max_ratio,
                "phases": {PHASES[i]: {"ms": self.kernel_ms[i], "launches": self.launches[i], "keys": self.keys[i]}
                           for i in range(NPHASE) if self.launches[i] or self.keys[i]}}


SYNTH = {"uniform": 0, "normal": 1, "distinct": 2, "sorted": 3, "reversed": 4, "index": 5, "mixture": 6,
         "piecewise": 7, "zipf": 8}


def fill(t, kind, seed=0, param=0):
    """Device-side synthetic input (include/s512_synth.h) into a torch tensor."""
    k = _ref_kind(t)
    _check(getattr(lib, f"s512x_fill_{k}")(_ptr(t), t.numel(), SYNTH[kind], seed, param, _stream_of(t)))
    return t


def sort_ex(a, options=None, profile=False):
    """Sort with options and statistics (device tensors); returns Stats."""
    k = _kind(a)
    opt = options or Options()
    if profile:
        opt.profile = 1
    st = Stats()
    _check(getattr(lib, f"s512_sort_{k}_dev_ex")(_ptr(a), a.numel(), ctypes.byref(opt), ctypes.byref(st),
                                                 _stream_of(a)))
    return st


def kv_sort_ex(keys, values, options=None, profile=False):
    opt = options or Options()
    if profile:
        opt.profile = 1
    st = Stats()
    _check(lib.s512kv_sort_i32_dev_ex(_ptr(keys), _ptr(values), keys.numel(), ctypes.byref(opt), ctypes.byref(st),
                                      _stream_of(keys)))
    return st


def declared_symbols():
    """Names declared in include/*.h (parsed from the headers)."""
    import glob
    import re
    names = set()
    for h in glob.glob(os.path.join(_HERE, "..", "include", "*.h")):
        for m in re.finditer(r"^\s*(?:s512_status|const char\*|uint32_t)\s+(s512\w+)\s*\(", open(h).read(), re.M):
            names.add(m.group(1))
    return names


def release():
    """Free the current device's cached workspaces (the next call allocates again)."""
    _check(lib.s512_release())


def sort_learned(a, profile=False, **kw):
    """Sort with the learned CDF partitioner (fallback: sampled splitters)."""
    opt = Options()
    opt.classifier = LEARNED
    for k, v in kw.items():
        setattr(opt, k, v)
    return sort_ex(a, opt, profile=profile)


def mlp_loss_grad(params, u, t):
    """GPU fp32 loss and gradient of the CDF model (device float32 tensors)."""
    import torch
    g = torch.empty(lib.s512_mlp_nparams(), dtype=torch.float32, device=params.device)
    loss = torch.empty(1, dtype=torch.float32, device=params.device)
    _check(lib.s512_mlp_loss_grad_f32(_ptr(params), _ptr(u), _ptr(t), u.numel(), _ptr(g), _ptr(loss),
                                      _stream_of(params)))
    return float(loss.item()), g


def _check(status):
    if status != 0:
        raise S512Error(status, lib.s512_last_error().decode(errors="replace"))


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _stream_of(t):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


_DT = {np.dtype(np.int32): "i32", np.dtype(np.float64): "f64", np.dtype(np.int64): "i64",
       np.dtype(np.uint64): "u64", np.dtype(np.uint32): "u32"}


def _kind(x):
    if _is_torch(x):
        import torch
        k = {torch.int32: "i32", torch.float64: "f64", torch.int64: "i64", torch.uint64: "u64",
             torch.uint32: "u32"}.get(x.dtype)
    else:
        k = _DT.get(np.dtype(x.dtype))
    if k is None:
        raise TypeError(f"unsupported dtype {x.dtype}: sorts cover int32 and float64 (the reference API) and "
                        f"int64, uint64, uint32 keys")
    return k


def _ref_kind(x):
    """Kind of an array for the entry points that cover the reference's key types only."""
    k = _kind(x)
    if k not in ("i32", "f64"):
        raise TypeError(f"unsupported dtype {x.dtype}: this entry point covers int32 and float64 (the reference API)")
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
    """Sort512kv::Sort: sort keys ascending, values follow their keys.
    int32 keys with int32 values are the reference's layout; float64, int64
    and uint64 keys, and 8-byte values (int64 / uint64 / float64), are an
    extension."""
    k = _kind(keys)
    if k == "u32":
        raise TypeError("key/value sorts take int32, float64, int64 or uint64 keys")
    n = keys.numel() if _is_torch(keys) else keys.size
    m = values.numel() if _is_torch(values) else values.size
    if n != m:
        raise ValueError("keys and values differ in length")
    vk = _kind(values)
    if vk in ("i64", "u64", "f64"):
        stream = _stream_of(keys) if _dev(keys) else None
        _check(getattr(lib, f"s512kv_sort_{k}_v64")(_ptr(keys), _ptr(values), n, stream))
        return keys, values
    if vk != "i32":
        raise TypeError("values must be int32 or an 8-byte type")
    if _is_torch(keys) and keys.is_cuda:
        _check(getattr(lib, f"s512kv_sort_{k}_dev")(_ptr(keys), _ptr(values), n, _stream_of(keys)))
    else:
        _check(getattr(lib, f"s512kv_sort_{k}")(_ptr(keys), _ptr(values), n))
    return keys, values


def topk(keys, k, out=None):
    """The k smallest keys in ascending order (keys unchanged).  `out` (same
    dtype, length k) may be given; by default it is allocated beside the keys."""
    kd = _kind(keys)
    if kd == "u32":
        raise TypeError("topk covers int32, float64, int64 and uint64 keys")
    n = _n(keys)
    if out is None:
        if _is_torch(keys):
            import torch
            out = torch.empty(k, dtype=keys.dtype, device=keys.device)
        else:
            out = np.empty(k, dtype=keys.dtype)
    if _kind(out) != kd or _n(out) != k:
        raise ValueError("out must have the keys' dtype and length k")
    stream = _stream_of(keys) if _dev(keys) else None
    _check(getattr(lib, f"s512_topk_{kd}")(_ptr(keys), n, k, _ptr(out), stream))
    return out


def select(keys, k):
    """The k-th smallest key (0-based), as a Python number."""
    kd = _kind(keys)
    if kd == "u32":
        raise TypeError("select covers int32, float64, int64 and uint64 keys")
    ct = {"i32": ctypes.c_int32, "f64": ctypes.c_double, "i64": ctypes.c_int64, "u64": ctypes.c_uint64}[kd]
    v = ct()
    stream = _stream_of(keys) if _dev(keys) else None
    _check(getattr(lib, f"s512_select_{kd}")(_ptr(keys), _n(keys), k, ctypes.cast(ctypes.byref(v), _vp), stream))
    return v.value


def argsort(keys, out=None):
    """Positions that sort `keys` ascending (int32; keys unchanged; equal keys
    in no particular order).  `out` (int32, same length) may be given; by
    default it is allocated beside the keys (device tensor or numpy array)."""
    k = _kind(keys)
    if k == "u32":
        raise TypeError("argsort covers int32, float64, int64 and uint64 keys")
    n = _n(keys)
    if out is None:
        if _is_torch(keys):
            import torch
            out = torch.empty(n, dtype=torch.int32, device=keys.device)
        else:
            out = np.empty(n, dtype=np.int32)
    if _kind(out) != "i32" or _n(out) != n:
        raise ValueError("out must be int32 of the keys' length")
    stream = _stream_of(keys) if _dev(keys) else None
    _check(getattr(lib, f"s512_argsort_{k}")(_ptr(keys), _ptr(out), n, stream))
    return out


def _n(x):
    return x.numel() if _is_torch(x) else x.size


def _dev(x):
    return _is_torch(x) and x.is_cuda


def sort_merge(a):
    """Sort512::SortOmpMerge / SortOmpMergeDeps / SortOmpParMerge: GPU merge sort."""
    k = _ref_kind(a)
    if _dev(a):
        _check(getattr(lib, f"s512_sort_merge_{k}_dev")(_ptr(a), _n(a), _stream_of(a)))
    else:
        _check(getattr(lib, f"s512_sort_merge_{k}")(_ptr(a), _n(a)))
    return a


def partition(a, left, right, pivot):
    """Sort512::Partition512: returns left + #{a[left..right] <= pivot}."""
    k = _ref_kind(a)
    out = _c_size()
    cast = int if k == "i32" else float
    if _dev(a):
        _check(getattr(lib, f"s512_partition_{k}_dev")(_ptr(a), left, right, cast(pivot), ctypes.byref(out),
                                                       _stream_of(a)))
    else:
        _check(getattr(lib, f"s512_partition_{k}")(_ptr(a), left, right, cast(pivot), ctypes.byref(out)))
    return out.value


def partition_noalloc(a, left, right, pivot):
    """Partition512 of a float64 device tensor with the library's fixed
    workspace (no allocation; at most 8,384,480 keys per call)."""
    out = _c_size()
    _check(lib.s512_partition_f64_dev_noalloc(_ptr(a), left, right, float(pivot), ctypes.byref(out), _stream_of(a)))
    return out.value


def kv_partition(keys, values, left, right, pivot):
    """Sort512kv::Partition512."""
    out = _c_size()
    if _dev(keys):
        _check(lib.s512kv_partition_i32_dev(_ptr(keys), _ptr(values), left, right, int(pivot), ctypes.byref(out),
                                            _stream_of(keys)))
    else:
        _check(lib.s512kv_partition_i32(_ptr(keys), _ptr(values), left, right, int(pivot), ctypes.byref(out)))
    return out.value


def smallsort16v(a):
    """Sort512::SmallSort16V (<= 256 int32 / <= 128 float64 keys)."""
    _check(getattr(lib, f"s512_smallsort16v_{_ref_kind(a)}")(_ptr(a), _n(a)))
    return a


def kv_smallsort16v(keys, values):
    """Sort512kv::SmallSort16V (<= 256 pairs)."""
    _check(lib.s512kv_smallsort16v_i32(_ptr(keys), _ptr(values), _n(keys)))
    return keys, values


def merge_inplace(a, center):
    """ParallelInplace::parallelMergeInPlace: merge sorted a[:center] and a[center:]."""
    k = _ref_kind(a)
    if _dev(a):
        _check(getattr(lib, f"s512_merge_inplace_{k}_dev")(_ptr(a), _n(a), center, _stream_of(a)))
    else:
        _check(getattr(lib, f"s512_merge_inplace_{k}")(_ptr(a), _n(a), center))
    return a


def _offsets_ok(keys, offsets):
    if _is_torch(offsets):
        import torch
        if offsets.dtype not in (torch.int64, torch.uint64):
            raise TypeError("offsets must be int64 / uint64")
        if _dev(offsets) != _dev(keys):
            raise ValueError("keys and offsets must both be on the device or both on the host")
    else:
        if np.dtype(offsets.dtype) not in (np.dtype(np.int64), np.dtype(np.uint64)):
            raise TypeError("offsets must be int64 / uint64")
    if _n(offsets) < 1:
        raise ValueError("offsets holds nseg + 1 entries")


def segsort(keys, offsets, stats=False, options=None):
    """Sort every segment keys[offsets[i]:offsets[i+1]] independently, in place."""
    k = _ref_kind(keys)
    _offsets_ok(keys, offsets)
    nseg = _n(offsets) - 1
    if _dev(keys):
        st = Stats()
        _check(getattr(lib, f"s512_segsort_{k}_dev_ex")(_ptr(keys), _ptr(offsets), nseg,
                                                        ctypes.byref(options) if options is not None else None,
                                                        ctypes.byref(st), _stream_of(keys)))
        return st if stats else keys
    if options is not None:
        raise ValueError("options apply to device arrays")
    _check(getattr(lib, f"s512_segsort_{k}")(_ptr(keys), _ptr(offsets), nseg))
    return keys


class SegsortPlan:
    """A captured segmented sort for fixed device tensors (keys, offsets)."""

    def __init__(self, keys, offsets):
        k = _ref_kind(keys)
        _offsets_ok(keys, offsets)
        if not _dev(keys):
            raise ValueError("a plan needs device tensors")
        self.keys, self.offsets = keys, offsets  # keep them alive
        self.h = _vp()
        _check(lib.s512_segsort_plan_create(0 if k == "i32" else 1, _ptr(keys), _ptr(offsets), _n(offsets) - 1,
                                            _n(keys), ctypes.byref(self.h)))

    def launch(self):
        st = Stats()
        _check(lib.s512_segsort_plan_launch(self.h, ctypes.byref(st), _stream_of(self.keys)))
        return st

    def close(self):
        if self.h:
            lib.s512_segsort_plan_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lognormal_offsets(nseg, median=512.0, sigma=1.5, lo=1, hi=1 << 20, seed=0):
    """Offsets (uint64, nseg + 1 entries) of nseg segments whose sizes are
    lognormal(median, sigma) rounded and clipped to [lo, hi]; numpy's PCG64
    seeded with `seed` (synthetic batched-sort workload)."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.rint(median * np.exp(sigma * rng.standard_normal(nseg))), lo, hi).astype(np.uint64)
    off = np.zeros(nseg + 1, dtype=np.uint64)
    np.cumsum(sizes, out=off[1:])
    return off
