"""GPU tests of the 64-bit and unsigned key sorts (int64, uint64, uint32; no
reference counterpart).  Ascending order of integers is unique, so every
result must equal numpy's sort bit for bit.  Inputs reach every tile class
and both partition levels, with the types' extremes (INT64_MIN/MAX,
UINT64_MAX, 0), heavy duplicates, sorted and reversed runs; host and device
arrays; the kernel shapes selectable by options; and 268,435,456 int64 keys
by size-independent properties."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DTYPES = [np.int64, np.uint64, np.uint32]
SIZES = [0, 1, 2, 100, 1000, 4097, 70_000, 1_000_003, 6_000_001]


def _gen(dtype, kind, n, seed):
    rng = np.random.default_rng(seed)
    info = np.iinfo(dtype)
    if kind == "full":
        a = rng.integers(info.min, info.max, size=n, dtype=dtype, endpoint=True)
    elif kind == "dups":
        a = rng.integers(0, 50, size=n).astype(dtype)
    elif kind == "sorted":
        a = np.sort(rng.integers(info.min, info.max, size=n, dtype=dtype, endpoint=True))
    elif kind == "reversed":
        a = np.sort(rng.integers(info.min, info.max, size=n, dtype=dtype, endpoint=True))[::-1].copy()
    elif kind == "extremes":
        pool = np.array([info.min, info.max, 0, 1, info.max - 1, info.min + 1], dtype=dtype)
        a = pool[rng.integers(0, len(pool), size=n)]
    else:
        raise ValueError(kind)
    return a


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("dtype", DTYPES)
def test_wide_keys_sizes_device(s512, dev, dtype, n):
    import torch
    a = _gen(dtype, "full", n, 1)
    t = torch.from_numpy(a.copy()).to(dev)
    s512.sort(t)
    assert np.array_equal(t.cpu().numpy(), np.sort(a))


@pytest.mark.parametrize("kind", ["dups", "sorted", "reversed", "extremes"])
@pytest.mark.parametrize("n", [3000, 300_000, 3_000_001])
@pytest.mark.parametrize("dtype", DTYPES)
def test_wide_keys_distributions(s512, dev, dtype, kind, n):
    import torch
    a = _gen(dtype, kind, n, 2)
    t = torch.from_numpy(a.copy()).to(dev)
    s512.sort(t)
    assert np.array_equal(t.cpu().numpy(), np.sort(a))


@pytest.mark.parametrize("dtype", DTYPES)
def test_wide_keys_host_arrays(s512, dtype):
    for n in (5000, 9_000_001):
        a = _gen(dtype, "full", n, 3)
        want = np.sort(a)
        s512.sort(a)
        assert np.array_equal(a, want)


@pytest.mark.parametrize("opts", [{"scatter_variant": 5}, {"count_variant": 4}, {"tile_variant": 1},
                                  {"max_fanout_log2": 6, "target_leaf": 512}])
@pytest.mark.parametrize("dtype", [np.int64, np.uint32])
def test_wide_keys_options(s512, dev, dtype, opts):
    import torch
    a = _gen(dtype, "full", 2_000_003, 4)
    t = torch.from_numpy(a.copy()).to(dev)
    o = s512.Options()
    for k, v in opts.items():
        setattr(o, k, v)
    st = s512.sort_ex(t, o)
    assert st.levels >= 1
    assert np.array_equal(t.cpu().numpy(), np.sort(a))


def test_wide_keys_learned_rejected(s512, dev):
    import torch
    t = torch.zeros(1000, dtype=torch.int64, device=dev)
    with pytest.raises(s512.S512Error):
        s512.sort_learned(t)


def test_reference_only_entry_points_reject_wide_keys(s512):
    with pytest.raises(TypeError):
        s512.partition(np.zeros(10, dtype=np.int64), 0, 9, 0)
    with pytest.raises(TypeError):
        s512.merge_inplace(np.zeros(10, dtype=np.uint32), 5)


def test_int64_full_size(s512, dev):
    """268,435,456 int64 keys over the full range: non-decreasing, same multiset."""
    import torch
    n = 268_435_456
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    t = torch.randint(-(1 << 63), (1 << 63) - 1, (n,), dtype=torch.int64, device=dev, generator=g)

    def h(x):
        x = x * -7046029254386353131
        x = x ^ (x >> 29)
        x = x * -4658895280553007687
        return int((x ^ (x >> 32)).sum().item())
    h0 = h(t)
    s512.sort(t)
    assert bool(torch.all(t[1:] >= t[:-1]).item())
    assert h(t) == h0
