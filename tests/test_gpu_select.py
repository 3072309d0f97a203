"""GPU tests of top-k and selection (no reference counterpart).  The k
smallest keys in ascending order are a prefix of the full ascending sort, so
every result is compared bit for bit with the first k keys of the keys-only
sort (itself checked against the oracle).  Cases: all four key types, k from
1 to n, heavy ties at the threshold (all keys equal), the types' extremes,
host and device memory, and the argument checks."""
import numpy as np
import pytest

from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu


def _keys(orc, kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "f64":
        return orc.gen_f64("normal", n, seed)
    if kind == "i32":
        return orc.gen_i32_full(n, seed)
    if kind == "i64":
        return rng.integers(-(1 << 63), (1 << 63) - 1, size=n, dtype=np.int64, endpoint=True)
    if kind == "u64":
        return rng.integers(0, (1 << 64) - 1, size=n, dtype=np.uint64, endpoint=True)
    raise ValueError(kind)


def _bits(a):
    return a.view(np.uint64) if a.dtype == np.float64 else a


def _sorted(s512, a):
    b = a.copy()
    s512.sort(b)
    return b


@pytest.mark.parametrize("n", [1, 100, 5000, 1_000_003, 5_000_001])
@pytest.mark.parametrize("kind", ["i32", "f64", "i64", "u64"])
def test_topk_and_select_device(s512, orc, dev, kind, n):
    a = _keys(orc, kind, n, 4)
    want = _sorted(s512, a)
    t = to_dev(a, dev)
    for k in sorted({1, min(10, n), max(1, n // 1000), max(1, n // 2), n}):
        got = to_host(s512.topk(t, k))
        assert np.array_equal(_bits(got), _bits(want[:k])), (kind, n, k)
        kth = s512.select(t, k - 1)
        assert _bits(np.array([kth], dtype=a.dtype))[0] == _bits(want[k - 1:k])[0], (kind, n, k)
    assert np.array_equal(to_host(t), a)  # keys unchanged


@pytest.mark.parametrize("kind", ["i32", "f64"])
def test_topk_ties_and_extremes(s512, orc, dev, kind):
    n = 2_000_000
    eq = np.full(n, 7, dtype=np.int32 if kind == "i32" else np.float64)
    t = to_dev(eq, dev)
    assert np.array_equal(to_host(s512.topk(t, n // 2)), eq[: n // 2])
    if kind == "i32":
        pool = np.array([np.iinfo(np.int32).min, -1, 0, 1, np.iinfo(np.int32).max], dtype=np.int32)
    else:
        pool = np.array([-np.inf, -1.0, -0.0, 0.0, 1.0, np.inf], dtype=np.float64)
    a = pool[np.random.default_rng(1).integers(0, len(pool), n)]
    want = _sorted(s512, a)
    t = to_dev(a, dev)
    for k in (1, 1000, n // 3, n - 1, n):
        assert np.array_equal(_bits(to_host(s512.topk(t, k))), _bits(want[:k])), k


def test_topk_host_and_mixed(s512, orc, dev):
    import torch
    n = 9_000_001
    a = _keys(orc, "f64", n, 8)
    want = _sorted(s512, a)
    k = 123_457
    got = s512.topk(a.copy(), k)  # host keys, host out
    assert np.array_equal(_bits(got), _bits(want[:k]))
    out = torch.empty(k, dtype=torch.float64, device=dev)  # host keys, device out
    s512.topk(a.copy(), k, out=out)
    assert np.array_equal(_bits(to_host(out)), _bits(want[:k]))
    got = s512.topk(to_dev(a, dev), k)  # device keys, device out
    assert np.array_equal(_bits(to_host(got)), _bits(want[:k]))
    assert s512.select(a.copy(), n - 1) == want[-1]


def test_topk_argument_errors(s512, dev):
    import torch
    t = torch.zeros(10, dtype=torch.float64, device=dev)
    with pytest.raises(s512.S512Error):
        s512.topk(t, 11)
    with pytest.raises(s512.S512Error):
        s512.select(t, 10)
    assert s512.topk(t, 0).numel() == 0
