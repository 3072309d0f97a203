"""GPU tests of key/value sorts with 64-bit keys and of argsort (no reference
counterpart).

- kv with float64 / int64 / uint64 keys and int32 values: the keys come out
  bit for bit as the keys-only sort (itself checked against the oracle) and
  the (key bits, value) multiset is kept -- the reference's kv contract
  (sort512kv.hpp:6132: values follow keys, equal keys in any order).
- argsort: a permutation of 0..n-1 whose gather of the keys equals the
  keys-only sort bit for bit, the keys unchanged; every combination of host
  and device memory for keys and positions.
"""
import numpy as np
import pytest

from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 100, 5000, 70_000, 1_000_003, 5_000_001]


def _keys(orc, kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "f64":
        a = orc.gen_f64("normal", n, seed)
        if n > 10:  # ties, signed zeros, infinities
            a[rng.integers(0, n, n // 10)] = 0.5
            a[rng.integers(0, n, 3)] = -0.0
            a[rng.integers(0, n, 3)] = 0.0
            a[rng.integers(0, n, 2)] = np.inf
            a[rng.integers(0, n, 2)] = -np.inf
        return a
    if kind == "i64":
        return rng.integers(-(1 << 63), (1 << 63) - 1, size=n, dtype=np.int64) // rng.integers(1, 1 << 20, size=n)
    if kind == "u64":
        return rng.integers(0, (1 << 64) - 1, size=n, dtype=np.uint64, endpoint=True) >> np.uint64(rng.integers(0, 40))
    if kind == "i32":
        return (orc.gen_i32_full(n, seed) % 100_000).astype(np.int32)
    raise ValueError(kind)


def _bits(a):
    return a.view(np.uint64) if a.dtype == np.float64 else a.astype(np.uint64) if a.dtype == np.int32 else a.view(np.uint64)


def _keys_sorted(s512, a):
    b = a.copy()
    s512.sort(b)
    return b


def _pairs(k, v):
    kb = _bits(k)
    o = np.lexsort((v, kb))
    return kb[o], v[o]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("kind", ["f64", "i64", "u64"])
def test_kv_wide_keys(s512, orc, dev, kind, n):
    k = _keys(orc, kind, n, 3)
    v = np.arange(n, dtype=np.int32)
    tk, tv = to_dev(k, dev), to_dev(v, dev)
    s512.kv_sort(tk, tv)
    gk, gv = to_host(tk), to_host(tv)
    assert np.array_equal(_bits(gk), _bits(_keys_sorted(s512, k)))
    pk, pv = _pairs(gk, gv)
    qk, qv = _pairs(k, v)
    assert np.array_equal(pk, qk) and np.array_equal(pv, qv)


@pytest.mark.parametrize("kind", ["f64", "i64"])
def test_kv_wide_keys_host(s512, orc, kind):
    n = 6_000_007
    k = _keys(orc, kind, n, 5)
    v = np.arange(n, dtype=np.int32)
    k0 = k.copy()
    s512.kv_sort(k, v)
    assert np.array_equal(_bits(k), _bits(_keys_sorted(s512, k0)))
    assert np.array_equal(_bits(k0[v]), _bits(k))  # each value still names its key's position


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("kind", ["i32", "f64", "i64", "u64"])
def test_argsort_device(s512, orc, dev, kind, n):
    k = _keys(orc, kind, n, 7)
    tk = to_dev(k, dev)
    idx = to_host(s512.argsort(tk))
    assert np.array_equal(to_host(tk), k)  # keys unchanged
    assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.int32))
    assert np.array_equal(_bits(k[idx]), _bits(_keys_sorted(s512, k)))


@pytest.mark.parametrize("where", ["host-host", "dev-host", "host-dev"])
def test_argsort_host_and_mixed(s512, orc, dev, where):
    import torch
    n = 3_000_001
    k = _keys(orc, "f64", n, 9)
    keys = k.copy() if where.startswith("host") else to_dev(k, dev)
    out = np.empty(n, dtype=np.int32) if where.endswith("host") else torch.empty(n, dtype=torch.int32, device=dev)
    s512.argsort(keys, out=out)
    idx = out if isinstance(out, np.ndarray) else to_host(out)
    assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.int32))
    assert np.array_equal(_bits(k[idx]), _bits(_keys_sorted(s512, k)))


def test_argsort_64m_properties(s512, dev):
    import torch
    n = 64 << 20
    t = torch.empty(n, dtype=torch.float64, device=dev)
    s512.fill(t, "normal", seed=2)
    idx = s512.argsort(t)
    g = t[idx.long()]
    assert bool(torch.all(g[1:] >= g[:-1]).item())
    assert int(idx.to(torch.int64).sum().item()) == n * (n - 1) // 2
    assert bool(torch.equal(torch.sort(idx).values, torch.arange(n, dtype=torch.int32, device=dev)))


def _pairs64(k, v):
    kb = _bits(k)
    vb = v.view(np.uint64)
    o = np.lexsort((vb, kb))
    return kb[o], vb[o]


@pytest.mark.parametrize("n", [2, 5000, 1_000_003, 5_000_001])
@pytest.mark.parametrize("kind", ["i32", "f64", "i64", "u64"])
def test_kv_8byte_values(s512, orc, dev, kind, n):
    """8-byte values (arbitrary bit patterns) follow their keys."""
    rng = np.random.default_rng(12)
    k = _keys(orc, kind, n, 13)
    v = rng.integers(-(1 << 63), (1 << 63) - 1, size=n, dtype=np.int64, endpoint=True)
    tk, tv = to_dev(k, dev), to_dev(v, dev)
    s512.kv_sort(tk, tv)
    gk, gv = to_host(tk), to_host(tv)
    assert np.array_equal(_bits(gk), _bits(_keys_sorted(s512, k)))
    pk, pv = _pairs64(gk, gv)
    qk, qv = _pairs64(k, v)
    assert np.array_equal(pk, qk) and np.array_equal(pv, qv)


def test_kv_8byte_values_host_and_mixed(s512, orc, dev):
    n = 7_000_001
    k = _keys(orc, "f64", n, 14)
    v = np.arange(n, dtype=np.float64) * 0.5  # float64 payload passed as 8-byte values
    k0, v0 = k.copy(), v.copy()
    s512.kv_sort(k, v)  # both on the host
    assert np.array_equal(_bits(k), _bits(_keys_sorted(s512, k0)))
    assert np.array_equal(_bits(k0[(v * 2).astype(np.int64)]), _bits(k))
    k = k0.copy()
    tv = to_dev(v0, dev)
    s512.kv_sort(k, tv)  # keys on the host, values on the device
    gv = to_host(tv)
    assert np.array_equal(_bits(k0[(gv * 2).astype(np.int64)]), _bits(k))
