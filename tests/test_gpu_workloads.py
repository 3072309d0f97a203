"""GPU tests of the full-size secondary workloads and their input generator.

- The Zipf key generator (key/value workload) against the Zipf law itself:
  rank frequencies r^-1.1 / H(N, 1.1) computed here in closed form, and the
  rank-to-key map a bijection of [0, N).
- Key/value sort (Sort512kv::Sort, sort512kv.hpp) of Zipf keys: exact against
  numpy at 8M pairs; at 536,870,912 pairs by size-independent properties
  (keys non-decreasing, the (key, value) multiset kept, values a permutation).
- Plain sorts at 268,435,456 keys (fp64 five distributions, int32 four):
  non-decreasing, multiset kept.
- Learned partitioner at 1,073,741,824 fp64 keys: mixture accepted, the
  piecewise-constant input rejected (fallback), both outputs sorted
  permutations of the input.
"""
import numpy as np
import pytest

from gpu_util import to_host

pytestmark = pytest.mark.gpu


def _hash64(x):
    import torch
    x = x * -7046029254386353131
    x = x ^ (x >> 29)
    x = x * -4658895280553007687
    return int((x ^ (x >> 32)).sum().item())


def test_zipf_generator_follows_zipf_law(s512, dev):
    import torch
    N, n = 1000, 20_000_000
    t = torch.empty(n, dtype=torch.int32, device=dev)
    s512.fill(t, "zipf", seed=5, param=N)
    a = to_host(t)
    assert a.min() >= 0 and a.max() < N
    counts = np.bincount(a, minlength=N)
    assert np.count_nonzero(counts) == N  # every rank drawn; the key map is onto [0, N)
    r = np.arange(1, N + 1, dtype=np.float64)
    p = r ** -1.1 / np.sum(r ** -1.1)
    got = np.sort(counts)[::-1] / n
    # the 20 most frequent keys: within 4 standard deviations of n p_r
    sd = np.sqrt(p[:20] * (1 - p[:20]) / n)
    assert np.all(np.abs(got[:20] - p[:20]) < 4 * sd + 1e-12), (got[:5], p[:5])
    # the whole law: chi-square over ranks (df = N - 1), with the observed
    # order of frequencies taken as the rank order
    exp = p * n
    chi2 = np.sum((np.sort(counts)[::-1] - exp) ** 2 / exp)
    assert chi2 < N + 6 * np.sqrt(2 * N), chi2


def test_zipf_generator_bijection_of_ranks(s512, dev):
    """Ranks mapped to keys by a bijection: with N = 2^12 ranks and enough
    draws every key appears, and seeds give different maps."""
    import torch
    N = 1 << 12
    t = torch.empty(4_000_000, dtype=torch.int32, device=dev)
    s512.fill(t, "zipf", seed=1, param=N)
    a = to_host(t)
    c = np.bincount(a, minlength=N)
    assert np.count_nonzero(c) == N and a.max() < N
    s512.fill(t, "zipf", seed=2, param=N)
    b = to_host(t)
    assert np.argmax(np.bincount(b, minlength=N)) != np.argmax(c) or not np.array_equal(a, b)


def test_kv_zipf_exact(s512, orc, dev):
    import torch
    n = 8_000_000
    k = torch.empty(n, dtype=torch.int32, device=dev)
    s512.fill(k, "zipf", seed=3, param=1 << 24)
    v = torch.arange(n, dtype=torch.int32, device=dev)
    k0, v0 = to_host(k), to_host(v)
    s512.kv_sort(k, v)
    gk, gv = to_host(k), to_host(v)
    assert np.array_equal(gk, np.sort(k0))
    assert np.array_equal(orc.pair_multiset(gk, gv), orc.pair_multiset(k0, v0))


def test_kv_zipf_full_size(s512, dev):
    """536,870,912 pairs (the Sort512kv workload at its full size)."""
    import torch
    n = 536_870_912
    k = torch.empty(n, dtype=torch.int32, device=dev)
    s512.fill(k, "zipf", seed=3, param=1 << 24)
    v = torch.arange(n, dtype=torch.int32, device=dev)
    h0 = _hash64(k.to(torch.int64) * (1 << 32) + v.to(torch.int64))
    s512.kv_sort(k, v)
    assert bool(torch.all(k[1:] >= k[:-1]).item())
    assert _hash64(k.to(torch.int64) * (1 << 32) + v.to(torch.int64)) == h0
    # values: still a permutation of 0..n-1
    assert int(v.to(torch.int64).sum().item()) == n * (n - 1) // 2
    assert _hash64(v.to(torch.int64)) == _hash64(torch.arange(n, dtype=torch.int64, device=dev))


@pytest.mark.parametrize("kind", ["mixture", "piecewise"])
def test_learned_1g(s512, dev, kind):
    import torch
    n = 1 << 30
    t = torch.empty(n, dtype=torch.float64, device=dev)
    s512.fill(t, kind, seed=7)
    h0 = _hash64(t.view(torch.int64))
    st = s512.sort_learned(t)
    assert bool(torch.all(t[1:] >= t[:-1]).item())
    assert _hash64(t.view(torch.int64)) == h0
    if kind == "piecewise":
        assert st.learned_fallback == 1 and st.classifier == 0, st.as_dict()
    else:
        assert st.learned_fallback == 0 and st.classifier == 1, st.as_dict()


@pytest.mark.parametrize("kind", ["uniform", "normal", "distinct", "sorted", "reversed"])
def test_plain_f64_full_size(s512, dev, kind):
    """268,435,456 fp64 keys (the benchmark workload): non-decreasing and the
    same multiset of bit patterns as the input."""
    import torch
    n = 268_435_456
    t = torch.empty(n, dtype=torch.float64, device=dev)
    s512.fill(t, kind, seed=0, param=1000)
    h0 = _hash64(t.view(torch.int64))
    st = s512.sort_ex(t)
    assert st.levels >= 1
    assert bool(torch.all(t[1:] >= t[:-1]).item())
    assert _hash64(t.view(torch.int64)) == h0


@pytest.mark.parametrize("kind", ["uniform", "distinct", "sorted", "reversed"])
def test_plain_i32_full_size(s512, dev, kind):
    """268,435,456 int32 keys over the full range."""
    import torch
    n = 268_435_456
    t = torch.empty(n, dtype=torch.int32, device=dev)
    s512.fill(t, kind, seed=0, param=1000)
    h0 = _hash64(t.to(torch.int64))
    s512.sort(t)
    assert bool(torch.all(t[1:] >= t[:-1]).item())
    assert _hash64(t.to(torch.int64)) == h0
