"""GPU parity for the rest of the reference API through the C ABI:
Partition512 (sort512.hpp:5729/5821, sort512kv.hpp:5958), SmallSort16V
(sort512.hpp:2255/5312, sort512kv.hpp:5235), the merge-based sorts
(SortOmpMerge/MergeDeps/ParMerge, sort512.hpp:6006-6174) and the in-place
merge (parallelInplace.hpp:209-306), against the oracle.
"""
import numpy as np
import pytest

from gpu_util import assert_keys_equal, to_dev, to_host

pytestmark = pytest.mark.gpu


def _check_partition(a0, a, left, right, pivot, bound):
    seg0, seg = a0[left:right + 1], a[left:right + 1]
    assert bound == left + int(np.count_nonzero(seg0 <= pivot))
    assert np.all(a[left:bound] <= pivot) and not np.any(a[bound:right + 1] <= pivot)
    assert np.array_equal(a[:left], a0[:left]) and np.array_equal(a[right + 1:], a0[right + 1:])
    assert np.array_equal(np.sort(seg), np.sort(seg0))


def test_partition_reference_sweep(s512, orc, dev):
    """The reference test sweep: every n in 1..1000 (and 2^k up to 1024),
    pivot n/2, drand48 input (sort512test.cpp:1451-1483); bound equals the
    oracle's restated Partition512."""
    for dt in (np.int32, np.float64):
        orc.srand48(0)
        for n in list(range(1, 1001)) + [1 << k for k in range(11)]:
            a0 = orc.rand_vec(n, dt)
            pivot = dt(n // 2)
            want = orc.partition(a0.copy(), 0, n - 1, int(pivot) if dt == np.int32 else float(pivot))
            t = to_dev(a0, dev)
            b = s512.partition(t, 0, n - 1, pivot)
            assert b == want, (dt, n)
            _check_partition(a0, to_host(t), 0, n - 1, pivot, b)


@pytest.mark.parametrize("n", [5_000, 100_000, 3_000_017])
def test_partition_large_subrange(s512, orc, dev, n):
    rng = np.random.default_rng(n)
    for dt in (np.int32, np.float64):
        a0 = (rng.integers(-10**6, 10**6, n)).astype(dt)
        left, right = n // 7, n - n // 5
        pivot = dt(a0[n // 2])
        t = to_dev(a0, dev)
        b = s512.partition(t, left, right, pivot)
        a = to_host(t)
        _check_partition(a0, a, left, right, pivot, b)
        # stable on both sides
        seg0 = a0[left:right + 1]
        assert np.array_equal(a[left:b], seg0[seg0 <= pivot])
        assert np.array_equal(a[b:right + 1], seg0[seg0 > pivot])


@pytest.mark.parametrize("low_share", [0.02, 0.3, 0.5, 0.7, 0.98])
@pytest.mark.parametrize("dt", [np.int32, np.float64])
def test_partition_both_directions(s512, orc, dev, low_share, dt):
    """The larger side stays in place (a pass from the left when most keys are
    <= pivot, from the right otherwise), the smaller one goes through
    scratch: bound, both sides, stability and the untouched outside."""
    n = 2_000_003
    rng = np.random.default_rng(int(low_share * 1000))
    a0 = rng.integers(0, 10**6, n + 200).astype(dt)
    pivot = dt(int(low_share * 10**6))
    left, right = 100, n + 99
    t = to_dev(a0, dev)
    b = s512.partition(t, left, right, pivot)
    a = to_host(t)
    _check_partition(a0, a, left, right, pivot, b)
    seg0 = a0[left:right + 1]
    assert np.array_equal(a[left:b], seg0[seg0 <= pivot])
    assert np.array_equal(a[b:right + 1], seg0[seg0 > pivot])


@pytest.mark.parametrize("low_at_samples", [True, False])
def test_partition_misleading_sample(s512, orc, dev, low_at_samples):
    """The direction comes from 4096 evenly spaced keys; an input whose keys at
    exactly those places belong to the smaller side sends the larger side
    through scratch.  The result must not depend on it."""
    n = 1 << 22
    a0 = np.full(n, 5.0 if low_at_samples else -5.0)
    idx = (np.arange(4096, dtype=np.uint64) * np.uint64(n) // np.uint64(4096)).astype(np.int64)
    a0[idx] = -5.0 if low_at_samples else 5.0
    t = to_dev(a0, dev)
    b = s512.partition(t, 0, n - 1, 0.0)
    a = to_host(t)
    _check_partition(a0, a, 0, n - 1, 0.0, b)


def test_partition_kv_multichunk(s512, orc, dev):
    rng = np.random.default_rng(8)
    n = 3_000_001
    k0 = rng.integers(-1000, 1000, n).astype(np.int32)
    v0 = np.arange(n, dtype=np.int32)
    for pivot in (-900, 0, 900):
        tk, tv = to_dev(k0, dev), to_dev(v0, dev)
        b = s512.kv_partition(tk, tv, 0, n - 1, pivot)
        k, v = to_host(tk), to_host(tv)
        _check_partition(k0, k, 0, n - 1, pivot, b)
        lo = k0 <= pivot
        assert np.array_equal(v[:b], v0[lo]) and np.array_equal(v[b:], v0[~lo])  # stable pairs


def test_partition_full_size_f64(s512, orc, dev):
    """268,435,456 fp64 keys, one pivot (the median of a uniform input): the
    bound by count, each side on its own side, the multiset by hash."""
    import torch
    n = 1 << 28
    t = torch.empty(n, dtype=torch.float64, device=dev)
    s512.fill(t, "uniform", seed=3)
    pivot = 0.5
    want = int((t <= pivot).sum().item())
    h0 = int((t.view(torch.int64) * -7046029254386353131).sum().item())
    b = s512.partition(t, 0, n - 1, pivot)
    assert b == want
    assert bool((t[:b] <= pivot).all().item()) and not bool((t[b:] <= pivot).any().item())
    assert int((t.view(torch.int64) * -7046029254386353131).sum().item()) == h0


def test_partition_edge_cases(s512, orc, dev):
    a0 = np.arange(10, dtype=np.int32)
    t = to_dev(a0, dev)
    assert s512.partition(t, 5, 4, 3) == 5           # empty range
    assert s512.partition(t, 0, 9, -100) == 0         # nothing <= pivot
    assert s512.partition(t, 0, 9, 100) == 10         # everything <= pivot
    f = np.array([1.0, np.nan, -0.0, 0.0, 2.0, -np.inf, np.inf])
    tf = to_dev(f, dev)
    b = s512.partition(tf, 0, f.size - 1, 0.0)       # NaN is never <= pivot
    assert b == orc.partition(f.copy(), 0, f.size - 1, 0.0) == 3


def test_partition_host_pointer(s512, orc):
    rng = np.random.default_rng(1)
    a0 = rng.integers(-50, 50, 20_000).astype(np.int32)
    a = a0.copy()
    b = s512.partition(a, 100, 19_000, 7)
    _check_partition(a0, a, 100, 19_000, 7, b)


def test_kv_partition(s512, orc, dev):
    orc.srand48(0)
    for n in list(range(1, 400, 7)) + [4096, 100_003]:
        k0 = orc.rand_vec(n, np.int32)
        v0 = (k0.astype(np.int64) * 100 + 1).astype(np.int32)
        pivot = n // 2
        want = orc.kv_partition(k0.copy(), v0.copy(), 0, n - 1, pivot)
        tk, tv = to_dev(k0, dev), to_dev(v0, dev)
        b = s512.kv_partition(tk, tv, 0, n - 1, pivot)
        k, v = to_host(tk), to_host(tv)
        assert b == want
        _check_partition(k0, k, 0, n - 1, pivot, b)
        assert np.array_equal(v, (k.astype(np.int64) * 100 + 1).astype(np.int32))


def test_smallsort_every_length(s512, orc, dev):
    """Every length the reference tests (1..256 int, 1..128 double), 5
    drand48 arrays each; output equals the oracle's SmallSort16V."""
    for dt, maxn in ((np.int32, 256), (np.float64, 128)):
        orc.srand48(0)
        for n in range(1, maxn + 1):
            for _ in range(5):
                a0 = orc.rand_vec(n, dt)
                want = orc.smallsort(a0.copy())
                t = to_dev(a0, dev)
                s512.smallsort16v(t)
                assert_keys_equal(to_host(t), want, f"{dt} n={n}")


def test_smallsort_rejects_long(s512):
    with pytest.raises(s512.S512Error):
        s512.smallsort16v(np.zeros(257, dtype=np.int32))
    with pytest.raises(s512.S512Error):
        s512.smallsort16v(np.zeros(129, dtype=np.float64))


def test_kv_smallsort_every_length(s512, orc, dev):
    orc.srand48(0)
    for n in range(1, 257):
        k0 = orc.rand_vec(n, np.int32)
        v0 = (k0.astype(np.int64) * 100 + 1).astype(np.int32)
        wk, _ = orc.kv_smallsort(k0.copy(), v0.copy())
        tk, tv = to_dev(k0, dev), to_dev(v0, dev)
        s512.kv_smallsort16v(tk, tv)
        k, v = to_host(tk), to_host(tv)
        assert np.array_equal(k, wk), n
        assert np.array_equal(v, (k.astype(np.int64) * 100 + 1).astype(np.int32)), n


@pytest.mark.parametrize("n", [1, 2, 1000, 1024, 16384, 16385, 100_000, 1 << 20, 2_500_001])
@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_sort_merge(s512, orc, dev, n, dtype):
    """SortOmpMerge-family drop-in: same keys as the restated Sort512::Sort."""
    a0 = orc.gen_i32_full(n, n) if dtype == np.int32 else orc.gen_f64("normal", n, n)
    want = orc.closed_sort(a0.copy())
    t = to_dev(a0, dev)
    s512.sort_merge(t)
    assert_keys_equal(to_host(t), want, f"merge sort n={n}")


@pytest.mark.parametrize("na,nb", [(1, 1), (1, 5000), (5000, 1), (2048, 2048), (3000, 7001),
                                   (1 << 19, 1 << 20), (1_000_003, 999_999)])
@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_merge_inplace(s512, orc, dev, na, nb, dtype):
    """parallelMergeInPlace: merge of two sorted runs equals the oracle's
    stable merge, bit for bit (heavy duplicates included)."""
    rng = np.random.default_rng(na * 7 + nb)
    A = np.sort(rng.integers(0, 1000, na)).astype(dtype)
    B = np.sort(rng.integers(0, 1000, nb)).astype(dtype)
    a0 = np.concatenate([A, B])
    want = orc.merge(a0.copy(), na)
    t = to_dev(a0, dev)
    s512.merge_inplace(t, na)
    got = to_host(t)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_merge_inplace_findmedian_cases(s512, orc, dev):
    """Small merges like the ones FindMedian splits (parallelInplace.hpp:77-134)."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        na, nb = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        a0 = np.concatenate([np.sort(rng.integers(0, 20, na)), np.sort(rng.integers(0, 20, nb))]).astype(np.int32)
        want = orc.merge(a0.copy(), na)
        t = to_dev(a0, dev)
        s512.merge_inplace(t, na)
        assert np.array_equal(to_host(t), want)


def test_dropin_headers_reference_test_pattern():
    """tests/cpp/shim_test.cpp: the reference's test pattern through the C++
    drop-in headers (host arrays), exit status 0."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "shim_test")
    assert os.path.exists(exe), "tests/cpp/shim_test not built (make -C tests/cpp)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]


def test_release_and_reuse(s512, orc, dev):
    """Cached workspaces freed between calls: every kind of call still works after."""
    import torch
    a = orc.gen_f64("normal", 3_000_001, 2)
    want = orc.closed_sort(a.copy())
    for _ in range(2):
        t = to_dev(a, dev)
        s512.sort(t)
        assert np.array_equal(to_host(t), want)
        h = a.copy()
        s512.sort(h)  # host staging
        assert np.array_equal(h, want)
        assert np.array_equal(to_host(s512.topk(to_dev(a, dev), 100)), want[:100])
        s512.release()
    torch.cuda.synchronize()
