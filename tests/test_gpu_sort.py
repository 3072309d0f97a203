"""GPU parity: Sort512::Sort / Sort512kv::Sort through the C ABI vs the oracle.

Reference behaviour pinned (sort512.hpp:5960-5963, sort512kv.hpp:6132-6135):
the output is the unique ascending permutation of the input; key/value pairs
stay together.  The oracle's restated quicksort and its closed-form qsort
both produce it; sizes here are ones the oracle finishes in seconds.
"""
import numpy as np
import pytest

from gpu_util import assert_keys_equal, to_dev, to_host

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 2, 3, 17, 255, 256, 257, 1000, 4096, 8191, 8192, 8193, 16384, 16385,
         65536, 100_000, 262_145, 1 << 20, 3_000_001]


def _inputs(orc, dtype, kind, n, seed):
    if dtype == np.int32:
        if kind == "uniform":
            return orc.gen_i32_full(n, seed)
        if kind == "small":
            return (orc.gen_i32_full(n, seed) & 0xFF).astype(np.int32)
        if kind == "sorted":
            return np.arange(n, dtype=np.int32) - n // 2
        if kind == "reversed":
            return (n // 2 - np.arange(n, dtype=np.int64)).astype(np.int32)
        if kind == "equal":
            return np.full(n, 7, dtype=np.int32)
        if kind == "extremes":
            a = orc.gen_i32_full(n, seed)
            a[::3] = np.iinfo(np.int32).max
            a[1::5] = np.iinfo(np.int32).min
            return a
    else:
        if kind in ("uniform", "normal", "distinct", "sorted", "reversed"):
            return orc.gen_f64(kind, n, seed)
        if kind == "equal":
            return np.full(n, -2.5)
        if kind == "zeros":
            a = orc.gen_f64("normal", n, seed)
            a[::2] = 0.0
            a[1::4] = -0.0
            return a
        if kind == "extremes":
            a = orc.gen_f64("normal", n, seed) * 1e300
            a[::7] = np.finfo(np.float64).max
            a[1::7] = -np.finfo(np.float64).max
            a[2::11] = np.finfo(np.float64).tiny / 4
            return a
    raise ValueError(kind)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_sort_uniform_sizes(s512, orc, dev, n, dtype):
    a = _inputs(orc, dtype, "uniform", n, seed=11 + n)
    want = orc.closed_sort(a.copy())
    t = to_dev(a, dev)
    s512.sort(t)
    assert_keys_equal(to_host(t), want, f"n={n} {np.dtype(dtype).name}")


@pytest.mark.parametrize("kind", ["small", "sorted", "reversed", "equal", "extremes"])
@pytest.mark.parametrize("n", [300, 9000, 200_000, 2_000_000])
def test_sort_i32_distributions(s512, orc, dev, kind, n):
    a = _inputs(orc, np.int32, kind, n, seed=5)
    want = orc.closed_sort(a.copy())
    t = to_dev(a, dev)
    s512.sort(t)
    assert_keys_equal(to_host(t), want, f"{kind} n={n}")


@pytest.mark.parametrize("kind", ["normal", "distinct", "sorted", "reversed", "equal", "zeros", "extremes"])
@pytest.mark.parametrize("n", [300, 9000, 200_000, 2_000_000])
def test_sort_f64_distributions(s512, orc, dev, kind, n):
    a = _inputs(orc, np.float64, kind, n, seed=6)
    want = orc.closed_sort(a.copy())
    t = to_dev(a, dev)
    s512.sort(t)
    assert_keys_equal(to_host(t), want, f"{kind} n={n}")


def test_config1_against_restated_reference(s512, orc, dev):
    """1,048,576 int32 keys, full int32 range, seed 0: bit-exact vs the
    oracle's restatement of Sort512::Sort (quicksort + bitonic leaves)."""
    a = orc.gen_i32_full(1 << 20, 0)
    want = orc.sort(a.copy())
    assert np.array_equal(want, np.sort(a))
    t = to_dev(a, dev)
    s512.sort(t)
    assert_keys_equal(to_host(t), want, "config 1")


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_sort_host_pointer(s512, orc, dtype):
    """Plain entry point with host memory: staged through the GPU."""
    a = _inputs(orc, dtype, "uniform", 123_457, seed=3)
    want = orc.closed_sort(a.copy())
    s512.sort(a)
    assert_keys_equal(a, want, "host pointer")


# ---------------------------------------------------------------- key/value
def _kv_check(orc, k0, v0, k, v, what):
    want_k = orc.closed_sort(k0.copy())
    assert_keys_equal(k, want_k, what + " keys")
    assert np.array_equal(orc.pair_multiset(k, v), orc.pair_multiset(k0, v0)), what + " pairs"


@pytest.mark.parametrize("n", [0, 1, 2, 100, 256, 257, 5000, 8192, 8193, 70_000, 1 << 20])
def test_kv_sort_reference_invariant(s512, orc, dev, n):
    """sort512test.cpp:1409-1427: values = key*100+1 must still match."""
    orc.srand48(0)
    k0 = orc.rand_vec(n, np.int32)
    v0 = (k0.astype(np.int64) * 100 + 1).astype(np.int32)
    tk, tv = to_dev(k0, dev), to_dev(v0, dev)
    s512.kv_sort(tk, tv)
    k, v = to_host(tk), to_host(tv)
    _kv_check(orc, k0, v0, k, v, f"kv n={n}")
    assert np.array_equal(v, (k.astype(np.int64) * 100 + 1).astype(np.int32))


@pytest.mark.parametrize("kind", ["uniform", "small", "equal", "extremes", "sorted"])
@pytest.mark.parametrize("n", [300, 20_000, 1_500_000])
def test_kv_sort_distributions(s512, orc, dev, kind, n):
    k0 = _inputs(orc, np.int32, kind, n, seed=9)
    v0 = np.arange(n, dtype=np.int32)
    tk, tv = to_dev(k0, dev), to_dev(v0, dev)
    s512.kv_sort(tk, tv)
    _kv_check(orc, k0, v0, to_host(tk), to_host(tv), f"kv {kind} n={n}")


def test_kv_intmax_keys_keep_their_values(s512, orc, dev):
    """Real INT_MAX keys in short leaves keep their own values (the reference
    pads leaves with (INT_MAX, INT_MAX) pairs and can emit a pad instead)."""
    rng = np.random.default_rng(4)
    for n in (5, 17, 100, 255):
        k0 = rng.integers(-100, 100, n).astype(np.int32)
        k0[rng.integers(0, n, max(1, n // 4))] = np.iinfo(np.int32).max
        v0 = np.arange(n, dtype=np.int32) + 1000
        tk, tv = to_dev(k0, dev), to_dev(v0, dev)
        s512.kv_sort(tk, tv)
        _kv_check(orc, k0, v0, to_host(tk), to_host(tv), f"INT_MAX n={n}")


def _enc_sorted_f64(a):
    """Ascending by the library's total order on doubles: IEEE order, -0.0
    before +0.0, negative NaNs first and positive NaNs last (by bit pattern)."""
    u = a.view(np.uint64)
    key = np.where((u >> np.uint64(63)) == 1, ~u, u | np.uint64(1 << 63))
    return a[np.argsort(key, kind="stable")]


@pytest.mark.parametrize("n", [100, 9000, 300_000, 2_000_000])
def test_sort_f64_nan_negzero_fallback(s512, orc, dev, n):
    """NaN and -0.0 keys: the min/max leaf kernels must not run; the encoded
    integer path gives the documented total order bit for bit."""
    a = orc.gen_f64("normal", n, 21)
    a[::97] = np.nan
    a[5::101] = -np.nan
    a[7::89] = -0.0
    a[9::83] = 0.0
    want = _enc_sorted_f64(a.copy())
    t = to_dev(a, dev)
    s512.sort(t)
    got = to_host(t)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("bits", [0x8000000000000000, 0x7FF8000000000000, 0xFFF8000000000001, 0x7FF0000000000001])
def test_sort_f64_single_special_key(s512, orc, dev, bits):
    """One -0.0, quiet NaN, negative NaN or signalling NaN among 300,000 keys
    (a partition level; the first count pass tests every key's class) is
    enough to keep the min/max leaves off: bit-exact total order.  A +0.0
    beside it makes the -0.0 case matter (-0.0 sorts first)."""
    a = orc.gen_f64("normal", 300_000, 23)
    a[98_765] = 0.0
    a.view(np.uint64)[123_457] = np.uint64(bits)
    want = _enc_sorted_f64(a.copy())
    t = to_dev(a, dev)
    st = s512.sort_ex(t)
    assert st.levels >= 1
    assert np.array_equal(to_host(t).view(np.uint64), want.view(np.uint64)), hex(bits)


@pytest.mark.parametrize("n", [50_000, 5_000_000])
def test_sort_f64_subnormal_inf_fast_path(s512, orc, dev, n):
    """No NaN / -0.0, so the v_min_f64/v_max_f64 leaves run: subnormals,
    +-inf and +0.0 must come back bit for bit."""
    a = orc.gen_f64("normal", n, 22) * 1e-310
    a[::13] = np.inf
    a[1::17] = -np.inf
    a[2::19] = 0.0
    a[3::23] = np.finfo(np.float64).max
    want = np.sort(a)
    t = to_dev(a, dev)
    s512.sort(t)
    assert np.array_equal(to_host(t).view(np.uint64), want.view(np.uint64))


def _presorted(kind, n, rng):
    """fp64 inputs in order, in reverse, or almost so."""
    if kind == "sorted":
        return np.sort(rng.standard_normal(n))
    if kind == "reversed":
        return np.sort(rng.standard_normal(n))[::-1].copy()
    if kind == "reversed_dups":  # non-increasing with runs of equal keys
        return np.repeat(np.sort(rng.integers(0, n // 7, n // 7 + 1))[::-1], 7)[:n].astype(np.float64)
    if kind == "sorted_swaps":  # sorted but for a few transpositions (no leaf may skip its network)
        a = np.sort(rng.standard_normal(n))
        i = rng.integers(0, n, 64)
        j = np.minimum(i + rng.integers(1, 50, 64), n - 1)
        a[i], a[j] = a[j].copy(), a[i].copy()
        return a
    if kind == "reversed_swaps":
        a = np.sort(rng.standard_normal(n))[::-1].copy()
        i = rng.integers(0, n, 64)
        j = np.minimum(i + rng.integers(1, 50, 64), n - 1)
        a[i], a[j] = a[j].copy(), a[i].copy()
        return a
    if kind == "sawtooth":  # ascending runs of 1000 keys
        return np.tile(np.arange(1000, dtype=np.float64), n // 1000 + 1)[:n]
    if kind == "organ":  # ascending then descending
        h = np.arange(n // 2, dtype=np.float64)
        return np.concatenate([h, h[::-1], [0.5] * (n - 2 * (n // 2))])
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["sorted", "reversed", "reversed_dups", "sorted_swaps", "reversed_swaps",
                                  "sawtooth", "organ"])
@pytest.mark.parametrize("n", [1500, 9000, 300_001, 3_000_001])
def test_sort_f64_runs_in_order(s512, dev, kind, n):
    """Inputs with order already in them, where the scatter's long runs take
    the plain scatter (the full-line scatter's gate) and the leaves come out
    nearly sorted or nearly reversed.  Against numpy's sort bit for bit: whole
    inputs in order or reversed (with and without equal keys), the same with a
    few transpositions, and inputs made of runs."""
    rng = np.random.default_rng(sum(map(ord, kind)) + n)
    a = _presorted(kind, n, rng)
    t = to_dev(a, dev)
    s512.sort(t)
    assert np.array_equal(to_host(t).view(np.uint64), np.sort(a).view(np.uint64)), kind
