"""GPU parity of the one-pass partition levels (s512_onepass.hpp) against
the oracle's closed form.

The one-pass engine must give the same result as the reference's sort
(sort512.hpp:5960-5963: the unique ascending permutation; sort512kv.hpp:6132
for pairs: keys ascending, each value with its key).  Cases cover both root
modes (exact: every chunk resident; gapped: sample-sized regions), the
second level, the hand-off of oversized buckets to the multi-level engine,
the region overflow fallback, every key type and the duplicate-heavy,
sorted, signed-zero and extreme inputs the reference's tests use.
"""
import numpy as np
import pytest

from gpu_util import assert_keys_equal, to_dev, to_host
from test_gpu_sort import _inputs

pytestmark = pytest.mark.gpu


def _opts(s512, **kw):
    o = s512.Options()
    o.engine = kw.pop("engine", s512.ENGINE_ONE_PASS)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _sort(s512, dev, a, **kw):
    t = to_dev(a, dev)
    st = s512.sort_ex(t, _opts(s512, **kw))
    return to_host(t), st


# exact root (n <= resident chunks x 8192), gapped root above it; the
# smallest size is just above the largest tile (one tile needs no level)
SIZES = [16_385, 100_003, 1 << 20, 4_000_037, 9_000_000]


@pytest.mark.parametrize("kind", ["uniform", "normal", "distinct", "sorted", "reversed", "equal", "zeros", "extremes"])
@pytest.mark.parametrize("n", SIZES)
def test_onepass_f64(s512, orc, dev, kind, n):
    a = _inputs(orc, np.float64, kind, n, seed=31)
    got, st = _sort(s512, dev, a)
    assert_keys_equal(got, orc.closed_sort(a.copy()), f"one-pass f64 {kind} n={n}")
    assert st.engine == s512.ENGINE_ONE_PASS and st.levels >= 1


@pytest.mark.parametrize("kind", ["uniform", "small", "sorted", "reversed", "equal", "extremes"])
@pytest.mark.parametrize("n", SIZES)
def test_onepass_i32(s512, orc, dev, kind, n):
    a = _inputs(orc, np.int32, kind, n, seed=32)
    got, st = _sort(s512, dev, a)
    assert_keys_equal(got, orc.closed_sort(a.copy()), f"one-pass i32 {kind} n={n}")
    assert st.engine == s512.ENGINE_ONE_PASS


@pytest.mark.parametrize("kind", ["uniform", "small", "equal", "extremes", "sorted"])
@pytest.mark.parametrize("n", [100_000, 3_000_017, 6_000_000])
def test_onepass_kv(s512, orc, dev, kind, n):
    k0 = _inputs(orc, np.int32, kind, n, seed=33)
    v0 = np.arange(n, dtype=np.int32)
    tk, tv = to_dev(k0, dev), to_dev(v0, dev)
    st = s512.kv_sort_ex(tk, tv, _opts(s512))
    k, v = to_host(tk), to_host(tv)
    assert_keys_equal(k, orc.closed_sort(k0.copy()), f"kv {kind} n={n}")
    assert np.array_equal(orc.pair_multiset(k, v), orc.pair_multiset(k0, v0)), f"kv pairs {kind} n={n}"
    assert st.engine == s512.ENGINE_ONE_PASS


@pytest.mark.parametrize("dtype", [np.int64, np.uint64, np.uint32])
@pytest.mark.parametrize("n", [500_000, 5_000_000])
def test_onepass_wide_keys(s512, dev, dtype, n):
    rng = np.random.default_rng(34)
    info = np.iinfo(dtype)
    a = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    a[::101] = info.max
    a[1::103] = info.min
    got, st = _sort(s512, dev, a)
    assert np.array_equal(got, np.sort(a)), f"{np.dtype(dtype).name} n={n}"
    assert st.engine == s512.ENGINE_ONE_PASS


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_onepass_region_overflow_falls_back(s512, orc, dev, dtype):
    """Regions with almost no margin overflow; the root level leaves the
    input as it was and the two-pass levels sort it."""
    n = 5_000_011
    a = _inputs(orc, dtype, "uniform" if dtype == np.int32 else "normal", n, seed=35)
    got, st = _sort(s512, dev, a, region_slack=1)
    assert_keys_equal(got, orc.closed_sort(a.copy()), "overflow fallback")
    assert st.region_overflow == 1 and st.engine == s512.ENGINE_TWO_PASS


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
@pytest.mark.parametrize("logf,n", [(2, 5_000_000), (1, 9_000_000), (3, 4_100_000)])
def test_onepass_leftovers_to_multilevel(s512, orc, dev, dtype, logf, n):
    """Small fan-outs leave buckets above a tile after two one-pass levels
    (finished by the multi-level engine) and, with fan-out 2 on 9M keys,
    second-level segments too large for the exact mode (moved, then the
    multi-level engine)."""
    a = _inputs(orc, dtype, "uniform", n, seed=36 + logf)
    got, st = _sort(s512, dev, a, max_fanout_log2=logf)
    assert_keys_equal(got, orc.closed_sort(a.copy()), f"fan-out 2^{logf} n={n}")
    assert st.engine == s512.ENGINE_ONE_PASS and st.levels > 2


@pytest.mark.parametrize("n", [1 << 20, 6_000_000])
def test_onepass_nan_negzero(s512, orc, dev, n):
    """NaN and -0.0 seen by the root level: the leaves run the encoded
    integer network; the documented total order bit for bit."""
    from test_gpu_sort import _enc_sorted_f64
    a = orc.gen_f64("normal", n, 37)
    a[::97] = np.nan
    a[5::101] = -np.nan
    a[7::89] = -0.0
    want = _enc_sorted_f64(a.copy())
    got, _ = _sort(s512, dev, a)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_engines_agree(s512, orc, dev, dtype):
    """The same input through the two-pass and the one-pass engine: identical
    bits (the sorted order is unique)."""
    n = 7_340_033
    a = _inputs(orc, dtype, "uniform" if dtype == np.int32 else "normal", n, seed=38)
    one, st1 = _sort(s512, dev, a)
    two, st2 = _sort(s512, dev, a, engine=s512.ENGINE_TWO_PASS)
    assert st1.engine == s512.ENGINE_ONE_PASS and st2.engine == s512.ENGINE_TWO_PASS
    assert np.array_equal(one.view(np.uint8), two.view(np.uint8))
    assert_keys_equal(one, orc.closed_sort(a.copy()), "one-pass vs closed form")


def test_onepass_repeated_calls_reuse_workspace(s512, orc, dev):
    """Back-to-back calls of different sizes and modes on one device (the
    look-back words and tickets of each call start from zero)."""
    for n, kind in [(6_000_000, "normal"), (300_000, "uniform"), (4_500_000, "distinct"), (1 << 20, "sorted"),
                    (6_000_000, "reversed")]:
        a = _inputs(orc, np.float64, kind, n, seed=39)
        got, _ = _sort(s512, dev, a)
        assert_keys_equal(got, orc.closed_sort(a.copy()), f"{kind} n={n}")
