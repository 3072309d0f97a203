"""GPU tests of the host-pointer entry points at sizes that take the pipelined
staging path: chunks of 32 MiB through a ring of three pinned buffers, with
worker threads copying between the caller's memory and the ring.

Sizes cover one partial chunk, several chunks with a ragged tail, and more
chunks than ring buffers; each result is checked against the oracle's closed
form (ascending order is unique, sort512.hpp:5960).  Key/value arrays are
staged in two slots, also with one array on the host and the other on the
device.  Partition512 and the merge stage only the addressed prefix.
"""
import numpy as np
import pytest

from gpu_util import assert_keys_equal, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [600_000, 5_000_001, 30_000_017])
@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_host_sort_pipelined(s512, orc, dtype, n):
    a = orc.gen_i32_full(n, 21) if dtype == np.int32 else orc.gen_f64("normal", n, 21)
    want = orc.closed_sort(a.copy())
    s512.sort(a)
    assert_keys_equal(a, want, f"host n={n}")
    s512.sort(a)  # a second call reuses the per-device staging buffers
    assert_keys_equal(a, want, f"host n={n}, sorted input")


def test_host_kv_pipelined(s512, orc):
    n = 20_000_003
    k = (orc.gen_i32_full(n, 4) & 0xFFFFF).astype(np.int32)
    v = np.arange(n, dtype=np.int32)
    k0, v0 = k.copy(), v.copy()
    s512.kv_sort(k, v)
    assert np.array_equal(k, orc.closed_sort(k0.copy()))
    assert np.array_equal(orc.pair_multiset(k, v), orc.pair_multiset(k0, v0))


def test_host_kv_mixed_host_and_device(s512, orc, dev):
    n = 9_000_001
    k = (orc.gen_i32_full(n, 6) & 0xFFFF).astype(np.int32)
    v = np.arange(n, dtype=np.int32)
    k0 = k.copy()
    tv = to_dev(v, dev)
    s512.kv_sort(k, tv)  # keys staged, values in place
    gv = to_host(tv)
    assert np.array_equal(k, orc.closed_sort(k0.copy()))
    assert np.array_equal(orc.pair_multiset(k, gv), orc.pair_multiset(k0, v))


def test_host_segsort_pipelined(s512, orc):
    off = s512.lognormal_offsets(40_000, seed=9)
    keys = orc.gen_f64("normal", int(off[-1]), 9)
    k0 = keys.copy()
    s512.segsort(keys, off)
    rng = np.random.default_rng(3)
    for i in rng.choice(len(off) - 1, 200, replace=False):
        assert np.array_equal(keys[off[i]:off[i + 1]], np.sort(k0[off[i]:off[i + 1]])), i


def test_host_partition_and_merge_pipelined(s512, orc):
    n = 12_000_000
    a = orc.gen_f64("uniform", n, 2)
    a0 = a.copy()
    right = n - 1000  # only [0, right] is staged; the tail must stay untouched
    bound = s512.partition(a, 0, right, 0.5)
    assert bound == int(np.count_nonzero(a0[:right + 1] <= 0.5))
    assert np.all(a[:bound] <= 0.5) and np.all(a[bound:right + 1] > 0.5)
    assert np.array_equal(a[right + 1:], a0[right + 1:])
    assert np.array_equal(np.sort(a[:right + 1]), np.sort(a0[:right + 1]))
    b = np.concatenate([np.sort(orc.gen_f64("uniform", 7_000_000, 3)), np.sort(orc.gen_f64("uniform", 6_000_001, 4))])
    want = np.sort(b)
    s512.merge_inplace(b, 7_000_000)
    assert np.array_equal(b, want)
