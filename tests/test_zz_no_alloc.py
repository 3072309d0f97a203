"""GPU: the allocation-free Partition512 entry point
(s512_partition_f64_dev_noalloc) against the ordinary one
(s512_partition_f64_dev, sort512.hpp:5821), through the C ABI.

Both run the same kernels, so the rearranged range and the bound must be
identical bit for bit; the fixed workspace keeps the previous call's contents,
so two calls in a row are compared, the second on a smaller range, plus the
rejection of a range beyond the workspace.
"""
import numpy as np
import pytest

from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu


def test_noalloc_partition_matches(s512, orc, dev):
    rng = np.random.default_rng(17)
    for n, pivot, left in ((6_000_011, 0.3, 20), (250_001, 0.8, 7), (1, 0.5, 0)):
        a0 = rng.random(n + left + 30)
        a0[::97] = 0.3  # keys equal to the pivot go left
        t1, t2 = to_dev(a0, dev), to_dev(a0, dev)
        b1 = s512.partition(t1, left, left + n - 1, pivot)
        b2 = s512.partition_noalloc(t2, left, left + n - 1, pivot)
        assert b1 == b2 == left + int(np.count_nonzero(a0[left:left + n] <= pivot)), n
        assert np.array_equal(to_host(t1).view(np.uint64), to_host(t2).view(np.uint64)), n


MAX_KEYS = 8_384_480  # the fixed 64 MiB workspace's limit for fp64 keys


def test_noalloc_partition_rejects_beyond_workspace(s512, dev):
    t = to_dev(np.zeros(16), dev)
    with pytest.raises(s512.S512Error, match="without allocation"):
        s512.partition_noalloc(t, 0, 20_000_000, 0.5)  # checked before any memory is touched
    b = s512.partition_noalloc(t, 0, 15, 0.5)  # the workspace still works afterwards
    assert b == 16


def test_noalloc_partition_at_the_limit(s512, dev):
    rng = np.random.default_rng(5)
    a0 = rng.random(MAX_KEYS + 1)
    t = to_dev(a0, dev)
    with pytest.raises(s512.S512Error, match=f"at most {MAX_KEYS} keys"):
        s512.partition_noalloc(t, 0, MAX_KEYS, 0.25)  # one key more than the workspace holds
    assert np.array_equal(to_host(t), a0)  # nothing changed
    t2 = to_dev(a0, dev)
    b1 = s512.partition(t, 0, MAX_KEYS - 1, 0.25)
    b2 = s512.partition_noalloc(t2, 0, MAX_KEYS - 1, 0.25)
    assert b1 == b2 == int(np.count_nonzero(a0[:MAX_KEYS] <= 0.25))
    assert np.array_equal(to_host(t).view(np.uint64), to_host(t2).view(np.uint64))
