"""GPU parity of every kernel shape the tuning options can select.

The count, scatter and tile-sort kernels come in several shapes (workgroup
threads, keys per thread, where the bucket of a key comes from); the
autotuner picks among them per dtype and size class.  The output must not
depend on the choice: each shape is checked against the oracle's closed form
(ascending order is unique, sort512.hpp:5960) on inputs that reach every tile
class and both partition levels, plus a skewed and a duplicate-heavy input
that switch on the classifier's sub-tables and equality buckets.
"""
import numpy as np
import pytest

from gpu_util import assert_keys_equal, to_dev, to_host

pytestmark = pytest.mark.gpu

N = 3_000_001


def _input(orc, dtype, kind):
    if dtype == np.int32:
        a = orc.gen_i32_full(N, 11)
        if kind == "distinct":
            a = (a % 1000).astype(np.int32)
        elif kind == "sorted":
            a = np.sort(a)
        return a
    return orc.gen_f64(kind, N, 11)


def _run(s512, dev, a, **opts):
    t = to_dev(a, dev)
    o = s512.Options()
    for k, v in opts.items():
        setattr(o, k, v)
    st = s512.sort_ex(t, o)
    return to_host(t), st


@pytest.mark.parametrize("kind", ["normal", "distinct", "sorted"])
@pytest.mark.parametrize("dtype", [np.int32, np.float64])
@pytest.mark.parametrize("variant", range(1, 12))
def test_scatter_variants(s512, orc, dev, dtype, kind, variant):
    a = _input(orc, dtype, kind)
    got, st = _run(s512, dev, a, scatter_variant=variant)
    assert_keys_equal(got, orc.closed_sort(a.copy()), f"scatter_variant={variant} {kind}")
    assert st.levels >= 1


@pytest.mark.parametrize("offset", [0, 1, 3])
@pytest.mark.parametrize("dtype", [np.int32, np.float64])
@pytest.mark.parametrize("variant", [9, 10, 11])
def test_line_buffered_scatter_unaligned(s512, orc, dev, dtype, variant, offset):
    """The line-buffered and full-line scatters find line boundaries from the
    destination's address: arrays that start 1 or 3 keys past a 128-byte boundary (and the
    scratch array, which is aligned) must sort the same; 9,437,187 keys give
    many chunks per level, two levels and a ragged last round."""
    import torch
    n = 9_437_187
    a = orc.gen_i32_full(n, 5) if dtype == np.int32 else orc.gen_f64("normal", n, 5)
    base = torch.empty(n + 8, dtype=torch.int32 if dtype == np.int32 else torch.float64, device=dev)
    t = base[offset:offset + n]
    t.copy_(torch.from_numpy(a))
    o = s512.Options()
    o.scatter_variant = variant
    st = s512.sort_ex(t, o)
    assert st.levels == 2
    assert_keys_equal(to_host(t), orc.closed_sort(a.copy()), f"scatter_variant={variant} offset={offset}")


@pytest.mark.parametrize("variant", [9, 10, 11])
def test_line_buffered_scatter_wide_and_segmented(s512, dev, variant):
    """int64 keys (8-byte lines of 16 keys) and a segmented batch through the
    line-buffered and full-line scatters, against numpy's sort (the ascending order is
    unique)."""
    import torch
    rng = np.random.default_rng(8)
    k = rng.integers(-(1 << 62), 1 << 62, 2_500_003, dtype=np.int64)
    t = to_dev(k, dev)
    o = s512.Options()
    o.scatter_variant = variant
    s512.sort_ex(t, o)
    assert np.array_equal(to_host(t), np.sort(k))
    sizes = np.array([1_200_000, 3, 800_001, 0, 2_000_000], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    a = rng.standard_normal(int(off[-1]))
    ta = to_dev(a, dev)
    toff = torch.from_numpy(off).to(dev)
    s512.segsort(ta, toff, options=o)
    want = np.concatenate([np.sort(a[off[i]:off[i + 1]]) for i in range(len(sizes))])
    assert np.array_equal(to_host(ta).view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("kind", ["normal", "distinct"])
@pytest.mark.parametrize("dtype", [np.int32, np.float64])
@pytest.mark.parametrize("variant", range(1, 6))
def test_count_variants(s512, orc, dev, dtype, kind, variant):
    a = _input(orc, dtype, kind)
    got, _ = _run(s512, dev, a, count_variant=variant)
    assert_keys_equal(got, orc.closed_sort(a.copy()), f"count_variant={variant} {kind}")


@pytest.mark.parametrize("n", [200, 700, 1500, 3000, 6000, 12000, 100_000, N])
@pytest.mark.parametrize("dtype", [np.int32, np.float64])
@pytest.mark.parametrize("variant", range(1, 4))
def test_tile_variants(s512, orc, dev, dtype, n, variant):
    a = _input(orc, dtype, "normal")[:n].copy()
    got, _ = _run(s512, dev, a, tile_variant=variant)
    assert_keys_equal(got, orc.closed_sort(a.copy()), f"tile_variant={variant} n={n}")


@pytest.mark.parametrize("variant", range(1, 4))
def test_tile_variants_kv(s512, orc, dev, variant):
    k = (orc.gen_i32_full(N, 5) & 0xFFFFF).astype(np.int32)
    v = np.arange(N, dtype=np.int32)
    tk, tv = to_dev(k, dev), to_dev(v, dev)
    o = s512.Options()
    o.tile_variant = variant
    s512.kv_sort_ex(tk, tv, o)
    gk, gv = to_host(tk), to_host(tv)
    assert np.array_equal(gk, orc.closed_sort(k.copy()))
    assert np.array_equal(orc.pair_multiset(gk, gv), orc.pair_multiset(k, v))


def test_fp64_minmax_tiles_with_variants(s512, orc, dev):
    """fp64 inputs without NaN/-0.0 take the v_min/v_max tile kernels."""
    a = orc.gen_f64("uniform", N, 3)
    for variant in range(1, 4):
        got, _ = _run(s512, dev, a, tile_variant=variant)
        assert np.array_equal(got.view(np.uint64), np.sort(a).view(np.uint64)), variant


def test_variant_out_of_range(s512, dev):
    import torch
    t = torch.zeros(10, dtype=torch.float64, device=dev)
    for field, bad in (("scatter_variant", 12), ("count_variant", 6), ("tile_variant", 4)):
        o = s512.Options()
        setattr(o, field, bad)
        with pytest.raises(s512.S512Error):
            s512.sort_ex(t, o)


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
@pytest.mark.parametrize("n", [9000, 100_000, 300_000, 1_000_000, 5_000_000, 9_000_000])
def test_tuned_defaults_every_class(s512, orc, dev, dtype, n):
    """Default options take the autotuned table entry of n's size class."""
    a = _input(orc, dtype, "normal")
    a = np.concatenate([a] * (n // len(a) + 1))[:n] if n > len(a) else a[:n].copy()
    got, _ = _run(s512, dev, a)
    assert_keys_equal(got, orc.closed_sort(a.copy()), f"tuned defaults n={n}")


def test_tuned_defaults_kv_and_segmented(s512, orc, dev):
    k = (orc.gen_i32_full(2_000_000, 9) & 0xFFFFFF).astype(np.int32)
    v = np.arange(k.size, dtype=np.int32)
    tk, tv = to_dev(k, dev), to_dev(v, dev)
    s512.kv_sort(tk, tv)
    gk, gv = to_host(tk), to_host(tv)
    assert np.array_equal(gk, orc.closed_sort(k.copy()))
    assert np.array_equal(orc.pair_multiset(gk, gv), orc.pair_multiset(k, v))
    import torch
    off = s512.lognormal_offsets(2000, seed=4)
    keys = orc.gen_f64("normal", int(off[-1]), 4)
    tk = to_dev(keys, dev)
    s512.segsort(tk, torch.from_numpy(off.view(np.int64)).to(dev), options=s512.Options())
    got = to_host(tk)
    for i in range(len(off) - 1):
        assert np.array_equal(got[off[i]:off[i + 1]], np.sort(keys[off[i]:off[i + 1]])), i


@pytest.mark.parametrize("variant", range(1, 4))
@pytest.mark.parametrize("dtype", [np.float64, np.int64])
def test_tile_16k_class_after_partition(s512, orc, dev, dtype, variant):
    """Buckets of 8193..16384 eight-byte keys left by a partition level go to
    the 16384-key tile: fan-out 2 halves 100,000 keys into buckets of about
    50,000, 25,000 and then 12,500 keys.  fp64 without NaN or -0.0 sorts them
    with min/max leaves, int64 with encoded compares; every tile shape."""
    n = 100_000
    if dtype == np.float64:
        a = orc.gen_f64("normal", n, 13)
    else:
        a = np.random.default_rng(13).integers(-(1 << 63), (1 << 63) - 1, n, dtype=np.int64)
    got, st = _run(s512, dev, a, tile_variant=variant, max_fanout_log2=1)
    assert st.levels >= 3, st.levels
    if dtype == np.float64:
        assert_keys_equal(got, orc.closed_sort(a.copy()), f"tile_variant={variant}")
    else:
        assert np.array_equal(got, np.sort(a)), f"tile_variant={variant}"


@pytest.mark.parametrize("variant", range(1, 4))
@pytest.mark.parametrize("logf", [3, 4, 5, 6, 7])
def test_fp64_minmax_leaf_classes_after_partition(s512, orc, dev, variant, logf):
    """One partition level of 100,000 normal doubles with fan-out 2^logf
    leaves buckets of about 12,500, 6,250, 3,125, 1,560 and 780 keys: every
    raw-double leaf shape from 512 to 16384 keys (with lane phases, the
    four-wave 8192-key tile) sorts leaves of a partition, bit-exact against
    numpy's sort."""
    a = orc.gen_f64("normal", 100_000, 17 + logf)
    got, st = _run(s512, dev, a, tile_variant=variant, max_fanout_log2=logf)
    assert st.levels >= 1, st.levels
    assert np.array_equal(got.view(np.uint64), np.sort(a).view(np.uint64)), (variant, logf)


def test_levels_with_thousands_of_next_segments(s512, dev):
    """Fan-out 2 on 268,435,456 int32 keys: about 14 partition levels, with
    8,192 segments handed to the last one (more than the 4,096 that come back
    with a level's counts, so the rest are read separately).  Sorted and the
    same multiset of keys."""
    import torch
    n = 1 << 28
    t = torch.empty(n, dtype=torch.int32, device=dev)
    s512.fill(t, "uniform", seed=21)
    x = t.to(torch.int64) * -7046029254386353131
    h0 = int((x ^ (x >> 29)).sum().item())
    o = s512.Options()
    o.max_fanout_log2 = 1
    st = s512.sort_ex(t, o)
    assert st.levels >= 13, st.levels
    assert bool(torch.all(t[1:] >= t[:-1]).item())
    x = t.to(torch.int64) * -7046029254386353131
    assert int((x ^ (x >> 29)).sum().item()) == h0
