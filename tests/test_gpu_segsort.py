"""GPU tests of the batched segmented sort (many arrays in one call).

Parity: every segment equals the oracle's loop of the restated Sort512::Sort
over the same segments (bit-exact for int32; by value with identical bits
except for the sign of zeros for float64).  At the full workload size
(200,000 lognormal segments, about 3e8 keys) the checks are size-independent:
every segment is sorted and keeps its multiset (per-segment sum of a 64-bit
hash of the key bits), plus an exact oracle check of a sample of segments.
"""
import time

import numpy as np
import pytest

from gpu_util import assert_keys_equal, to_dev, to_host

pytestmark = pytest.mark.gpu


def _keys(orc, s512, dev, dtype, n, seed, kind="normal"):
    import torch
    if dtype == np.int32:
        t = torch.empty(n, dtype=torch.int32, device=dev)
        s512.fill(t, "uniform" if kind == "normal" else kind, seed=seed, param=(3 if kind == "distinct" else 0))
    else:
        t = torch.empty(n, dtype=torch.float64, device=dev)
        s512.fill(t, kind, seed=seed, param=(3 if kind == "distinct" else 0))
    return t


def _check_against_oracle(orc, got, a0, off, what):
    want = a0.copy()
    orc.segsort(want, off)
    assert_keys_equal(got, want, what)


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
@pytest.mark.parametrize("kind", ["normal", "distinct"])
def test_segsort_lognormal_vs_oracle(s512, orc, dev, dtype, kind):
    import torch
    off = s512.lognormal_offsets(3000, seed=5)
    n = int(off[-1])
    t = _keys(orc, s512, dev, dtype, n, seed=9, kind=kind)
    a0 = to_host(t)
    st = s512.segsort(t, torch.from_numpy(off.view(np.int64)).to(dev), stats=True)
    _check_against_oracle(orc, to_host(t), a0, off, f"{kind} {np.dtype(dtype).name}")
    assert st.levels == 1  # segments above the largest tile exist at this size
    print(f"\nsegsort {kind} {np.dtype(dtype).name}: n={n} tiles={st.tiles} engine buckets={st.learned_levels}")


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_segsort_edge_sizes(s512, orc, dev, dtype):
    """Empty and single-key segments, every tile-class boundary, segments
    just above the largest tile, and keys before the first segment."""
    import torch
    caps = [256, 512, 1024, 2048, 4096, 8192] + ([16384] if dtype == np.int32 else [])
    sizes = [0, 1, 0, 2, 3, 63, 64, 65]
    for c in caps:
        sizes += [c - 1, c, c + 1]
    sizes += [0, 100_000, 1, 300_001, 0]
    lead = 17  # keys before off[0] are not in any segment
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64) + lead
    n = int(off[-1])
    t = _keys(orc, s512, dev, dtype, n, seed=3)
    a0 = to_host(t)
    s512.segsort(t, torch.from_numpy(off.view(np.int64)).to(dev))
    got = to_host(t)
    assert np.array_equal(got[:lead].view(np.uint8), a0[:lead].view(np.uint8))
    _check_against_oracle(orc, got, a0, off, "edge sizes")


def test_segsort_oversized_buckets_second_level(s512, orc, dev):
    """One 6M-key fp64 segment: the first partition level leaves buckets
    above the largest tile; the sequence's second level finishes them, so no
    bucket goes to the host engine, and the result is exact."""
    import torch
    sizes = [5, 6_000_000, 1000, 9000]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    t = _keys(orc, s512, dev, np.float64, int(off[-1]), seed=21, kind="uniform")
    a0 = to_host(t)
    st = s512.segsort(t, torch.from_numpy(off.view(np.int64)).to(dev), stats=True)
    assert st.levels == 2 and st.learned_levels == 0, st.as_dict()
    _check_against_oracle(orc, to_host(t), a0, off, "oversized buckets")


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_segsort_oversized_after_two_levels_finish_on_host_engine(s512, orc, dev, dtype):
    """4-way levels (max_fanout_log2 = 2) leave buckets above the largest
    tile even after the second level: those finish in the multi-level engine
    after the sequence, and the result is still exact."""
    import torch
    sizes = [3, 2_000_003, 40_000, 0, 700_001]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    t = _keys(orc, s512, dev, dtype, int(off[-1]), seed=22)
    a0 = to_host(t)
    o = s512.Options()
    o.max_fanout_log2 = 2
    st = s512.segsort(t, torch.from_numpy(off.view(np.int64)).to(dev), stats=True, options=o)
    assert st.levels == 2 and st.learned_levels > 0, st.as_dict()
    _check_against_oracle(orc, to_host(t), a0, off, "oversized after two levels")


def test_segsort_special_values(s512, orc, dev):
    """NaN, +-inf, -0.0/+0.0 and denormals inside segments of every path:
    order by the library's key order (the reference leaves NaN undefined)."""
    import torch
    rng = np.random.default_rng(4)
    sizes = [70, 3000, 20000]
    n = sum(sizes)
    a = rng.standard_normal(n)
    idx = rng.integers(0, n, 600)
    specials = np.array([np.nan, -np.nan, np.inf, -np.inf, 0.0, -0.0, 5e-324, -5e-324])
    a[idx] = specials[np.arange(600) % len(specials)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    t = to_dev(a, dev)
    s512.segsort(t, torch.from_numpy(off.view(np.int64)).to(dev))
    got = to_host(t)
    # order-preserving encoding of the bit patterns (-0.0 before +0.0, NaNs by bits)
    for i in range(len(sizes)):
        seg = a[off[i]:off[i + 1]].view(np.uint64)
        enc = np.where(seg >> np.uint64(63), ~seg, seg | np.uint64(1 << 63))
        want = seg[np.argsort(enc, kind="stable")]
        assert np.array_equal(got[off[i]:off[i + 1]].view(np.uint64), want), i


def test_segsort_host_arrays(s512, orc):
    off = s512.lognormal_offsets(500, median=100, sigma=1.0, seed=2)
    a = orc.gen_i32_full(int(off[-1]), 8)
    a0 = a.copy()
    s512.segsort(a, off)
    _check_against_oracle(orc, a, a0, off, "host arrays")


def test_segsort_invalid_offsets_leave_keys(s512, dev):
    import torch
    t = torch.arange(1000, 0, -1, dtype=torch.int32, device=dev)
    before = to_host(t).copy()
    off = torch.tensor([0, 500, 400, 1000], dtype=torch.int64, device=dev)
    with pytest.raises(s512.S512Error) as e:
        s512.segsort(t, off)
    assert e.value.status == 1
    assert np.array_equal(to_host(t), before)
    s512.segsort(t, torch.tensor([0], dtype=torch.int64, device=dev))  # no segments: no-op
    assert np.array_equal(to_host(t), before)


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_segsort_graph_plan_replays(s512, orc, dev, dtype):
    """The captured plan, replayed three times on new keys with eager work in
    between, equals the oracle every time (and the direct entry point)."""
    import torch
    off = s512.lognormal_offsets(2000, seed=11)
    n = int(off[-1])
    toff = torch.from_numpy(off.view(np.int64)).to(dev)
    t = _keys(orc, s512, dev, dtype, n, seed=1)
    plan = s512.SegsortPlan(t, toff)
    try:
        for rep in range(3):
            s512.fill(t, "uniform" if dtype == np.int32 else "normal", seed=100 + rep)
            a0 = to_host(t)
            direct = t.clone()
            s512.segsort(direct, toff)
            plan.launch()
            got = to_host(t)
            _check_against_oracle(orc, got, a0, off, f"replay {rep}")
            assert np.array_equal(got.view(np.uint8), to_host(direct).view(np.uint8))
            torch.cuda.synchronize()
    finally:
        plan.close()


def _seg_hash_sums(x, off):
    v = x.view(np.uint64) if x.dtype == np.float64 else x.view(np.uint32).astype(np.uint64)
    h = v * np.uint64(0x9E3779B97F4A7C15)
    h ^= h >> np.uint64(29)
    h *= np.uint64(0xBF58476D1CE4E5B9)
    h ^= h >> np.uint64(32)
    p = np.zeros(x.size + 1, dtype=np.uint64)
    np.cumsum(h, out=p[1:])
    return p[off[1:].astype(np.int64)] - p[off[:-1].astype(np.int64)]


def _segments_sorted(x, off):
    if x.dtype == np.float64:
        ok = x[1:] >= x[:-1]
    else:
        ok = x[1:] >= x[:-1]
    starts = off[1:-1].astype(np.int64)
    starts = starts[(starts > 0) & (starts < x.size)]
    ok[starts - 1] = True
    return bool(ok.all())


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_segsort_full_workload(s512, orc, dev, dtype, capsys):
    """200,000 arrays, sizes lognormal(median 512, sigma 1.5) clipped to
    [1, 2^20]; int32 uniform / fp64 normal keys."""
    import torch
    off = s512.lognormal_offsets(200_000, seed=0)
    n = int(off[-1])
    toff = torch.from_numpy(off.view(np.int64)).to(dev)
    t = _keys(orc, s512, dev, dtype, n, seed=0)
    a0 = to_host(t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = s512.segsort(t, toff, stats=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    got = to_host(t)
    assert _segments_sorted(got, off)
    assert np.array_equal(_seg_hash_sums(got, off), _seg_hash_sums(a0, off))
    rng = np.random.default_rng(1)
    for i in rng.choice(200_000, 64, replace=False):
        seg = a0[off[i]:off[i + 1]].copy()
        orc.sort(seg)
        assert_keys_equal(got[off[i]:off[i + 1]], seg, f"segment {i}")
    with capsys.disabled():
        print(f"\nsegsort 200k lognormal {np.dtype(dtype).name}: {n} keys, {dt * 1e3:.2f} ms "
              f"({n / dt / 1e9:.2f} G keys/s incl. host sync), engine buckets={st.learned_levels}")
    assert st.learned_levels == 0, "every bucket finishes inside the sequence"


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_single_array_plan_stays_in_graph(s512, dev, dtype):
    """One array of 2^25 keys as a single segment: the plan's two 128-way levels
    leave buckets of about 2,048 keys, so every bucket is finished inside the
    captured sequence (no host engine call) and replays sort correctly."""
    import torch
    n = 1 << 25
    rng = np.random.default_rng(25)
    a = rng.standard_normal(n) if dtype == np.float64 else rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32)
    t = torch.from_numpy(a).to(dev)
    off = torch.tensor([0, n], dtype=torch.int64, device=dev)
    plan = s512.SegsortPlan(t, off)
    want = np.sort(a)
    for rep in range(2):
        t.copy_(torch.from_numpy(a))
        st = plan.launch()
        assert st.learned_levels == 0, st.as_dict()  # buckets left to the host engine
        assert np.array_equal(t.cpu().numpy().view(np.uint8), want.view(np.uint8)), rep
