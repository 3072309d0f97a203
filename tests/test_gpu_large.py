"""GPU tests of sorts past 2^32 keys (the reference's IndexType is size_t,
sort512.hpp:5960-5963).

An array of 2^32 keys or more is sorted as pieces of 2^31 keys, each an
ordinary engine sort, then merged by merge-path rounds.  4,294,971,395
(2^32 + 4,099) keys give pieces of 2^31, 2^31 and 4,099 keys, so both merge
rounds run and one of them pairs a full run with a short one.  Checked by
size-independent properties, in slices that keep the temporaries small: every
key not below its predecessor (slices overlap by one key), and an
order-independent multiset hash equal to the input's.  The merge sort
(SortOmpMerge family) takes the same sizes.  Key/value sorts still take fewer
than 2^32 pairs, and say so.
"""
import pytest

pytestmark = pytest.mark.gpu

N = (1 << 32) + 4099
SLICE = 1 << 28


def _hash(t):
    import torch
    x = t.view(torch.int64) if t.dtype == torch.float64 else t.to(torch.int64)
    x = x * -7046029254386353131
    x = x ^ (x >> 29)
    x = x * -4658895280553007687
    return int((x ^ (x >> 32)).sum().item())


def _multiset(t):
    return sum(_hash(t[i:i + SLICE]) for i in range(0, t.numel(), SLICE)) & ((1 << 64) - 1)


def _sorted(t):
    import torch
    n = t.numel()
    for i in range(0, n - 1, SLICE):
        s = t[i:min(n, i + SLICE + 1)]
        if not bool(torch.all(s[1:] >= s[:-1]).item()):
            return False
    return True


@pytest.mark.parametrize("dtype,kind", [("int32", "uniform"), ("float64", "normal"), ("int32", "distinct")])
def test_sort_past_2_32_keys(s512, dev, dtype, kind):
    import torch
    t = torch.empty(N, dtype=getattr(torch, dtype), device=dev)
    s512.fill(t, kind, seed=31, param=1000)
    h0 = _multiset(t)
    print(f"\nsort of {N} {dtype} keys ({kind})", flush=True)
    st = s512.sort_ex(t)
    assert st.levels >= 2, st.as_dict()
    assert _sorted(t)
    assert _multiset(t) == h0
    del t
    torch.cuda.empty_cache()


def test_merge_sort_past_2_32_keys(s512, dev):
    import torch
    t = torch.empty(N, dtype=torch.int32, device=dev)
    s512.fill(t, "uniform", seed=32)
    h0 = _multiset(t)
    print(f"\nmerge sort of {N} int32 keys", flush=True)
    s512.sort_merge(t)
    assert _sorted(t)
    assert _multiset(t) == h0
    del t
    torch.cuda.empty_cache()


def test_kv_past_2_32_pairs_rejected(s512, dev):
    import torch
    k = torch.empty(N, dtype=torch.int32, device=dev)
    with pytest.raises(s512.S512Error, match="2\\^32"):
        s512.kv_sort_ex(k, k)
    del k
    torch.cuda.empty_cache()
