"""Pins the oracle (CPU restatement) before anything is checked against it.

Each restated function is checked against at least one of: the reference's
own known-answer vectors, outputs of the compiled reference stored in
tests/golden/ref_vectors.npz, the live compiled reference (oracle/_ref, when
built), libc, a closed form, or brute force.  A test that only retyped the
oracle's formula would pin nothing, so every check here compares against an
independent source.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "ref_vectors.npz"))


@pytest.fixture(scope="module")
def kat():
    return json.load(open(os.path.join(GOLD, "kat_reference_tests.json")))


@pytest.fixture(scope="module")
def ref(orc):
    r = orc.ref_lib()
    if r is None:
        pytest.skip("compiled reference not built (oracle/_ref needs /root/reference)")
    if not orc.cpu_has_avx512():
        pytest.skip("host CPU lacks AVX-512; the compiled reference cannot run")
    return r


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- generators
def test_drand48_matches_libc(orc):
    """createRandVec's generator is POSIX drand48: compare with libc's."""
    libc = ctypes.CDLL(None)
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (0, 1, 12345):
        libc.srand48(seed)
        orc.srand48(seed)
        for _ in range(2000):
            assert orc.drand48() == libc.drand48()


def test_splitmix64_known_values(orc):
    """SplitMix64 finaliser (Steele et al. 2014) recomputed in Python ints."""
    M = (1 << 64) - 1

    def sm(seed, i):
        z = (seed * 0x9E3779B97F4A7C15 + (i + 1) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    for seed in (0, 1, 7, 2**63 + 5):
        for i in (0, 1, 2, 1000, 2**40):
            assert orc.splitmix64(seed, i) == sm(seed, i)


def test_generators_shape(orc):
    u = orc.gen_f64("uniform", 100_000, 0)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.01
    z = orc.gen_f64("normal", 200_000, 1)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01 and np.isfinite(z).all()
    d = orc.gen_f64("distinct", 100_000, 2)
    assert len(np.unique(d)) == 1000
    s = orc.gen_f64("sorted", 1000, 0)
    assert np.all(np.diff(s) > 0)
    r = orc.gen_f64("reversed", 1000, 0)
    assert np.all(np.diff(r) < 0)
    a = orc.gen_i32_full(100_000, 0)
    assert a.min() < -2**30 and a.max() > 2**30


# ---------------------------------------------------------------- KATs
def test_kat_vectors(orc, kat):
    """The known-answer vectors of sort512test.cpp, through the restated
    networks (SmallSort16V covers the 1- and 2-vector networks)."""
    for case in kat["cases"]:
        dt = np.int32 if case["dtype"] == "int32" else np.float64
        a = np.array(case["input"], dtype=dt)
        orc.smallsort(a)
        assert np.array_equal(a, np.array(case["expect"], dtype=dt)), case["source"]
    for case in kat["kv_cases"]:
        k = np.array(case["keys"], dtype=np.int32)
        v = (k.astype(np.int64) * 100 + 1).astype(np.int32)
        orc.kv_smallsort(k, v)
        assert np.array_equal(k, np.array(case["expect"], dtype=np.int32)), case["source"]
        assert np.array_equal(v, k * 100 + 1)


# ---------------------------------------------------------------- golden
def _split(arr, lengths):
    out, o = [], 0
    for n in lengths:
        out.append(arr[o:o + n])
        o += n
    return out


def test_smallsort_matches_reference_outputs(orc, gold):
    """Every length 1..256 (int) / 1..128 (double): same output as the
    compiled reference's SmallSort16V, inputs reproduced by drand48."""
    for dt, name, maxn in ((np.int32, "i32", 256), (np.float64, "f64", 128)):
        lens = list(range(1, maxn + 1))
        ins = _split(gold[f"small_{name}_in"], lens)
        outs = _split(gold[f"small_{name}_out"], lens)
        orc.srand48(0)
        for n, a_in, a_out in zip(lens, ins, outs):
            a = orc.rand_vec(n, dt)
            assert np.array_equal(a, a_in), f"drand48 input differs at n={n}"
            orc.smallsort(a)
            assert np.array_equal(a, a_out), f"{name} n={n}"


def test_kv_smallsort_matches_reference(orc, gold):
    lens = list(range(1, 257))
    ins = _split(gold["kvsmall_in"], lens)
    kouts = _split(gold["kvsmall_kout"], lens)
    vouts = _split(gold["kvsmall_vout"], lens)
    for n, k_in, k_out, v_out in zip(lens, ins, kouts, vouts):
        k = k_in.copy()
        v = (k.astype(np.int64) * 100 + 1).astype(np.int32)
        orc.kv_smallsort(k, v)
        assert np.array_equal(k, k_out), n
        # the value rule the reference tests check (sort512test.cpp:1581-1595)
        assert np.array_equal(v, (k.astype(np.int64) * 100 + 1).astype(np.int32))
        assert np.array_equal(np.sort(v), np.sort(v_out))


def test_partition_bound_matches_reference(orc, gold):
    """Partition512 return value for n = 1..1000, pivot n/2, and the
    partition property (sort512test.cpp:41-60, 1462-1471)."""
    for dt, name in ((np.int32, "i32"), (np.float64, "f64")):
        bounds = gold[f"part_{name}_bound"]
        orc.srand48(0)
        for n in range(1, 1001):
            a = orc.rand_vec(n, dt)
            a0 = a.copy()
            pivot = dt(n // 2)
            b = orc.partition(a, 0, n - 1, int(pivot) if dt == np.int32 else float(pivot))
            assert b == bounds[n - 1], (name, n)
            assert np.all(a[:b] <= pivot) and np.all(a[b:] > pivot)
            assert np.array_equal(np.sort(a), np.sort(a0))


def test_partition_subrange_and_closed_form(orc):
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = int(rng.integers(1, 300))
        a = rng.integers(-50, 50, n).astype(np.int32)
        l = int(rng.integers(0, n))
        r = int(rng.integers(l, n))
        pivot = int(rng.integers(-60, 60))
        a0 = a.copy()
        b = orc.partition(a, l, r, pivot)
        assert b == l + int(np.count_nonzero(a0[l:r + 1] <= pivot))
        assert np.array_equal(a[:l], a0[:l]) and np.array_equal(a[r + 1:], a0[r + 1:])
        assert np.all(a[l:b] <= pivot) and np.all(a[b:r + 1] > pivot)


def test_sort_matches_reference_outputs(orc, gold):
    for dt, name in ((np.int32, "i32"), (np.float64, "f64")):
        lens = [1 << k for k in range(11)]
        ins = _split(gold[f"sort_{name}_in"], lens)
        outs = _split(gold[f"sort_{name}_out"], lens)
        for a_in, a_out in zip(ins, outs):
            a = a_in.copy()
            orc.sort(a)
            assert np.array_equal(a, a_out)


def test_config1_digest(orc, gold):
    """Config 1 (1M int32 full range, seed 0): the restated Sort512::Sort
    output hashes to the compiled reference's."""
    a = orc.gen_i32_full(1 << 20, 0)
    orc.sort(a)
    assert hashlib.sha256(a.tobytes()).digest() == gold["cfg1_sha256"].tobytes()


def test_reference_quirk_inf_reproduced(orc, gold):
    """Q1: the reference pads leaves with DBL_MAX, so +inf comes back as
    DBL_MAX; the restatement reproduces it (the GPU library does not)."""
    q = np.array([np.inf, 1.0, 2.0, 3.0, 4.0])
    orc.smallsort(q)
    assert np.array_equal(q, gold["quirk_inf_out"])
    assert q[-1] == np.finfo(np.float64).max


def test_findmedian_matches_reference(orc, gold):
    for trial, na, nb, mA, mB in gold["findmedian"]:
        a = gold[f"fm_{trial}"].astype(np.int32)
        assert orc.findmedian(a, int(na)) == (mA, mB), trial


def test_sort_closed_form_pathological(orc):
    """Duplicates, sorted and reversed inputs: the restated quicksort equals
    the closed form (np.sort) at sizes where it is still fast."""
    rng = np.random.default_rng(0)
    cases = [rng.integers(0, 3, 20000).astype(np.int32), np.arange(30000, dtype=np.int32)[::-1].copy(),
             np.full(5000, 7, dtype=np.int32), rng.standard_normal(50000)]
    for a in cases:
        want = np.sort(a)
        orc.sort(a)
        assert np.array_equal(a, want)


def test_kv_sort_closed_form(orc):
    rng = np.random.default_rng(1)
    k = rng.integers(-1000, 1000, 100000).astype(np.int32)
    v = np.arange(k.size, dtype=np.int32)
    k0, v0 = k.copy(), v.copy()
    orc.kv_sort(k, v)
    assert np.array_equal(k, np.sort(k0))
    assert np.array_equal(orc.pair_multiset(k, v), orc.pair_multiset(k0, v0))


def test_pair_multiset_independent(orc):
    rng = np.random.default_rng(2)
    k = rng.integers(-5, 5, 1000).astype(np.int32)
    v = rng.integers(-9, 9, 1000).astype(np.int32)
    got = orc.pair_multiset(k, v)
    want = sorted(zip(k.tolist(), v.tolist()))
    dec = [(int(np.int32(np.uint32((x >> 32) ^ 0x80000000))), int(np.int32(np.uint32(x & 0xFFFFFFFF)))) for x in got]
    assert sorted(dec) == want
    assert list(got) == sorted(got)


# ---------------------------------------------------------------- merge
def test_merge_and_corank_brute_force(orc):
    rng = np.random.default_rng(5)
    for _ in range(300):
        na, nb = int(rng.integers(0, 50)), int(rng.integers(0, 50))
        A = np.sort(rng.integers(0, 20, na)).astype(np.int32)
        B = np.sort(rng.integers(0, 20, nb)).astype(np.int32)
        a = np.concatenate([A, B])
        orc.merge(a, na)
        assert np.array_equal(a, np.sort(np.concatenate([A, B])))
        # stable merge by brute force: A wins ties
        tagged = sorted([(x, 0, i) for i, x in enumerate(A)] + [(x, 1, i) for i, x in enumerate(B)])
        for diag in range(na + nb + 1):
            took_a = sum(1 for t in tagged[:diag] if t[1] == 0)
            assert orc.corank(A, B, diag) == took_a


def test_findmedian_live_reference(orc, ref):
    rng = np.random.default_rng(7)
    for _ in range(500):
        na, nb = int(rng.integers(0, 100)), int(rng.integers(0, 100))
        if na + nb == 0:
            continue
        for dt in (np.int32, np.float64):
            a = np.concatenate([np.sort(rng.integers(0, 40, na)), np.sort(rng.integers(0, 40, nb))]).astype(dt)
            mA, mB = ctypes.c_int(), ctypes.c_int()
            fn = ref.ref_findmedian_i32 if dt == np.int32 else ref.ref_findmedian_f64
            fn(_p(a), na, na + nb, ctypes.byref(mA), ctypes.byref(mB))
            assert orc.findmedian(a, na) == (mA.value, mB.value)


# ---------------------------------------------------------------- live reference
@pytest.mark.parametrize("n", [1, 2, 200, 257, 4096, 100_000, 1_000_003])
def test_sort_live_reference(orc, ref, n):
    for dt, fn in ((np.int32, ref.ref_sort_i32), (np.float64, ref.ref_sort_f64)):
        a = orc.gen_i32_full(n, n) if dt == np.int32 else orc.gen_f64("normal", n, n)
        b = a.copy()
        fn(_p(b), n)
        orc.sort(a)
        assert np.array_equal(a, b)


def test_kv_live_reference(orc, ref):
    orc.srand48(0)
    for n in (3, 100, 256, 1000, 50_000):
        k = orc.rand_vec(n, np.int32)
        v = (k.astype(np.int64) * 100 + 1).astype(np.int32)
        k1, v1, k2, v2 = k.copy(), v.copy(), k.copy(), v.copy()
        ref.ref_kv_sort(_p(k1), _p(v1), n)
        orc.kv_sort(k2, v2)
        assert np.array_equal(k1, k2)
        assert np.array_equal(orc.pair_multiset(k1, v1), orc.pair_multiset(k2, v2))


def test_partition_live_reference(orc, ref):
    rng = np.random.default_rng(9)
    for _ in range(300):
        n = int(rng.integers(1, 3000))
        a = rng.integers(-1000, 1000, n).astype(np.int32)
        l = int(rng.integers(0, n))
        r = int(rng.integers(l, n))
        pivot = int(rng.integers(-1100, 1100))
        b1, b2 = a.copy(), a.copy()
        got = ref.ref_partition_i32(_p(b1), l, r, pivot)
        assert orc.partition(b2, l, r, pivot) == got
        assert np.array_equal(np.sort(b1), np.sort(b2))


def test_kv_partition_live_reference(orc, ref):
    """orc_partition_kv against the compiled Sort512kv::Partition512
    (sort512kv.hpp:5958): the same bound, the same (key, value) multiset in
    each side, and the closed form (bound = left + #{keys <= pivot}); values
    are key*100+1 so that a pair split apart fails."""
    rng = np.random.default_rng(12)
    for _ in range(300):
        n = int(rng.integers(1, 3000))
        k = rng.integers(-1000, 1000, n).astype(np.int32)
        v = (k.astype(np.int64) * 100 + 1).astype(np.int32)
        l = int(rng.integers(0, n))
        r = int(rng.integers(l, n))
        pivot = int(rng.integers(-1100, 1100))
        k1, v1, k2, v2 = k.copy(), v.copy(), k.copy(), v.copy()
        got = ref.ref_kv_partition(_p(k1), _p(v1), l, r, pivot)
        b = orc.kv_partition(k2, v2, l, r, pivot)
        assert b == got == l + int(np.count_nonzero(k[l:r + 1] <= pivot))
        assert np.all(k2[l:b] <= pivot) and np.all(k2[b:r + 1] > pivot)
        assert np.array_equal(k2[:l], k[:l]) and np.array_equal(k2[r + 1:], k[r + 1:])
        assert np.array_equal(v2, (k2.astype(np.int64) * 100 + 1).astype(np.int32))
        for lo, hi in ((l, b), (b, r + 1)):
            assert np.array_equal(orc.pair_multiset(k1[lo:hi], v1[lo:hi]), orc.pair_multiset(k2[lo:hi], v2[lo:hi]))


@pytest.mark.parametrize("dtype", [np.int32, np.float64])
def test_segsort_matches_numpy_per_segment(orc, dtype):
    """orc_segsort (the caller's loop of Sort512::Sort, sort512.hpp:5960)
    against numpy's sort of each segment (the ascending order is unique):
    empty, one-key, leaf-sized and multi-level segments; decreasing offsets
    are refused."""
    rng = np.random.default_rng(13)
    sizes = np.array([0, 1, 2, 3, 17, 128, 129, 256, 257, 0, 5000, 70_000, 1, 300_001], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(off[-1])
    a = rng.integers(-(1 << 31), 1 << 31, n).astype(np.int32) if dtype == np.int32 else rng.standard_normal(n)
    a[70_000:70_100] = a[70_000]  # repeated keys inside a segment
    got = orc.segsort(a.copy(), off)
    want = np.concatenate([np.sort(a[off[i]:off[i + 1]]) for i in range(len(sizes))])
    assert np.array_equal(got.view(np.uint64 if dtype == np.float64 else np.uint32),
                          want.view(np.uint64 if dtype == np.float64 else np.uint32))
    bad = off.copy()
    bad[3], bad[4] = bad[4], bad[3]
    with pytest.raises(AssertionError):
        orc.segsort(a.copy(), bad)


def test_segsort_live_reference_loop(orc, ref):
    """orc_segsort against the compiled reference's Sort looped over the same
    segments, on one thread and under OpenMP."""
    rng = np.random.default_rng(14)
    sizes = rng.integers(0, 3000, 400).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    a = rng.standard_normal(int(off[-1]))
    want = orc.segsort(a.copy(), off)
    for par in (0, 1):
        b = a.copy()
        ref.ref_segsort_loop_f64(_p(b), _p(off), len(sizes), par)
        assert np.array_equal(b.view(np.uint64), want.view(np.uint64)), par


# ---------------------------------------------------------------- MLP
def test_mlp_gradient_finite_differences(orc):
    """Central differences on the fp64 restatement of the model gradient
    (points kept away from the ramp kinks, where the derivative jumps)."""
    H = 6
    rng = np.random.default_rng(11)
    p = rng.standard_normal(orc.mlp_nparams(H)) * 0.7
    x = rng.uniform(-1, 1, 40)
    t = np.sort(rng.uniform(0, 1, 40))
    loss, g = orc.mlp_loss_grad(p, H, x, t)
    eps = 1e-7
    for j in range(p.size):
        pp, pm = p.copy(), p.copy()
        pp[j] += eps
        pm[j] -= eps
        num = (orc.mlp_loss_grad(pp, H, x, t)[0] - orc.mlp_loss_grad(pm, H, x, t)[0]) / (2 * eps)
        assert abs(num - g[j]) <= 1e-5 * max(1.0, abs(num)), (j, num, g[j])


def _ramp_net_interp(orc, H, knots):
    """Parameters for which the net is the piecewise-linear interpolation of
    the points (knots[j], j/H): layer-1 unit j ramps across [k_j, k_{j+1}],
    layer 2 passes units through, layer 3 averages them."""
    p = np.zeros(orc.mlp_nparams(H))
    w = 1.0 / np.diff(knots)
    p[:H] = np.sqrt(w)
    p[H:2 * H] = -knots[:-1] * w
    p[2 * H:2 * H + H * H] = np.eye(H).ravel()
    b2 = 2 * H + H * H
    p[b2 + H:b2 + 2 * H] = np.sqrt(1.0 / H)
    return p


def test_mlp_forward_closed_form_interpolation(orc):
    """With the quantile-ramp parameters the model equals np.interp through
    the knots (a textbook closed form), and is monotone for any parameters."""
    H = 8
    knots = np.sort(np.random.default_rng(3).uniform(-2, 3, H + 1))
    p = _ramp_net_interp(orc, H, knots)
    xs = np.linspace(-3, 4, 701)
    ys = np.array([orc.mlp_forward(p, H, x) for x in xs])
    assert np.allclose(ys, np.interp(xs, knots, np.arange(H + 1) / H), atol=1e-12)
    rng = np.random.default_rng(12)
    for _ in range(5):
        p2 = rng.standard_normal(orc.mlp_nparams(H))
        ys2 = np.array([orc.mlp_forward(p2, H, x) for x in np.linspace(-3, 3, 301)])
        assert np.all(np.diff(ys2) >= -1e-12)
