"""GPU tests of the learned partitioner (no reference counterpart).

- The fp32 training kernel's loss and gradient equal the oracle's fp64
  restatement of the same model (tolerance written below).
- A learned sort returns exactly the keys of the plain sort (bit-exact):
  the model only changes bucket sizes.
- On the piecewise-constant adversarial input the model is rejected and the
  sampled-splitter fallback runs (asserted through the stats): its 32 narrow
  pieces each put 1/64 of the keys in a range far narrower than any cell of
  the model's CDF table, so one first-level bucket holds 8x the mean and the
  imbalance check falls back before any key moves.  On 1000 distinct values
  the table sees keys on repeated values and is not used.
- Training gains: the trained network's CDF fit against the one-step
  network's, and what it does to the second level's leaves.
- The A/B test records the learned path's time against sampled splitters on
  the skewed mixture input.
"""
import time

import numpy as np
import pytest

from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu

# fp32 kernel vs fp64 oracle: relative to the gradient's largest component
GRAD_RTOL = 2e-4


def _params(orc, H, rng):
    p = rng.standard_normal(orc.mlp_nparams(H)) * 0.6
    # keep a share of the ramps in their linear region so gradients are non-trivial
    p[:H] = rng.uniform(1.0, 3.0, H)
    p[H:2 * H] = rng.uniform(-2.0, 0.5, H)
    return p


@pytest.mark.parametrize("m", [1000, 256, 100])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_training_gradient_matches_oracle(s512, orc, dev, seed, m):
    """m <= 256 runs the single-workgroup training kernel with zero steps (the
    kernel the learned sort trains with); m = 1000 the multi-workgroup step."""
    import torch
    H = 32
    rng = np.random.default_rng(seed)
    p = _params(orc, H, rng)
    u = np.sort(rng.uniform(0, 1, m))
    t = np.sort(rng.uniform(0, 1, m))
    loss_ref, g_ref = orc.mlp_loss_grad(p, H, u, t)
    tp = torch.tensor(p, dtype=torch.float32, device=dev)
    loss, g = s512.mlp_loss_grad(tp, torch.tensor(u, dtype=torch.float32, device=dev),
                                 torch.tensor(t, dtype=torch.float32, device=dev))
    # the oracle evaluated at the fp32-rounded parameters and points
    p32, u32, t32 = (np.float64(np.float32(v)) for v in (p, u, t))
    loss_ref, g_ref = orc.mlp_loss_grad(p32, H, u32, t32)
    g = g.cpu().numpy().astype(np.float64)
    scale = np.abs(g_ref).max()
    assert scale > 0
    assert abs(loss - loss_ref) <= 1e-4 * max(loss_ref, 1e-3)
    bad = np.abs(g - g_ref) > GRAD_RTOL * scale
    assert not bad.any(), (np.flatnonzero(bad)[:10], g[bad][:5], g_ref[bad][:5])


def _fill(s512, dev, kind, n, seed):
    import torch
    t = torch.empty(n, dtype=torch.float64, device=dev)
    s512.fill(t, kind, seed=seed, param=(1000 if kind == "distinct" else 0))
    return t


@pytest.mark.parametrize("kind", ["uniform", "normal", "mixture", "distinct", "sorted", "piecewise"])
@pytest.mark.parametrize("n", [300_000, 4_000_000])
def test_learned_sort_exact(s512, dev, kind, n):
    t = _fill(s512, dev, kind, n, seed=7)
    a0 = to_host(t)
    # a 512-way first level (the tuned defaults may pick fewer buckets at
    # these sizes; the piecewise input is adversarial relative to 512)
    st = s512.sort_learned(t, max_fanout_log2=9, target_leaf=1024)
    got = to_host(t)
    assert np.array_equal(got.view(np.uint64), np.sort(a0).view(np.uint64)), st.as_dict()
    if kind == "piecewise":
        # 32 pieces of width 1e-9 each hold 1/64 of the keys: 8x a first-level
        # bucket; the model must be rejected
        assert st.learned_fallback == 1 and st.classifier == 0, st.as_dict()
    if kind in ("uniform", "normal", "mixture", "sorted"):
        assert st.learned_fallback == 0 and st.classifier == 1, st.as_dict()
        assert st.learned_max_bucket_ratio < 2.0, st.as_dict()
    print(f"\n{kind} n={n}: classifier={st.classifier} fallback={st.learned_fallback} "
          f"max_bucket_ratio={st.learned_max_bucket_ratio:.2f} loss={st.model_loss:.3e}")


def test_learned_sort_int32(s512, orc, dev):
    a0 = orc.gen_i32_full(3_000_000, 5)
    t = to_dev(a0, dev)
    st = s512.sort_learned(t)
    assert np.array_equal(to_host(t), np.sort(a0)), st.as_dict()
    assert st.classifier == 1


@pytest.mark.parametrize("kind", ["normal", "mixture"])
def test_learned_vs_splitters_ab(s512, dev, capsys, kind):
    """Same input (64M keys) through both classifiers; both exact, times and
    the learned path's outcome printed for the record."""
    import torch
    n = 64 << 20
    src = _fill(s512, dev, kind, n, seed=11)
    ref = None
    times = {}
    for name in ("splitters", "learned", "splitters", "learned"):
        t = src.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = s512.sort_learned(t) if name == "learned" else s512.sort_ex(t)
        torch.cuda.synchronize()
        times[name] = time.perf_counter() - t0
        if ref is None:
            ref = t.clone()
        assert torch.equal(t.view(torch.int64), ref.view(torch.int64))
        if name == "learned":
            lst = st.as_dict()
    with capsys.disabled():
        print(f"\nA/B {kind} 64M fp64: splitters {times['splitters']*1e3:.2f} ms, "
              f"learned {times['learned']*1e3:.2f} ms, ratio {times['splitters']/times['learned']:.3f}; "
              f"learned classifier={lst['classifier']} fallback={lst['learned_fallback']} "
              f"max_ratio={lst['learned_max_bucket_ratio']:.2f}")


@pytest.mark.parametrize("kind", ["normal", "mixture"])
def test_training_gain(s512, dev, capsys, kind):
    """What training buys.  The network's CDF, tabulated at the cell boundaries
    of the model's table (within a quarter of a leaf around the sample's own
    table), splits every root bucket at the second level.  Asserted: training
    lowers the CDF fit (full-batch MSE of the training points) against the
    network after one Adam step from its initialisation (the piecewise-linear
    interpolation through 33 quantiles), and the trained network's second-level
    leaves are no less balanced (largest leaf / mean, learned_leaf_max_ratio)
    than the one-step network's.  The first level uses the sample's own table,
    so its balance is the same for every training length.  (The band of a
    quarter leaf keeps the network from moving any leaf boundary by more than
    that: its fit is coarser than the sample's, so the sort gains nothing
    measurable from training -- the leaf ratios print equal.)"""
    n = 16 << 20
    src = _fill(s512, dev, kind, n, seed=5)
    res = {}
    for steps in (1, 30, 200):
        t = src.clone()
        st = s512.sort_learned(t, train_steps=steps, max_fanout_log2=9, target_leaf=1024)
        # every leaf of the model's second level fits a tile: no third level
        assert st.classifier == 1 and st.learned_levels == 2 and st.levels == 2, st.as_dict()
        res[steps] = (st.model_loss, st.learned_max_bucket_ratio, st.learned_leaf_max_ratio)
    with capsys.disabled():
        print(f"\n{kind} 16M: " + ", ".join(f"{k} steps: loss {v[0]:.2e} root max/mean {v[1]:.3f} "
                                           f"leaf max/mean {v[2]:.3f}" for k, v in res.items()))
    assert res[30][0] < 0.5 * res[1][0], res
    assert res[200][0] < res[30][0], res
    assert res[30][1] == res[1][1] == res[200][1], res
    assert res[30][2] <= 1.02 * res[1][2], res
