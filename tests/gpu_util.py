"""Helpers shared by the GPU parity tests."""
import numpy as np


def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def assert_keys_equal(got, want, what=""):
    """Bit-exact for int32; for float64 compare by value (+0.0 == -0.0) and
    require identical bit patterns everywhere else."""
    assert got.shape == want.shape, what
    if got.dtype == np.float64:
        eq = (got == want)
        if not eq.all():
            bad = np.flatnonzero(~eq)[:5]
            raise AssertionError(f"{what}: {len(np.flatnonzero(~eq))} keys differ, first at {bad}: "
                                 f"{got[bad]} vs {want[bad]}")
        nz = want != 0
        assert np.array_equal(got[nz].view(np.uint64), want[nz].view(np.uint64)), what
    else:
        if not np.array_equal(got, want):
            bad = np.flatnonzero(got != want)[:5]
            raise AssertionError(f"{what}: keys differ at {bad}: {got[bad]} vs {want[bad]}")
