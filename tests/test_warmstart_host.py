"""Host logic of the warm-start training driver (no GPU): random crops are drawn over
the whole of each volume, whatever its shape, and pair input with target voxel for
voxel."""
import numpy as np
import torch

from rtk_amd import warmstart_train as W


def test_random_crops_cover_each_volume_and_stay_paired():
    shape = (96, 80, 72)
    vol = torch.arange(np.prod(shape), dtype=torch.float32).reshape(shape)
    pairs = [(vol, -vol)]
    rng = np.random.default_rng(3)
    lo, hi = np.full(3, 10 ** 9), np.zeros(3, int)
    for _ in range(200):
        x, t = W.random_crops(pairs, 32, 2, rng)
        assert x.shape == (2, 32, 32, 32) and torch.equal(t, -x)
        for c in x:
            flat = int(c[0, 0, 0].item())
            z, y, w = np.unravel_index(flat, shape)
            assert torch.equal(c, vol[z:z + 32, y:y + 32, w:w + 32])
            lo, hi = np.minimum(lo, (z, y, w)), np.maximum(hi, (z, y, w))
    # origins reach both ends of every axis
    assert np.all(lo == 0) and np.all(hi == np.array(shape) - 32)
