"""Host logic of the warm-start training driver (no GPU): random crops are drawn over
the whole of each volume, whatever its shape, and pair input with target voxel for
voxel; with a validation score the training keeps its best-scoring parameters."""
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


class _FakeNet:
    """Stands in for the network: a parameter vector that every step moves by one."""

    def __init__(self):
        self.p = np.zeros(3, np.float32)

    def train_step(self, x, t, lr):
        self.p += 1.
        return float(lr)

    def params(self):
        return self.p.copy()

    def set_params(self, p):
        self.p = np.asarray(p, np.float32).copy()


def test_train_keeps_the_best_scoring_checkpoint():
    vol = torch.zeros((40, 40, 40))
    net = _FakeNet()
    # scores at steps 3, 6, 9, 10 (the last step is always scored): lowest at step 6
    scores = {3: 0.5, 6: 0.1, 9: 0.3, 10: 0.2}
    losses = W.train(net, [(vol, vol)], 10, crop=8, batch=1, val=lambda n: scores[int(n.p[0])], val_every=3)
    assert len(losses) == 10
    assert np.all(net.p == 6.)
    # without a score function the last parameters stay
    net = _FakeNet()
    W.train(net, [(vol, vol)], 10, crop=8, batch=1)
    assert np.all(net.p == 10.)
