"""HIP warm-start CNN (BF16 MFMA convolutions, FP32 accumulation) vs the FP64
oracle: forward residual and parameter gradients, plus a training smoke test.
Tolerances are those of BF16 activations/weights (8-bit mantissa)."""
import numpy as np
import pytest
import torch

from helpers import rel_l2

pytestmark = pytest.mark.gpu

SCALE = 50.0


def _problem(seed, shape=(2, 7, 10, 21)):
    from oracle import cnn_ref
    rng = np.random.default_rng(seed)
    p = cnn_ref.init_params(seed, scale=SCALE * 20)  # a residual branch that matters
    x = (0.02 + 0.01 * rng.normal(size=shape)).astype(np.float32)
    t = (x + 0.002 * rng.normal(size=shape)).astype(np.float32)
    return p, x, t


def test_cnn_forward_matches_oracle(hip):
    from oracle import cnn_ref
    from rtk_amd.warmstart import WarmStartCNN
    p, x, _ = _problem(3)
    net = WarmStartCNN(params=p, scale=SCALE)
    y = net(torch.from_numpy(x).cuda()).cpu().numpy()
    yr, _ = cnn_ref.forward(p.astype(np.float64), x.astype(np.float64), SCALE)
    res, res_ref = (y - x) * SCALE, (yr - x) * SCALE
    e = rel_l2(res, res_ref)
    assert e < 3e-2, f"residual branch rel L2 {e:.3g}"
    np.testing.assert_allclose(net.params(), p.astype(np.float32), rtol=0, atol=0)


def test_cnn_gradients_match_oracle(hip):
    from oracle import cnn_ref
    from rtk_amd.warmstart import WarmStartCNN
    p, x, t = _problem(4)
    net = WarmStartCNN(params=p, scale=SCALE)
    loss, _ = net.loss_and_grads(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    g = net.grads()
    lr, gr = cnn_ref.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), SCALE)
    assert abs(loss - lr) < 5e-2 * abs(lr)
    lay, _ = cnn_ref.layout()
    for name, o, s in lay:
        n = int(np.prod(s))
        a, b = g[o:o + n], gr[o:o + n]
        if np.linalg.norm(b) == 0:
            continue
        e = rel_l2(a, b)
        assert e < 6e-2, f"{name}: gradient rel L2 {e:.3g}"


def test_cnn_training_reduces_loss(hip):
    from rtk_amd.warmstart import WarmStartCNN
    rng = np.random.default_rng(5)
    x = torch.from_numpy((0.02 + 0.01 * rng.normal(size=(2, 16, 16, 32))).astype(np.float32)).cuda()
    # learnable target: a local smoothing of the input
    k = torch.ones(1, 1, 3, 3, 3, device="cuda") / 27
    t = torch.nn.functional.conv3d(x[:, None], k, padding=1)[:, 0].contiguous()
    net = WarmStartCNN(scale=SCALE, seed=1)
    losses = [net.train_step(x, t, lr=2e-3) for _ in range(60)]
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])


def test_tiled_inference_equals_full_volume(hip):
    """Overlapping tiles with an 8-voxel halo (the large-volume inference path)
    reproduce whole-volume inference bit for bit: the 6-voxel receptive field is
    covered and every voxel's sums run in the same order."""
    from oracle import cnn_ref
    from rtk_amd.warmstart import WarmStartCNN
    p = cnn_ref.init_params(5, scale=SCALE * 20)
    net = WarmStartCNN(params=p, scale=SCALE)
    rng = np.random.default_rng(6)
    x = torch.from_numpy((0.02 + 0.01 * rng.normal(size=(45, 70, 83))).astype(np.float32)).cuda()
    full = net(x)
    tiled = net.forward_tiled(x, tile=32, halo=8)
    assert torch.equal(full, tiled)
    with pytest.raises(ValueError):
        net.forward_tiled(x, tile=32, halo=4)
