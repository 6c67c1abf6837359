"""Pins the FP64 CNN oracle (oracle/cnn_ref.py): its forward against torch's float64
CPU conv3d (an independent implementation) and its analytic gradient against
central finite differences of its own forward."""
import numpy as np
import torch

from oracle import cnn_ref


def _torch_forward(p, x, scale):
    q = cnn_ref.unpack(p)
    # torch conv3d weight [out, in, kz, ky, kx]; tap t = (dz*3 + dy)*3 + dx
    def w5(w):  # [T, Co, Ci] -> [Co, Ci, 3, 3, 3]
        return torch.from_numpy(w.reshape(3, 3, 3, w.shape[1], w.shape[2]).transpose(3, 4, 0, 1, 2).copy())

    a = torch.from_numpy(scale * x)[:, None]
    a = torch.relu(torch.nn.functional.conv3d(a, w5(q["w1"][..., None]), torch.from_numpy(q["b1"]), padding=1))
    for l in range(2, 6):
        a = torch.relu(torch.nn.functional.conv3d(a, w5(q[f"w{l}"]), torch.from_numpy(q[f"b{l}"]), padding=1))
    r = torch.nn.functional.conv3d(a, w5(q["w6"][:, None, :]), torch.from_numpy(q["b6"]), padding=1)[:, 0]
    return x + r.numpy() / scale


def test_forward_matches_torch_conv3d():
    rng = np.random.default_rng(0)
    p = cnn_ref.init_params(1, scale=50.)
    x = rng.normal(size=(2, 5, 6, 7)) * 0.02
    y, _ = cnn_ref.forward(p, x, 50.)
    np.testing.assert_allclose(y, _torch_forward(p, x, 50.), rtol=1e-10, atol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = cnn_ref.init_params(2, scale=50.)
    x = rng.normal(size=(1, 4, 5, 6)) * 0.02
    t = x + rng.normal(size=x.shape) * 0.001
    loss, g = cnn_ref.loss_and_grad(p, x, t, 50.)
    lay, _ = cnn_ref.layout()
    idx = []
    for name, o, s in lay:  # a few entries of every block
        n = int(np.prod(s))
        idx += list(o + rng.choice(n, size=min(n, 3), replace=False))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (cnn_ref.loss_and_grad(pp, x, t, 50.)[0] - cnn_ref.loss_and_grad(pm, x, t, 50.)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1., abs(fd)) + 1e-9, (i, fd, g[i])
