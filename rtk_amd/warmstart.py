"""Learned CG warm start: a 3D residual CNN on the HIP back end.

The network maps an FDK volume x to y = x + conv6(relu(...relu(conv1(s x))...)) / s
(six 3x3x3 convolutions, 32 channels; BF16 MFMA kernels with FP32 accumulation, FP32
master weights and Adam, rtk_amd/csrc/cnn.hip).  CG then starts from y instead of
zero; plain CG from zero remains the fallback.  There is no counterpart in the
reference: RTK's CG accepts any input volume as X0 (the resume path of
applications/rtkconjugategradient/rtkconjugategradient.cxx:61-77), which is where the
prediction enters.
"""
import ctypes
import os

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

C, T = 32, 27


def default_weights_path():
    """The network trained by rtk_amd.warmstart_train on the sparse-view workload
    (256^3 @1 mm, 90 views of 512^2, I0 = 1e4), shipped with the package."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart_c4.npz")


def scale_weights_path():
    """The network the at-scale workload (512^3 @0.5 mm, 180 views of 1024^2 @0.4 mm,
    I0 = 5e3) uses by default, shipped with the package: trained on the sparse-view
    workload toward the 50-iteration CG volume (python -m rtk_amd.warmstart_train
    --cg-iter 50), the one that transferred best to the at-scale workload."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart_c5.npz")


def num_params():
    return lib().rtkhip_cnn_num_params()


def init_params(seed=0, out_scale=1e-3):
    """He-normal initialisation, last layer small (the network starts near identity)."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2 / 27), T * C), np.zeros(C)]
    for _ in range(4):
        parts += [rng.normal(0, np.sqrt(2 / (27 * C)), T * C * C), np.zeros(C)]
    parts += [rng.normal(0, out_scale, T * C), np.zeros(1)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == num_params()
    return p


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class WarmStartCNN:
    """Inference and training of the warm-start network (device-resident)."""

    def __init__(self, params=None, scale=50.0, seed=0):
        # the library is held by the instance: at interpreter exit the module's own
        # globals may already be cleared when the destructor runs
        self._dll = lib()
        self._h = self._dll.rtkhip_cnn_create()
        if not self._h:
            raise _lib.RtkHipError(_lib.RTKHIP_ERR_NO_DEVICE, lib().rtkhip_last_error().decode())
        self.scale = float(scale)
        self.step = 0
        self.set_params(init_params(seed) if params is None else params)

    def __del__(self):
        h = getattr(self, "_h", None)
        dll = getattr(self, "_dll", None)
        if h and dll is not None:
            dll.rtkhip_cnn_destroy(h)
            self._h = None

    def set_params(self, p):
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float32).ravel())
        if p.size != num_params():
            raise ValueError(f"expected {num_params()} parameters, got {p.size}")
        check(lib().rtkhip_cnn_set_params(self._h, p.ctypes.data_as(ctypes.c_void_p), _stream()))
        self.step = 0

    def params(self):
        p = np.empty(num_params(), np.float32)
        check(lib().rtkhip_cnn_get_params(self._h, p.ctypes.data_as(ctypes.c_void_p), None, _stream()))
        return p

    def grads(self):
        g = np.empty(num_params(), np.float32)
        check(lib().rtkhip_cnn_get_params(self._h, None, g.ctypes.data_as(ctypes.c_void_p), _stream()))
        return g

    @staticmethod
    def _batched(x):
        if x.dim() == 3:
            x = x[None]
        if x.dim() != 4 or not x.is_cuda or x.dtype != torch.float32:
            raise ValueError("expected a CUDA float32 tensor [B, Z, Y, X] or [Z, Y, X]")
        return x.contiguous()

    def forward(self, x):
        """Network output for x ([Z, Y, X] or [B, Z, Y, X]); same shape as x."""
        squeeze = x.dim() == 3
        xb = self._batched(x)
        self._x = xb  # kept alive for backward
        y = torch.empty_like(xb)
        dims = _lib.int3((xb.shape[3], xb.shape[2], xb.shape[1]))
        check(lib().rtkhip_cnn_forward(self._h, xb.shape[0], dims, ctypes.c_void_p(xb.data_ptr()),
                                       ctypes.c_void_p(y.data_ptr()), ctypes.c_float(self.scale), _stream()))
        self._y = y
        return y[0] if squeeze else y

    __call__ = forward

    def forward_tiled(self, x, tile=128, halo=8):
        """Inference on a [Z, Y, X] volume in overlapping blocks: each tile^3 block is
        run with `halo` voxels of context on every side that lies inside the volume
        (zero padding only at the volume's own faces), and its centre is kept.  Six
        3x3x3 layers see 6 voxels, so any halo >= 6 gives exactly the full-volume
        result with the activations of one block resident at a time."""
        if halo < 6:
            raise ValueError("halo must cover the network's receptive radius (6 voxels)")
        if x.dim() != 3:
            raise ValueError("forward_tiled takes one [Z, Y, X] volume")
        Z, Y, X = x.shape
        y = torch.empty_like(x)
        for z0 in range(0, Z, tile):
            for y0 in range(0, Y, tile):
                for x0 in range(0, X, tile):
                    z1, y1, x1 = min(z0 + tile, Z), min(y0 + tile, Y), min(x0 + tile, X)
                    za, ya, xa = max(z0 - halo, 0), max(y0 - halo, 0), max(x0 - halo, 0)
                    zb, yb, xb = min(z1 + halo, Z), min(y1 + halo, Y), min(x1 + halo, X)
                    out = self.forward(x[za:zb, ya:yb, xa:xb].contiguous())
                    y[z0:z1, y0:y1, x0:x1] = out[z0 - za:z1 - za, y0 - ya:y1 - ya, x0 - xa:x1 - xa]
        return y

    def loss_and_grads(self, x, target):
        """Forward + backward; returns (loss, gradient vector) without updating."""
        y = self.forward(x)
        t = self._batched(target)
        loss = ctypes.c_double()
        check(lib().rtkhip_cnn_backward(self._h, ctypes.c_void_p(self._y.data_ptr()), ctypes.c_void_p(t.data_ptr()),
                                        ctypes.byref(loss), _stream()))
        return loss.value, y

    def backward_to(self, target):
        """Backward pass of the last forward toward `target` (the gradient of
        mean((scale (y - target))^2)); returns that loss."""
        t = self._batched(target)
        loss = ctypes.c_double()
        check(lib().rtkhip_cnn_backward(self._h, ctypes.c_void_p(self._y.data_ptr()), ctypes.c_void_p(t.data_ptr()),
                                        ctypes.byref(loss), _stream()))
        return loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        """One Adam update from the last backward pass."""
        self.step += 1
        check(lib().rtkhip_cnn_adam_step(self._h, ctypes.c_float(lr), ctypes.c_float(beta1), ctypes.c_float(beta2),
                                         ctypes.c_float(eps), self.step, _stream()))

    def train_step(self, x, target, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        """One Adam step on mean((scale (net(x) - target))^2); returns the loss."""
        loss, _ = self.loss_and_grads(x, target)
        self.step += 1
        check(lib().rtkhip_cnn_adam_step(self._h, ctypes.c_float(lr), ctypes.c_float(beta1), ctypes.c_float(beta2),
                                         ctypes.c_float(eps), self.step, _stream()))
        return loss

    def save(self, path):
        np.savez(path, params=self.params(), scale=np.float64(self.scale))

    @classmethod
    def load(cls, path):
        d = np.load(path, allow_pickle=False)
        return cls(params=d["params"], scale=float(d["scale"]))
