"""FP64 oracle of the warm-start CNN (test infrastructure only).

The network has no counterpart in the reference; this module is the correctness
reference for the HIP kernels of rtk_amd/csrc/cnn.hip.  Convolutions are cross-correlations with zero
padding, tap t = (dz*3 + dy)*3 + dx, offset (dx-1, dy-1, dz-1); arrays are
channels-last ``[b, z, y, x, c]``.  Pinned by finite differences of its own forward
(tests/test_oracle_cnn.py) and by torch's float64 CPU conv3d as an independent route.
"""
import numpy as np

C = 32
T = 27


def layout():
    """(name, offset, shape) of the packed FP32 parameter vector."""
    out, o = [], 0
    for name, shape in [("w1", (T, C)), ("b1", (C,))] + \
            [(f"{k}{l}", s) for l in range(2, 6) for k, s in (("w", (T, C, C)), ("b", (C,)))] + \
            [("w6", (T, C)), ("b6", (1,))]:
        n = int(np.prod(shape))
        out.append((name, o, shape))
        o += n
    return out, o


def unpack(p):
    lay, n = layout()
    assert p.size == n
    return {name: p[o:o + int(np.prod(s))].reshape(s).astype(np.float64) for name, o, s in lay}


def pack(d):
    lay, n = layout()
    p = np.zeros(n)
    for name, o, s in lay:
        p[o:o + int(np.prod(s))] = np.asarray(d[name], dtype=np.float64).ravel()
    return p


def init_params(seed, scale=1.0):
    """He-style random initialisation (last layer small: the net starts near identity)."""
    rng = np.random.default_rng(seed)
    d = {"w1": rng.normal(0, np.sqrt(2 / 27), (T, C)), "b1": np.zeros(C)}
    for l in range(2, 6):
        d[f"w{l}"] = rng.normal(0, np.sqrt(2 / (27 * C)), (T, C, C))
        d[f"b{l}"] = np.zeros(C)
    d["w6"] = rng.normal(0, 1e-3 * scale, (T, C))
    d["b6"] = np.zeros(1)
    return pack(d)


def _taps():
    for t in range(T):
        yield t, (t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1)


def _shift(a, d):
    """a[b, z, y, x, ...] sampled at (x+dx, y+dy, z+dz) with zeros outside."""
    dx, dy, dz = d
    out = np.zeros_like(a)
    Z, Y, X = a.shape[1:4]

    def sl(delta, n):
        return (slice(max(0, -delta), n - max(0, delta)), slice(max(0, delta), n - max(0, -delta)))

    (oz, iz), (oy, iy), (ox, ix) = sl(dz, Z), sl(dy, Y), sl(dx, X)
    out[:, oz, oy, ox] = a[:, iz, iy, ix]
    return out


def conv(a, w, b):
    """a [B,Z,Y,X,Ci], w [T,Co,Ci] (or [T,Co] for Ci=1 as a[...,None]), b [Co]."""
    out = np.zeros(a.shape[:4] + (w.shape[1],)) + b
    for t, d in _taps():
        out += np.einsum("bzyxi,oi->bzyxo", _shift(a, d), w[t])
    return out


def forward(p, x, scale):
    """x [B,Z,Y,X] FP64 -> (y, activations).  y = x + conv6(...)/scale."""
    q = unpack(p)
    acts = []
    a = conv((scale * x)[..., None], q["w1"][..., None], q["b1"])
    a = np.maximum(a, 0.)
    acts.append(a)
    for l in range(2, 6):
        a = np.maximum(conv(a, q[f"w{l}"], q[f"b{l}"]), 0.)
        acts.append(a)
    r = conv(a, q["w6"][:, None, :], q["b6"])[..., 0]
    return x + r / scale, acts


def loss_and_grad(p, x, target, scale):
    """L = mean((scale (y - target))^2) and dL/dp (packed), FP64, by the adjoint."""
    q = unpack(p)
    y, acts = forward(p, x, scale)
    n = y.size
    d = y - target
    loss = np.mean((scale * d) ** 2)
    gy = 2 * scale ** 2 * d / n
    g = {}
    gr = gy / scale  # d/d(residual)
    a5 = acts[-1]
    g["b6"] = np.array([gr.sum()])
    g["w6"] = np.stack([np.einsum("bzyx,bzyxi->i", gr, _shift(a5, dd)) for t, dd in _taps()])
    # data gradient of layer 6: dA5[u] = sum_t w6[t] gr[u - d(t)]
    da = np.zeros_like(a5)
    for t, (dx, dy, dz) in _taps():
        da += _shift(gr[..., None], (-dx, -dy, -dz)) * q["w6"][t]
    dz_ = da * (a5 > 0)
    for l in range(5, 1, -1):
        ain = acts[l - 2]
        g[f"b{l}"] = dz_.sum(axis=(0, 1, 2, 3))
        g[f"w{l}"] = np.stack([np.einsum("bzyxo,bzyxi->oi", dz_, _shift(ain, dd)) for t, dd in _taps()])
        da = np.zeros_like(ain)
        for t, (dx, dy, dz) in _taps():
            da += np.einsum("bzyxo,oi->bzyxi", _shift(dz_, (-dx, -dy, -dz)), q[f"w{l}"][t])
        dz_ = da * (ain > 0)
    xs = (scale * x)[..., None]
    g["b1"] = dz_.sum(axis=(0, 1, 2, 3))
    g["w1"] = np.stack([np.einsum("bzyxo,bzyx->o", dz_, _shift(xs, dd)[..., 0]) for t, dd in _taps()])
    return loss, pack(g)
