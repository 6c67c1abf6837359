"""The image data manager of the C ABI (rtkhip_buffer_*): a CG reconstruction chained
through device buffers (forward projection -> multiplication by the weights ->
backprojection -> CG update, RTK's operator of
include/rtkReconstructionConjugateGradientOperator.hxx:194-286 driven as the
ConjugateGradientImageFilter loop does) uploads each host image once and copies
nothing else until the result is read back; the result equals the CG filter's on the
same inputs.  Also the manager's state rules: a host write re-uploads on the next device
access, a device write re-downloads on the next host access, and nothing moves when the
accessed side is current."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import circular_rows, hip_geometry, rel_l2

pytestmark = pytest.mark.gpu


class Buf:
    def __init__(self, L, nbytes, host=None):
        self.L = L
        self.h = L.rtkhip_buffer_create(nbytes)
        assert self.h, L.rtkhip_last_error()
        self.host = host
        if host is not None:
            assert L.rtkhip_buffer_set_host(self.h, host.ctypes.data_as(ctypes.c_void_p)) == 0

    def dev(self, write=False):
        p = ctypes.c_void_p()
        assert self.L.rtkhip_buffer_device(self.h, int(write), ctypes.byref(p), None) == 0
        return p

    def read(self, write=False):
        p = ctypes.c_void_p()
        assert self.L.rtkhip_buffer_host(self.h, int(write), ctypes.byref(p), None) == 0
        return p

    def transfers(self):
        u, d = ctypes.c_longlong(), ctypes.c_longlong()
        assert self.L.rtkhip_buffer_transfers(self.h, ctypes.byref(u), ctypes.byref(d)) == 0
        return u.value, d.value

    def __del__(self):
        self.L.rtkhip_buffer_destroy(self.h)


def test_buffer_state_rules(hip):
    from rtk_amd import _lib
    L = _lib.lib()
    a = np.arange(1000, dtype=np.float32)
    b = Buf(L, a.nbytes, a)
    assert b.transfers() == (0, 0)
    d = b.dev()
    assert b.transfers() == (1, 0)
    b.dev()
    b.read()  # both sides current: nothing moves
    assert b.transfers() == (1, 0)
    t = torch.empty(1000, device="cuda")
    # a device write (doubling through the multiply entry point) makes the host stale
    two = torch.full((1000,), 2., device="cuda")
    assert L.rtkhip_multiply(1000, b.dev(), ctypes.c_void_p(two.data_ptr()), b.dev(write=True), None) == 0
    torch.cuda.synchronize()
    b.read()
    assert b.transfers() == (1, 1) and np.array_equal(a, 2 * np.arange(1000, dtype=np.float32))
    a[:] = 7.  # a host write, declared through a write access, re-uploads on the next device access
    b.read(write=True)
    b.dev()
    assert b.transfers() == (2, 1)
    torch.cuda.synchronize()
    del t


def test_cg_chain_through_buffers_copies_nothing_after_the_first_upload(hip):
    from rtk_amd import _lib, phantoms
    L = _lib.lib()
    n, nu, nv, nx = 40, 48, 40, 32
    rows = circular_rows(n, 500., 900.)
    g = hip_geometry(hip, rows)
    st = hip.ConstantImageSource((nu, nv, n), (6., 6., 1.), (-(nu - 1) * 3., -(nv - 1) * 3., 0.))
    vol = hip.ConstantImageSource((nx,) * 3, (5., 5., 5.), (-(nx - 1) * 2.5,) * 3)
    q = phantoms.random_ellipsoids(7, n=6, half_fov=70.)
    p_host = hip.project_quadrics(g, st, q).numpy()
    w_host = (0.5 + np.random.default_rng(1).random(p_host.shape)).astype(np.float32)
    x_host = np.zeros((nx,) * 3, np.float32)
    nvox, npix = x_host.size, p_host.size
    X, Pj, W = Buf(L, x_host.nbytes, x_host), Buf(L, p_host.nbytes, p_host), Buf(L, w_host.nbytes, w_host)
    T = Buf(L, p_host.nbytes)
    B, AX, R, P, AP = (Buf(L, x_host.nbytes) for _ in range(5))
    S = Buf(L, 4 * 8)
    c = hip.filters._context()
    sd, vd = st.descriptor(), vol.descriptor()

    def fp(src_vol, dst_stack):
        _lib.check(L.rtkhip_forward_project(c, g.handle, ctypes.byref(sd), None, dst_stack, ctypes.byref(vd), src_vol,
                                            None, None))

    def bp(src_stack, dst_vol):
        _lib.check(L.rtkhip_back_project(c, g.handle, ctypes.byref(vd), None, dst_vol, ctypes.byref(sd), src_stack,
                                         _lib.F32, 0, None))

    # B = BP(W p); R = B - BP(W FP(X)); P = R
    _lib.check(L.rtkhip_multiply(npix, W.dev(), Pj.dev(), T.dev(True), None))
    bp(T.dev(), B.dev(True))
    fp(X.dev(), T.dev(True))
    _lib.check(L.rtkhip_multiply(npix, W.dev(), T.dev(), T.dev(True), None))
    bp(T.dev(), AX.dev(True))
    _lib.check(L.rtkhip_cg_init(nvox, B.dev(), AX.dev(), R.dev(True), P.dev(True), S.dev(True), None))
    for _ in range(4):
        fp(P.dev(), T.dev(True))
        _lib.check(L.rtkhip_multiply(npix, W.dev(), T.dev(), T.dev(True), None))
        bp(T.dev(), AP.dev(True))
        _lib.check(L.rtkhip_cg_step(nvox, X.dev(True), R.dev(True), P.dev(True), AP.dev(), S.dev(True), None))
    torch.cuda.synchronize()
    # one upload of each host image, nothing else
    assert X.transfers() == (1, 0) and Pj.transfers() == (1, 0) and W.transfers() == (1, 0)
    for b in (T, B, AX, R, P, AP, S):
        assert b.transfers() == (0, 0)
    X.read()
    assert X.transfers() == (1, 1)
    # the same four iterations through the CG filter
    f = hip.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, vol)
    f.SetInput(1, st.like(torch.from_numpy(p_host).cuda()))
    f.SetInputWeights(st.like(torch.from_numpy(w_host).cuda()))
    f.SetGeometry(g)
    f.SetNumberOfIterations(4)
    f.Update()
    e = rel_l2(x_host, f.GetOutput().numpy())
    print(f"buffer chain vs CG filter: rel L2 {e:.3g}")
    assert np.abs(x_host).max() > 0 and e <= 1e-5
