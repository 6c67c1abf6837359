"""The kernel-level entry points called the way RTK's CUDA filters would call them.

Each test builds the per-view arrays with the formulas of the CUDA host code it
replaces, in numpy and in FP64, from the closed-form geometry of
documentation/docs/Geometry.md (the matrices the oracle's geometry is pinned to), and
with non-zero buffered-region indices on the stack and the volume:

  * rtkhip_fdk_backproject: src/rtkCudaFDKBackProjectionImageFilter.cxx:45-100
    (index-to-index matrix, region-index corrections, division by the perspective
    factor at the rotation centre), without the CUDA texture's 0.5 offset;
  * rtkhip_voxel_backproject: include/rtkCudaBackProjectionImageFilter.hxx:90-130
    (the same product, also normalised by the perspective factor there);
  * rtkhip_joseph_forward_project / rtkhip_joseph_backproject:
    include/rtkCudaForwardProjectionImageFilter.hxx:100-165 (projection index ->
    volume index matrices and source positions in volume indices), without the 0.5
    corner-convention offset;
  * rtkhip_fdk_weight: src/rtkCudaFDKWeightProjectionFilter.cxx:84-107 (7 values per
    view; the angular weight factor of include/rtkFDKWeightProjectionFilter.hxx:40-63);
  * rtkhip_dot against an FP64 numpy dot product.

Results are compared with the oracle's filters on the same images, with the region
index folded into the oracle image's origin.  Bars: 1e-5 relative L2 (the projector
parity bar), 1e-6 for the weighting."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import circular_rows, rel_l2

pytestmark = pytest.mark.gpu

NV = 24
ROWS = circular_rows(NV, sid=900., sdd=1400., start=7., proj_off=[(3. * np.sin(i), -2. * np.cos(i)) for i in range(NV)],
                     src_off=(4., -3.))
# stack: 72 x 56 pixels of 1.2 mm, buffered region index (2, 5, 0)
ST_SIZE, ST_SP, ST_IDX = (72, 56, NV), (1.2, 1.2, 1.), (2, 5, 0)
ST_ORG = (-(71 + 4) * 0.6, -(55 + 10) * 0.6, 0.)
# volume: 40 x 36 x 44 voxels of 2 mm, buffered region index (3, 0, 2)
VO_SIZE, VO_SP, VO_IDX = (40, 36, 44), (2., 2., 2.), (3, 0, 2)
VO_ORG = (-45., -35., -47.)


def _rot(ga, oop, ip):
    d = np.pi / 180
    c, s = np.cos(-ga * d), np.sin(-ga * d)
    ry = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    c, s = np.cos(-oop * d), np.sin(-oop * d)
    rx = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    c, s = np.cos(-ip * d), np.sin(-ip * d)
    rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    r = np.eye(4)
    r[:3, :3] = rz @ rx @ ry
    return r


def _proj_matrix(sid, sdd, ga, px, py, oop, ip, sx, sy):
    """3x4 projection matrix T(s - p) Mag T(-s) R of documentation/docs/Geometry.md."""
    t2 = np.array([[1, 0, sx - px], [0, 1, sy - py], [0, 0, 1.]])
    mag = np.array([[-sdd, 0, 0, 0], [0, -sdd, 0, 0], [0, 0, 1, -sid]], dtype=float)
    ts = np.eye(4)
    ts[0, 3], ts[1, 3] = -sx, -sy
    return t2 @ mag @ ts @ _rot(ga, oop, ip)


def _index_to_pp(org, sp):
    m = np.eye(4)
    for i in range(3):
        m[i, i], m[i, 3] = sp[i], org[i]
    return m


def _fold(org, sp, idx):
    """Origin of the same pixels seen as an image whose buffer starts at index 0."""
    return tuple(o + i * s for o, i, s in zip(org, idx, sp))


def bp_matrices(normalise):
    """The CUDA FDK / voxel backprojection host matrices (no texture offset)."""
    out = []
    vol_idx_to_pp = _index_to_pp(VO_ORG, VO_SP)
    pp_to_idx = np.array([[1 / ST_SP[0], 0, -ST_ORG[0] / ST_SP[0]], [0, 1 / ST_SP[1], -ST_ORG[1] / ST_SP[1]],
                          [0, 0, 1.]])
    idx_proj = np.eye(3)
    idx_proj[0, 2], idx_proj[1, 2] = -ST_IDX[0], -ST_IDX[1]
    idx_vol = np.eye(4)
    idx_vol[:3, 3] = VO_IDX
    # continuous index of the physical origin, relative to the buffered region
    rot_centre = np.array([-o / s - i for o, s, i in zip(VO_ORG, VO_SP, VO_IDX)])
    for r in ROWS:
        m = idx_proj @ pp_to_idx @ _proj_matrix(*r) @ vol_idx_to_pp @ idx_vol
        if normalise:
            m = m / (m[2, 3] + m[2, :3] @ rot_centre)
        out.append(m)
    return np.ascontiguousarray(np.array(out))


def joseph_arrays():
    """Projection index -> volume index matrices and source positions (volume indices)."""
    vol_pp_to_idx = np.linalg.inv(_index_to_pp(VO_ORG, VO_SP))
    vol_tr = np.eye(4)
    vol_tr[:3, 3] = [-i for i in VO_IDX]
    proj_tr = np.eye(4)
    proj_tr[:3, 3] = ST_IDX
    st_idx_to_pp = _index_to_pp(ST_ORG, ST_SP)
    mats, srcs = [], []
    for sid, sdd, ga, px, py, oop, ip, sx, sy in ROWS:
        rinv = np.linalg.inv(_rot(ga, oop, ip))
        det = np.eye(4)
        det[0, 3], det[1, 3], det[2, 3], det[2, 2] = px, py, sid - sdd, 0.
        m = vol_tr @ vol_pp_to_idx @ rinv @ det @ st_idx_to_pp @ proj_tr
        mats.append(m[:3])
        s = vol_tr @ vol_pp_to_idx @ rinv @ np.array([sx, sy, sid, 1.])
        srcs.append(s[:3])
    return np.ascontiguousarray(np.array(mats)), np.ascontiguousarray(np.array(srcs))


def _oracle_images(oracle):
    vol = oracle.image(VO_SIZE, _fold(VO_ORG, VO_SP, VO_IDX), VO_SP)
    st = oracle.image(ST_SIZE, _fold(ST_ORG, ST_SP, ST_IDX), ST_SP)
    return vol, st


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _data(seed):
    rng = np.random.default_rng(seed)
    proj = (1 + 0.3 * rng.normal(size=(NV, ST_SIZE[1], ST_SIZE[0]))).astype(np.float32)
    vol = (rng.uniform(0, 1, size=VO_SIZE[::-1])).astype(np.float32)
    return proj, vol


@pytest.mark.parametrize("fdk", [True, False])
def test_backprojectors_with_cuda_host_matrices(oracle, hip, fdk):
    from rtk_amd import _lib
    L = _lib.lib()
    proj, vol0 = _data(1)
    # both CUDA filters normalise by the perspective factor; the plain BP is also
    # given unnormalised matrices (its result does not depend on the scale of a row)
    variants = [True] if fdk else [True, False]
    ovol, ost = _oracle_images(oracle)
    ref = oracle.backproject(ROWS, ovol, vol0, ost, proj, fdk=fdk)
    p = torch.from_numpy(proj).cuda()
    vin = torch.from_numpy(vol0).cuda()
    for normalise in variants:
        m, mp = _lib.dbl(bp_matrices(normalise).ravel())
        out = torch.empty_like(vin)
        fn = L.rtkhip_fdk_backproject if fdk else L.rtkhip_voxel_backproject
        _lib.check(fn(_lib.int3(ST_SIZE), _lib.int3(VO_SIZE), mp, None, _ptr(vin), _ptr(out), _ptr(p), _lib.F32,
                      None))
        torch.cuda.synchronize()
        e = rel_l2(out.cpu().numpy(), ref)
        print(f"{'FDK' if fdk else 'voxel'} BP, {'normalised' if normalise else 'raw'} matrices: rel L2 {e:.3g}")
        assert e <= 1e-5


def test_joseph_pair_with_cuda_host_matrices(oracle, hip):
    from rtk_amd import _lib
    L = _lib.lib()
    proj, vol = _data(2)
    mats, srcs = joseph_arrays()
    m, mp = _lib.dbl(mats.ravel())
    s, sp = _lib.dbl(srcs.ravel())
    spacing, spp = _lib.dbl(np.array(VO_SP))
    ovol, ost = _oracle_images(oracle)
    ref_fp = oracle.joseph(ROWS, ost, proj, ovol, vol)
    ref_bp = oracle.joseph_back(ROWS, ovol, vol, ost, proj)
    p = torch.from_numpy(proj).cuda()
    v = torch.from_numpy(vol).cuda()
    out = torch.empty_like(p)
    _lib.check(L.rtkhip_joseph_forward_project(_lib.int3(ST_SIZE), _lib.int3(VO_SIZE), mp, sp, spp, 0., 1., _ptr(p),
                                               _ptr(out), _ptr(v), None, None))
    vout = torch.empty_like(v)
    _lib.check(L.rtkhip_joseph_backproject(_lib.int3(ST_SIZE), _lib.int3(VO_SIZE), mp, sp, spp, 0., 1., _ptr(v),
                                           _ptr(vout), _ptr(p), _lib.F32, None))
    torch.cuda.synchronize()
    e_fp, e_bp = rel_l2(out.cpu().numpy(), ref_fp), rel_l2(vout.cpu().numpy(), ref_bp)
    print(f"Joseph FP rel L2 {e_fp:.3g}, Joseph BP rel L2 {e_bp:.3g}")
    assert e_fp <= 1e-5 and e_bp <= 1e-5


def test_fdk_weight_with_cuda_host_geometry(oracle, hip):
    from rtk_amd import _lib
    L = _lib.lib()
    proj, _ = _data(3)
    _, src, angles, gaps, tilt = oracle.geometry(ROWS)
    g7 = []
    for k, (sid, sdd, ga, px, py, oop, ip, sx, sy) in enumerate(ROWS):
        # include/rtkFDKWeightProjectionFilter.hxx:40-63: gap * |sdd| / (2 sid^2) * |source|
        factor = gaps[k] * abs(sdd) / (2 * sid * sid) * np.linalg.norm(src[k])
        g7 += [sdd, sid, px, py, sy, factor, tilt[k]]
    gd, gp = _lib.dbl(np.array(g7))
    o2, op = _lib.dbl(np.array(_fold(ST_ORG, ST_SP, ST_IDX)[:2]))
    s2, sp2 = _lib.dbl(np.array(ST_SP[:2]))
    ovol, ost = _oracle_images(oracle)
    ref = oracle.fdk_weight(ROWS, ost, proj)
    p = torch.from_numpy(proj).cuda()
    out = torch.empty_like(p)
    _lib.check(L.rtkhip_fdk_weight(_lib.int3(ST_SIZE), gp, op, sp2, _ptr(p), _ptr(out), None))
    torch.cuda.synchronize()
    e = rel_l2(out.cpu().numpy(), ref)
    print(f"FDK weighting rel L2 {e:.3g}")
    assert e <= 1e-6


@pytest.mark.parametrize("n", [0, 1, 1000, 2 ** 20 + 3, 3 * 2 ** 22 + 5])
def test_dot_against_fp64_numpy(hip, n):
    from rtk_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(n)
    a = rng.normal(size=n).astype(np.float32)
    b = rng.normal(size=n).astype(np.float32)
    # one spare element keeps the device pointers non-null for n = 0 (an empty sum: 0)
    da = torch.from_numpy(np.append(a, np.float32(7.))).cuda()
    db = torch.from_numpy(np.append(b, np.float32(7.))).cuda()
    r = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    _lib.check(L.rtkhip_dot(n, _ptr(da), _ptr(db), _ptr(r), None))
    torch.cuda.synchronize()
    ref = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
    got = float(r.item())
    # FP64 sums of FP32 products (exact in FP64) in another order: rounding only
    assert abs(got - ref) <= 1e-12 * max(1., float(np.abs(a.astype(np.float64) * b).sum())), (got, ref)
