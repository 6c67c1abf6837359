"""Parity at the full benchmark shapes.

* FDK and the plain voxel backprojection at 512^3 @0.5 mm <- 720 x 1024^2 @0.4 mm
  (SID 1000, SDD 1536, per-view detector offsets jittered +-2 mm, analytic
  projections of a seeded 20-ellipsoid phantom): the whole volume is reconstructed
  on the GPU; three slabs of two y slices (y is the rotation axis: bottom edge,
  centre, upper half) are recomputed by the oracle from the detector rows they
  project onto.  The ramp kernel is 1-D, so detector rows are filtered
  independently, and the FDK weight depends on the pixel position only: the
  oracle's result from the row band is its result from the whole stack for those
  slices, to the rounding of the pixel coordinates (tests/test_oracle_fdk_cg.py).
* Joseph forward projection of the voxelised phantom (512^3) onto all 720 views;
  two views are recomputed by the oracle.
* Joseph backprojection at the sparse-view CG shape (256^3 @1 mm <- 90 x 512^2
  @0.8 mm): the adjoint identity with the forward projector over all 90 views, and
  the oracle on four of the views at the full per-view shape.
Bars: FDK 1e-4 relative L2 per slab; Joseph and the voxel backprojection 1e-5.
"""
import numpy as np
import pytest
import torch

from helpers import hip_geometry, rel_l2, row_band

pytestmark = pytest.mark.gpu

NX, NU, NV = 512, 1024, 720
VSP, PSP = 0.5, 0.4
VORG = -(NX - 1) * VSP / 2
PORG = -(NU - 1) * PSP / 2
SLABS = (0, 255, 383)  # first y index of each 2-slice slab


@pytest.fixture(scope="module")
def full(hip):
    rng = np.random.default_rng(20260001)
    offs = rng.uniform(-2., 2., size=(NV, 2))
    rows = np.array([[1000., 1536., i * 360. / NV, offs[i, 0], offs[i, 1], 0., 0., 0., 0.] for i in range(NV)])
    g = hip_geometry(hip, rows)
    q = hip.phantoms.random_ellipsoids(20260001, n=20, half_fov=NX * VSP / 2)
    proj = hip.project_quadrics(g, hip.ConstantImageSource((NU, NU, NV), (PSP, PSP, 1.), (PORG, PORG, 0.)), q)
    return rows, g, q, proj


def _slab_case(oracle, rows, proj, j):
    """Oracle images for the y slices j, j + 1 and the detector rows they reach."""
    y0 = VORG + j * VSP
    r0, r1 = row_band(oracle, rows, (VORG, y0, VORG), (-VORG, y0 + VSP, -VORG), PORG, PSP, NU)
    band = np.ascontiguousarray(proj.tensor[:, r0:r1, :].cpu().numpy())
    st = oracle.image((NU, r1 - r0, NV), (PORG, PORG + r0 * PSP, 0.), (PSP, PSP, 1.))
    ov = oracle.image((NX, 2, NX), (VORG, y0, VORG), (VSP,) * 3)
    return st, ov, band, (r0, r1)


def test_fdk_full_size_slabs(oracle, hip, full):
    rows, g, _, proj = full
    f = hip.HipFDKConeBeamReconstructionFilter()
    f.SetInput(0, hip.ConstantImageSource((NX,) * 3, (VSP,) * 3, (VORG,) * 3, 0.))
    f.SetInput(1, proj)
    f.SetGeometry(g)
    f.Update()
    out = f.GetOutput().tensor
    assert bool(torch.isfinite(out).all())
    for k in SLABS:
        st, ov, band, rb = _slab_case(oracle, rows, proj, k)
        ref = oracle.fdk(rows, ov, np.zeros((NX, 2, NX), np.float32), st, band, threads=16, gpf=13)
        got = out[:, k:k + 2, :].cpu().numpy()
        e = rel_l2(got, ref)
        print(f"FDK slab y={k}: rows {rb}, rel L2 {e:.3g}", flush=True)
        assert e <= 1e-4, (k, rb, e)


def test_voxel_backprojection_full_size_slab(oracle, hip, full):
    rows, g, _, proj = full
    f = hip.HipBackProjectionImageFilter()
    f.SetInput(0, hip.ConstantImageSource((NX,) * 3, (VSP,) * 3, (VORG,) * 3, 0.))
    f.SetInput(1, proj)
    f.SetGeometry(g)
    f.Update()
    k = SLABS[1]
    st, ov, band, rb = _slab_case(oracle, rows, proj, k)
    ref = oracle.backproject(rows, ov, np.zeros((NX, 2, NX), np.float32), st, band, threads=16)
    e = rel_l2(f.GetOutput().tensor[:, k:k + 2, :].cpu().numpy(), ref)
    print(f"voxel BP slab y={k}: rows {rb}, rel L2 {e:.3g}", flush=True)
    assert e <= 1e-5, (rb, e)


def test_joseph_forward_full_size_views(oracle, hip, full):
    rows, g, q, _ = full
    vol = hip.draw_quadrics(hip.ConstantImageSource((NX,) * 3, (VSP,) * 3, (VORG,) * 3, 0.), q)
    f = hip.HipForwardProjectionImageFilter()
    f.SetInput(0, hip.ConstantImageSource((NU, NU, NV), (PSP, PSP, 1.), (PORG, PORG, 0.)))
    f.SetInput(1, vol)
    f.SetGeometry(g)
    f.Update()
    out = f.GetOutput().tensor
    v = vol.tensor.cpu().numpy()
    ov = oracle.image((NX,) * 3, (VORG,) * 3, (VSP,) * 3)
    st = oracle.image((NU, NU, 1), (PORG, PORG, 0.), (PSP, PSP, 1.))
    for k in (0, 413):
        ref = oracle.joseph(rows, st, np.zeros((1, NU, NU), np.float32), ov, v, threads=16, geom_offset=k)
        e = rel_l2(out[k:k + 1].cpu().numpy(), ref)
        print(f"Joseph FP view {k}: rel L2 {e:.3g}", flush=True)
        assert e <= 1e-5, (k, e)


def test_joseph_backprojection_sparse_view_shape(oracle, hip):
    n, nx, nu = 90, 256, 512
    rows = np.array([[1000., 1536., i * 360. / n, 0, 0, 0, 0, 0, 0] for i in range(n)])
    g = hip_geometry(hip, rows)
    vorg, porg = -(nx - 1) / 2., -(nu - 1) * 0.8 / 2
    gen = torch.Generator(device="cuda").manual_seed(5)
    p = torch.rand((n, nu, nu), device="cuda", generator=gen) + 0.5
    x = torch.rand((nx,) * 3, device="cuda", generator=gen)
    proj = hip.Image(p, (porg, porg, 0.), (0.8, 0.8, 1.))
    bp = hip.HipJosephBackProjectionImageFilter()
    bp.SetInput(0, hip.ConstantImageSource((nx,) * 3, (1.,) * 3, (vorg,) * 3, 0.))
    bp.SetInput(1, proj)
    bp.SetGeometry(g)
    bp.Update()
    fp = hip.HipForwardProjectionImageFilter()
    fp.SetInput(0, hip.ConstantImageSource((nu, nu, n), (0.8, 0.8, 1.), (porg, porg, 0.)))
    fp.SetInput(1, hip.Image(x, (vorg,) * 3, (1.,) * 3))
    fp.SetGeometry(g)
    fp.Update()
    a = float((fp.GetOutput().tensor.double() * p.double()).sum())
    b = float((x.double() * bp.GetOutput().tensor.double()).sum())
    print(f"Joseph adjoint at 256^3 x 90 x 512^2: {a:.9g} vs {b:.9g}", flush=True)
    assert abs(a / b - 1) < 1e-5, (a, b)
    # the oracle on four views at the full per-view shape
    views = [0, 11, 45, 78]
    sub = hip_geometry(hip, rows[views])
    ps = p[views].contiguous()
    bp4 = hip.HipJosephBackProjectionImageFilter()
    bp4.SetInput(0, hip.ConstantImageSource((nx,) * 3, (1.,) * 3, (vorg,) * 3, 0.))
    bp4.SetInput(1, hip.Image(ps, (porg, porg, 0.), (0.8, 0.8, 1.)))
    bp4.SetGeometry(sub)
    bp4.Update()
    ref = oracle.joseph_back(rows[views], oracle.image((nx,) * 3, (vorg,) * 3, (1.,) * 3),
                             np.zeros((nx,) * 3, np.float32), oracle.image((nu, nu, 4), (porg, porg, 0.), (0.8, 0.8, 1.)),
                             ps.cpu().numpy(), threads=16)
    e = rel_l2(bp4.GetOutput().tensor.cpu().numpy(), ref)
    print(f"Joseph BP, 4 views at 256^3 <- 512^2: rel L2 {e:.3g}", flush=True)
    assert e <= 1e-5, e
