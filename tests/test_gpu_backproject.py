"""HIP voxel and FDK backprojection vs the oracle (the reference's CPU filters
restated), through the C ABI.  Bar: relative L2 <= 1e-5 against the float
oracle; the error against the double oracle is reported in the messages."""
import numpy as np
import pytest
import torch

from helpers import circular_rows, hip_geometry, rel_l2, to_device

pytestmark = pytest.mark.gpu

TOL = 1e-5

GEOMS = {
    "circular": circular_rows(23, 600., 1000., start=3.),
    "offsets": np.array([[600., 1000., 17 * i + 3, 2.5 * i - 4, -3. + 0.5 * i, 0, 0, 5. - 0.3 * i, 2.]
                         for i in range(23)]),
    # out-of-plane / in-plane angles: RTK's generic path (ITK linear interpolation)
    "tilted": np.array([[600., 1000., 16 * i, 1., -2., 4., 3., 0, 0] for i in range(23)]),
}


def _case(oracle, gname, seed=0, vsize=(37, 33, 29), index=(0, 0, 0), nu=61):
    rows = GEOMS[gname]
    vsp, vorig = (3., 2.5, 3.5), (-53., -40., -49.)
    nv, n = 47, len(rows)
    ssp, sorig = (3.2, 3.2), (-96., -73.)
    rng = np.random.default_rng(seed)
    proj = rng.normal(size=(n, nv, nu)).astype(np.float32)
    return rows, vsize, vsp, vorig, (nu, nv, n), ssp, sorig, proj


# nu = 64: rows of whole 16-byte chunks, the staging path with vector loads.  The
# volume's footprint crosses the detector's edges, so the column kernel's blocks are
# both inside the detector (no per-voxel test) and across its border.
@pytest.mark.parametrize("nu", [61, 64])
@pytest.mark.parametrize("fdk", [False, True])
@pytest.mark.parametrize("gname", list(GEOMS))
def test_backproject_parity(oracle, hip, gname, fdk, nu):
    rows, vsize, vsp, vorig, ssize, ssp, sorig, proj = _case(oracle, gname, nu=nu)
    rng = np.random.default_rng(1)
    v0 = rng.normal(size=vsize[::-1]).astype(np.float32)
    ov, os_ = oracle.image(vsize, vorig, vsp), oracle.image(ssize, sorig + (0.,), ssp + (1.,))
    ref_f = oracle.backproject(rows, ov, v0, os_, proj, fdk=fdk)
    ref_d = oracle.backproject(rows, ov, v0.astype(np.float64), os_, proj.astype(np.float64), fdk=fdk)
    g = hip_geometry(hip, rows)
    f = hip.HipFDKBackProjectionImageFilter() if fdk else hip.HipBackProjectionImageFilter()
    f.SetInput(0, to_device(hip, v0, vorig, vsp))
    f.SetInput(1, to_device(hip, proj, sorig + (0.,), ssp + (1.,)))
    f.SetGeometry(g)
    f.Update()
    out = f.GetOutput().numpy()
    e_f, e_d = rel_l2(out - v0, ref_f - v0), rel_l2(out - v0, ref_d - v0)
    assert e_f <= TOL, f"{gname} fdk={fdk}: rel L2 vs float oracle {e_f:.3g} (vs double {e_d:.3g})"


def test_buffered_region_offsets_and_view_subset(oracle, hip):
    # Volume with a non-zero buffered index, stack holding views 5..14 of a
    # 23-view geometry (as the FDK subset extraction produces).
    rows = GEOMS["offsets"]
    vsize, vsp, vorig, vindex = (20, 18, 16), (3., 3., 3.), (-60., -50., -45.), (7, 4, 9)
    nu, nv = 61, 47
    proj = np.random.default_rng(2).normal(size=(10, nv, nu)).astype(np.float32)
    # Oracle: equivalent image with the index folded into the origin, and a
    # geometry truncated to the stack's views.
    oorig = tuple(o + i * s for o, i, s in zip(vorig, vindex, vsp))
    ov = oracle.image(vsize, oorig, vsp)
    os_ = oracle.image((nu, nv, 10), (-96., -73., 0.), (3.2, 3.2, 1.))
    ref = oracle.backproject(rows[5:15], ov, np.zeros(vsize[::-1], np.float32), os_, proj, fdk=True)
    g = hip_geometry(hip, rows)
    f = hip.HipFDKBackProjectionImageFilter()
    f.SetInput(0, to_device(hip, np.zeros(vsize[::-1]), vorig, vsp, index=vindex))
    f.SetInput(1, to_device(hip, proj, (-96., -73., 0.), (3.2, 3.2, 1.), index=(0, 0, 5)))
    f.SetGeometry(g)
    f.Update()
    assert rel_l2(f.GetOutput().numpy(), ref) <= TOL


def test_in_place_and_empty_stack(oracle, hip):
    rows = GEOMS["circular"]
    g = hip_geometry(hip, rows)
    v0 = np.random.default_rng(3).normal(size=(29, 33, 37)).astype(np.float32)
    vol = to_device(hip, v0, (-53., -40., -49.), (3., 2.5, 3.5))
    proj = to_device(hip, np.ones((23, 47, 61)), (-96., -73., 0.), (3.2, 3.2, 1.))
    f = hip.HipBackProjectionImageFilter()
    f.SetInput(0, vol)
    f.SetInput(1, proj)
    f.SetGeometry(g)
    f.InPlaceOn()
    f.Update()
    assert f.GetOutput().tensor.data_ptr() == vol.tensor.data_ptr()
    # constant-1 projections add integer view counts
    d = vol.numpy() - v0
    assert np.all(np.abs(d - np.round(d)) < 2e-4) and abs(d.max() - 23) < 2e-4
    # zero views: output equals input
    empty = to_device(hip, np.zeros((1, 47, 61)), (-96., -73., 0.), (3.2, 3.2, 1.))
    empty.tensor = empty.tensor[:0]
    f2 = hip.HipBackProjectionImageFilter()
    f2.SetInput(0, vol)
    f2.SetInput(1, empty)
    f2.SetGeometry(g)
    with pytest.raises(hip.RtkHipError) as e:
        f2.Update()  # a 0-size image has no valid descriptor
    assert e.value.code == 1


@pytest.mark.parametrize("dtype,tol", [(torch.float16, 2e-3), (torch.bfloat16, 1.5e-2)])
def test_half_precision_stacks(oracle, hip, dtype, tol):
    """Projection stack stored in FP16/BF16 (FP32 math): error against the double oracle."""
    rows, vsize, vsp, vorig, ssize, ssp, sorig, proj = _case(oracle, "circular", seed=4)
    # smooth projections (what a ramp-filtered real stack looks like at this scale)
    u = np.linspace(-3, 3, ssize[0])
    proj = (np.cos(u)[None, None, :] * np.linspace(1, 2, ssize[1])[None, :, None]
            * np.ones((ssize[2], 1, 1))).astype(np.float32)
    ov, os_ = oracle.image(vsize, vorig, vsp), oracle.image(ssize, sorig + (0.,), ssp + (1.,))
    ref = oracle.backproject(rows, ov, np.zeros(vsize[::-1]), os_, proj.astype(np.float64), fdk=True)
    g = hip_geometry(hip, rows)
    f = hip.HipFDKBackProjectionImageFilter()
    f.SetInput(0, to_device(hip, np.zeros(vsize[::-1]), vorig, vsp))
    f.SetInput(1, to_device(hip, proj, sorig + (0.,), ssp + (1.,), dtype=dtype))
    f.SetGeometry(g)
    f.Update()
    e = rel_l2(f.GetOutput().numpy(), ref)
    assert e <= tol, f"{dtype}: rel L2 {e:.3g}"


def test_full_size_constant_counts(oracle, hip):
    """512^3-scale property at 1/4 linear size: constant-1 projections give, per voxel,
    the number of views whose footprint passes the border rule (exact integers)."""
    n = 90
    rows = circular_rows(n, 1000., 1536.)
    g = hip_geometry(hip, rows)
    vol = hip.ConstantImageSource((128, 128, 128), (2., 2., 2.), (-127., -127., -127.), 0.)
    proj = hip.ConstantImageSource((256, 256, n), (1.6, 1.6, 1.), (-204., -204., 0.), 1.)
    f = hip.HipBackProjectionImageFilter()
    f.SetInput(0, vol)
    f.SetInput(1, proj)
    f.SetGeometry(g)
    f.Update()
    got = f.GetOutput().numpy()
    ov, os_ = oracle.image((128,) * 3, (-127.,) * 3, (2.,) * 3), oracle.image((256, 256, n), (-204., -204., 0),
                                                                            (1.6, 1.6, 1.))
    ref = oracle.backproject(rows, ov, np.zeros((128,) * 3, np.float32), os_, np.ones((n, 256, 256), np.float32))
    assert np.all(np.abs(got - np.round(ref)) < 1e-3)
