"""CG with the projection stacks the backprojector reads stored in FP16 / BF16
(FP32 projector arithmetic, FP64 CG scalars), for RTK's default pair (Joseph FP +
voxel BP) and for the matched Joseph pair, against the FP64 oracle running the
same pair in double.  The 16-bit rounding of the stacks is the only extra error;
the bars bound it (FP16 keeps 11 significant bits, BF16 8)."""
import numpy as np
import pytest

from helpers import hip_geometry, rel_l2

pytestmark = pytest.mark.gpu

N = 60
ROWS = np.array([[600., 1200., i * 360. / N, 0, 0, 0, 0, 0, 0] for i in range(N)])


@pytest.fixture(scope="module")
def case(oracle, hip):
    from rtk_amd import phantoms
    q = np.concatenate([phantoms.ellipsoid((0, 0, 0), (90, 90, 90), None, 1.0)[None],
                        phantoms.ellipsoid((20, -10, 15), (30, 20, 25), None, 0.5)[None]])
    g = hip_geometry(hip, ROWS)
    proj = hip.project_quadrics(g, hip.ConstantImageSource((64, 64, N), (8., 8., 8.), (-255.,) * 3, 0.), q)
    ov = oracle.image((64,) * 3, (-127.,) * 3, (4.,) * 3)
    os_ = oracle.image((64, 64, N), (-255.,) * 3, (8.,) * 3)
    refs = {}
    for jbp in (False, True):
        refs[jbp] = oracle.conjugate_gradient(ROWS, ov, np.zeros((64,) * 3), os_, proj.numpy().astype(np.float64), 8,
                                              joseph_bp=jbp)[0]
    return g, proj, refs


@pytest.mark.parametrize("dtype,tol", [("f32", 1e-4), ("f16", 1e-3), ("bf16", 1e-2)])
@pytest.mark.parametrize("bp", ["voxel", "joseph"])
def test_cg_stack_dtype_vs_double_oracle(hip, case, dtype, tol, bp):
    g, proj, refs = case
    f = hip.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, hip.ConstantImageSource((64, 64, 64), (4., 4., 4.), (-127.,) * 3, 0.))
    f.SetInput(1, proj)
    f.SetGeometry(g)
    if bp == "joseph":
        f.SetBackProjectionFilter(f.BP_JOSEPH)
    f.SetStackDtype(dtype)
    f.SetNumberOfIterations(8)
    f.Update()
    e = rel_l2(f.GetOutput().numpy(), refs[bp == "joseph"])
    print(f"CG {bp} BP, {dtype} stacks: rel L2 vs FP64 oracle {e:.3g}")
    assert e <= tol


def test_joseph_bp_16bit_projections(oracle, hip):
    """The Joseph backprojector reading FP16/BF16 projections equals the FP32 one on
    the same rounded values (the loader is the only difference)."""
    import torch
    from helpers import to_device
    rng = np.random.default_rng(4)
    p = rng.normal(size=(N, 64, 64)).astype(np.float32)
    out = {}
    for dt in (torch.float16, torch.bfloat16):
        pr = torch.from_numpy(p).to(dt)
        f = hip.HipJosephBackProjectionImageFilter()
        f.SetInput(0, hip.ConstantImageSource((64, 64, 64), (4., 4., 4.), (-127.,) * 3, 0.))
        f.SetInput(1, hip.Image(pr.cuda(), (-255.,) * 3, (8.,) * 3))
        f.SetGeometry(hip_geometry(hip, ROWS))
        f.Update()
        f32 = hip.HipJosephBackProjectionImageFilter()
        f32.SetInput(0, hip.ConstantImageSource((64, 64, 64), (4., 4., 4.), (-127.,) * 3, 0.))
        f32.SetInput(1, to_device(hip, pr.float().numpy(), (-255.,) * 3, (8.,) * 3))
        f32.SetGeometry(hip_geometry(hip, ROWS))
        f32.Update()
        out[dt] = rel_l2(f.GetOutput().numpy(), f32.GetOutput().numpy())
    assert max(out.values()) <= 1e-6, out
