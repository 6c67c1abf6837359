"""The plain-FDK configuration at its full size: RTK's analytic Shepp-Logan phantom
scaled to a 64 mm radius, projected in FP64 over 180 views (360 deg, SID 1000 mm,
SDD 1536 mm) onto a 256 x 256 detector at 0.8 mm, reconstructed on a 128^3 volume
at 1 mm centred at the isocentre with a Hann-windowed ramp (cut 0.8).  The HIP FDK
must agree with the oracle's FDK in float and in double within 1e-4 relative L2."""
import numpy as np
import pytest

from helpers import hip_geometry, rel_l2, to_device

pytestmark = pytest.mark.gpu


def test_fdk_c1_shape_vs_float_and_double_oracle(oracle, hip):
    from rtk_amd import phantoms
    n = 180
    rows = np.array([[1000., 1536., i * 360. / n, 0, 0, 0, 0, 0, 0] for i in range(n)])
    q = phantoms.shepp_logan(64.)
    sorig, ssp = (-102., -102., 0.), (0.8, 0.8, 1.)
    os_ = oracle.image((256, 256, n), sorig, ssp)
    proj_d = oracle.project_quadrics(rows, os_, np.zeros((n, 256, 256)), q)  # FP64 analytic line integrals
    proj = proj_d.astype(np.float32)
    vorig = (-63.5,) * 3
    ov = oracle.image((128,) * 3, vorig, (1.,) * 3)
    ref_f = oracle.fdk(rows, ov, np.zeros((128,) * 3, np.float32), os_, proj, hann=0.8, gpf=13)
    ref_d = oracle.fdk(rows, ov, np.zeros((128,) * 3), os_, proj.astype(np.float64), hann=0.8, gpf=13)
    f = hip.HipFDKConeBeamReconstructionFilter()
    f.SetInput(0, hip.ConstantImageSource((128, 128, 128), (1., 1., 1.), vorig, 0.))
    f.SetInput(1, to_device(hip, proj, sorig, ssp))
    f.SetGeometry(hip_geometry(hip, rows))
    f.GetRampFilter().SetHannCutFrequency(0.8)
    f.Update()
    out = f.GetOutput().numpy()
    ef, ed = rel_l2(out, ref_f), rel_l2(out, ref_d)
    print(f"C1 FDK: rel L2 vs float oracle {ef:.3g}, vs double oracle {ed:.3g}")
    assert ef <= 1e-4 and ed <= 1e-4
