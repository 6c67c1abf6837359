"""The applications end to end: projections written as one MetaImage file per view
and an RTK geometry XML file go through rtkfdk / rtkconjugategradient; the volumes
they write equal those of the filters run directly on the same data, and the FDK
one matches the double oracle within 1e-4."""
import numpy as np
import pytest

from helpers import hip_geometry, rel_l2, to_device

pytestmark = pytest.mark.gpu

N, NU = 48, 96
ROWS = np.array([[1000., 1536., i * 360. / N, 0, 0, 0, 0, 0, 0] for i in range(N)])
SORIG, SSP = (-(NU - 1) * 1.6 / 2, -(NU - 1) * 1.6 / 2), (1.6, 1.6)


@pytest.fixture(scope="module")
def dataset(tmp_path_factory, oracle, hip):
    from rtk_amd import io, phantoms
    d = tmp_path_factory.mktemp("acq")
    q = phantoms.shepp_logan(60.)
    os_ = oracle.image((NU, NU, N), SORIG + (0.,), SSP + (1.,))
    proj = oracle.project_quadrics(ROWS, os_, np.zeros((N, NU, NU)), q).astype(np.float32)
    for i in range(N):
        io.write_image(str(d / f"proj{i:03d}.mha"), proj[i], SORIG, SSP)
    io.write_geometry(hip_geometry(hip, ROWS), str(d / "geometry.xml"))
    return d, proj


def test_rtkfdk_cli(dataset, oracle, hip):
    from rtk_amd import apps, io
    d, proj = dataset
    out = str(d / "fdk.mha")
    assert apps.rtkfdk(["-g", str(d / "geometry.xml"), "-p", str(d), "-r", r"proj\d+\.mha", "-o", out,
                        "--hann", "0.8", "--size", "64", "--spacing", "2"]) == 0
    vol, o, s = io.read_image(out)
    assert vol.shape == (64, 64, 64) and o == (-63.,) * 3 and s == (2.,) * 3
    f = hip.HipFDKConeBeamReconstructionFilter()
    f.SetInput(0, hip.ConstantImageSource((64,) * 3, (2.,) * 3, (-63.,) * 3))
    f.SetInput(1, to_device(hip, proj, SORIG + (0.,), SSP + (1.,)))
    f.SetGeometry(hip_geometry(hip, ROWS))
    f.GetRampFilter().SetHannCutFrequency(0.8)
    f.SetProjectionSubsetSize(16)  # rtkfdk's default --subsetsize: same summation order
    f.Update()
    assert np.array_equal(vol, f.GetOutput().numpy())
    ref = oracle.fdk(ROWS, oracle.image((64,) * 3, (-63.,) * 3, (2.,) * 3), np.zeros((64,) * 3),
                     oracle.image((NU, NU, N), SORIG + (0.,), SSP + (1.,)), proj.astype(np.float64), hann=0.8, gpf=13)
    assert rel_l2(vol, ref) <= 1e-4


def test_rtkconjugategradient_cli(dataset, hip):
    from rtk_amd import apps, io
    d, proj = dataset
    out = str(d / "cg.mha")
    assert apps.rtkconjugategradient(["-g", str(d / "geometry.xml"), "-p", str(d), "-r", r"proj\d+\.mha",
                                      "-o", out, "-n", "4", "--bp", "Joseph", "--size", "64", "--spacing", "2"]) == 0
    vol, _, _ = io.read_image(out)
    f = hip.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, hip.ConstantImageSource((64,) * 3, (2.,) * 3, (-63.,) * 3))
    f.SetInput(1, to_device(hip, proj, SORIG + (0.,), SSP + (1.,)))
    f.SetGeometry(hip_geometry(hip, ROWS))
    f.SetBackProjectionFilter(f.BP_JOSEPH)
    f.SetNumberOfIterations(4)
    f.Update()
    # FP32 atomics in the Joseph backprojector: equal to rounding, not bitwise
    assert rel_l2(vol, f.GetOutput().numpy()) <= 1e-5


def test_rtkfdk_cli_short_scan_displaced_detector(tmp_path, oracle, hip):
    """A 220 degree arc on a detector shifted by 40 mm: rtkfdk runs the displaced-
    detector and Parker weightings before FDK (applications/rtkfdk/rtkfdk.cxx:83-113)
    and writes what the filter chain gives; --nodisplaced skips the first one."""
    from rtk_amd import apps, io, phantoms
    n = 60
    rows = np.array([[1000., 1536., i * 220. / n, 40., 0, 0, 0, 0, 0] for i in range(n)])
    q = phantoms.shepp_logan(60.)
    os_ = oracle.image((NU, NU, n), SORIG + (0.,), SSP + (1.,))
    proj = oracle.project_quadrics(rows, os_, np.zeros((n, NU, NU)), q).astype(np.float32)
    for i in range(n):
        io.write_image(str(tmp_path / f"p{i:03d}.mha"), proj[i], SORIG, SSP)
    io.write_geometry(hip_geometry(hip, rows), str(tmp_path / "g.xml"))
    base = ["-g", str(tmp_path / "g.xml"), "-p", str(tmp_path), "-r", r"p\d+\.mha", "--size", "64", "--spacing", "2"]
    for nodisplaced in (False, True):
        out = str(tmp_path / f"fdk{int(nodisplaced)}.mha")
        assert apps.rtkfdk(base + ["-o", out] + (["--nodisplaced"] if nodisplaced else [])) == 0
        vol = io.read_image(out)[0]
        g = hip_geometry(hip, rows)
        ddf = hip.HipDisplacedDetectorImageFilter()
        ddf.SetInput(to_device(hip, proj, SORIG + (0.,), SSP + (1.,)))
        ddf.SetGeometry(g)
        ddf.SetDisable(nodisplaced)
        ddf.Update()
        assert ddf.GetOutput().size[0] == (NU if nodisplaced else 2 * NU)
        pssf = hip.HipParkerShortScanImageFilter()
        pssf.SetInput(ddf.GetOutput())
        pssf.SetGeometry(g)
        pssf.Update()
        assert pssf.IsShortScan()
        f = hip.HipFDKConeBeamReconstructionFilter()
        f.SetInput(0, hip.ConstantImageSource((64,) * 3, (2.,) * 3, (-63.,) * 3))
        f.SetInput(1, pssf.GetOutput())
        f.SetGeometry(g)
        f.SetProjectionSubsetSize(16)
        f.Update()
        assert np.array_equal(vol, f.GetOutput().numpy())


def test_cli_rejects_processing_outside_the_path(dataset, hip):
    from rtk_amd import apps
    d, _ = dataset
    base = ["-g", str(d / "geometry.xml"), "-p", str(d), "-r", r"proj\d+\.mha", "-o", str(d / "x.mha"),
            "--size", "16"]
    with pytest.raises(SystemExit, match="not supported"):
        apps.rtkfdk(base + ["--binning", "2", "2"])
    with pytest.raises(SystemExit, match="regularised"):
        apps.rtkconjugategradient(base + ["--tikhonov", "0.1"])
