"""RTK geometry XML and MetaImage I/O (CPU).

The geometry writer/reader follow src/rtkThreeDCircularProjectionGeometryXMLFile
{Writer,Reader}.cxx; the hand-written document below exercises the reader rules
those files define (global parameters, values persisting across projections,
case-insensitive names, the "Angle" alias, the matrix consistency check)."""
import os
import zlib

import numpy as np
import pytest


def _geom(rows):
    from rtk_amd import ThreeDCircularProjectionGeometry
    g = ThreeDCircularProjectionGeometry()
    for r in rows:
        g.AddProjection(*r)
    return g


def test_geometry_round_trip(tmp_path):
    from rtk_amd import io
    rows = np.array([[1000., 1536., 10. * i - 40, 2. - 0.5 * i, 1.5, 3. * i, -2. + i, 0.7, -0.3]
                     for i in range(12)])
    g = _geom(rows)
    g.SetCollimationOfLastProjection(-10., 12., -5., 6.)
    p = tmp_path / "g.xml"
    io.write_geometry(g, str(p))
    text = p.read_text()
    assert text.startswith('<?xml version="1.0"?>\n<!DOCTYPE RTKGEOMETRY>\n<RTKThreeDCircularGeometry version="3">')
    # constant parameters are written once, before the projections
    head = text.split("<Projection>")[0]
    assert "<SourceToIsocenterDistance>1000</SourceToIsocenterDistance>" in head
    assert "<ProjectionOffsetY>1.5</ProjectionOffsetY>" in head
    assert text.count("<Projection>") == 12 and text.count("<GantryAngle>") == 12
    h = io.read_geometry(str(p))
    r2 = h.rows()
    assert len(h) == 12
    # angles come back in [0, 360), everything else unchanged
    np.testing.assert_allclose(np.mod(rows[:, [2, 5, 6]], 360.), r2[:, [2, 5, 6]], atol=1e-12)
    np.testing.assert_allclose(rows[:, [0, 1, 3, 4, 7, 8]], r2[:, [0, 1, 3, 4, 7, 8]], atol=1e-12)
    for i in range(12):
        np.testing.assert_allclose(g.GetMatrix(i), h.GetMatrix(i), atol=1e-9)
    assert h.collimation[-1] == [-10., 12., -5., 6.]
    assert h.collimation[0] == [np.finfo(np.float64).max] * 4


def _doc(projections, version=3, head=""):
    return (f'<?xml version="1.0"?>\n<!DOCTYPE RTKGEOMETRY>\n<RTKThreeDCircularGeometry version="{version}">\n'
            f"{head}{projections}</RTKThreeDCircularGeometry>\n")


def test_reader_rules_of_the_reference(tmp_path):
    from rtk_amd import io
    ref = _geom([[800., 1200., 30., 0, 0, 0, 0, 0, 0], [800., 1200., 75., 4., 0, 0, 0, 0, 0],
                 [800., 1200., 120., 4., 0, 0, 0, 0, 0]])
    m = ref.GetMatrix(2)
    mat = "\n".join(" ".join(f"{v:.15g}" for v in m[j]) for j in range(3))
    doc = _doc("<Projection><GantryAngle>30</GantryAngle></Projection>\n"
               "<Projection><angle>75</angle><PROJECTIONOFFSETX>4</PROJECTIONOFFSETX></Projection>\n"
               f"<Projection><GantryAngle>120</GantryAngle><Matrix>{mat}</Matrix></Projection>\n",
               head="<SourceToIsocenterDistance>800</SourceToIsocenterDistance>"
                    "<SourceToDetectorDistance>1200</SourceToDetectorDistance>\n")
    p = tmp_path / "h.xml"
    p.write_text(doc)
    g = io.read_geometry(str(p))
    # ProjectionOffsetX set in projection 2 persists into projection 3
    np.testing.assert_allclose(g.rows(), ref.rows())
    # version 2 is read as well; others are refused
    p.write_text(_doc("<Projection><GantryAngle>0</GantryAngle></Projection>",
                      version=2, head="<SourceToIsocenterDistance>800</SourceToIsocenterDistance>"
                                      "<SourceToDetectorDistance>1200</SourceToDetectorDistance>"))
    assert len(io.read_geometry(str(p))) == 1
    p.write_text(_doc("<Projection><GantryAngle>0</GantryAngle></Projection>", version=1))
    with pytest.raises(ValueError, match="Incompatible version"):
        io.read_geometry(str(p))
    bad = mat.replace(mat.split()[3], str(float(mat.split()[3]) + 0.01), 1)
    p.write_text(_doc(f"<Projection><GantryAngle>120</GantryAngle><ProjectionOffsetX>4</ProjectionOffsetX>"
                      f"<Matrix>{bad}</Matrix></Projection>",
                      head="<SourceToIsocenterDistance>800</SourceToIsocenterDistance>"
                           "<SourceToDetectorDistance>1200</SourceToDetectorDistance>"))
    with pytest.raises(ValueError, match="not consistent"):
        io.read_geometry(str(p))


@pytest.mark.parametrize("ext", [".mha", ".mhd"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint16, np.int16])
def test_metaimage_round_trip(tmp_path, ext, dtype):
    from rtk_amd import io
    rng = np.random.default_rng(0)
    a = (rng.normal(size=(5, 7, 9)) * 100).astype(dtype)
    p = str(tmp_path / ("img" + ext))
    io.write_image(p, a, (-1.5, 2.25, 3.), (0.5, 0.75, 1.25))
    b, o, s = io.read_image(p)
    assert b.dtype == a.dtype and np.array_equal(a, b)
    assert o == (-1.5, 2.25, 3.) and s == (0.5, 0.75, 1.25)


def test_metaimage_compressed_big_endian_and_2d(tmp_path):
    from rtk_amd import io
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    hdr = ("ObjectType = Image\nNDims = 2\nBinaryData = True\nBinaryDataByteOrderMSB = True\n"
           "CompressedData = True\nOffset = 1 2\nElementSpacing = 0.4 0.4\nDimSize = 4 3\n"
           "ElementType = MET_FLOAT\nElementDataFile = LOCAL\n")
    p = tmp_path / "c.mha"
    p.write_bytes(hdr.encode() + zlib.compress(a.astype(">f4").tobytes()))
    b, o, s = io.read_image(str(p))
    assert np.array_equal(a, b) and o == (1., 2.) and s == (0.4, 0.4)
    p.write_bytes(hdr.replace("CompressedData = True", "TransformMatrix = 0 1 1 0").encode())
    with pytest.raises(ValueError, match="direction"):
        io.read_image(str(p))


def test_projection_file_order(tmp_path):
    from rtk_amd import io
    for n in (10, 9, 100, 1):
        (tmp_path / f"proj_{n}.mha").write_bytes(b"")
    (tmp_path / "other.txt").write_bytes(b"")
    names = [os.path.basename(f) for f in io.projection_files(str(tmp_path), r"proj_(\d+)\.mha")]
    assert names == ["proj_1.mha", "proj_10.mha", "proj_100.mha", "proj_9.mha"]
    names = [os.path.basename(f) for f in io.projection_files(str(tmp_path), r"proj_(\d+)\.mha", nsort=True,
                                                            submatch=1)]
    assert names == ["proj_1.mha", "proj_9.mha", "proj_10.mha", "proj_100.mha"]
