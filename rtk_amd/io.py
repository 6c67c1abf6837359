"""The reference's on-disk formats for this path: RTK's circular geometry XML
(RTKThreeDCircularGeometry, version 3) and MetaImage images (.mha with the data
inline, .mhd with a separate raw file).

Geometry XML follows src/rtkThreeDCircularProjectionGeometryXMLFileWriter.cxx:42-183
(what is written, in which order, angles in degrees in [0, 360), parameters that are
equal for every projection written once before the projections, defaults omitted,
15 significant digits) and src/rtkThreeDCircularProjectionGeometryXMLFileReader.cxx:
48-189 (version 2 or 3, case-insensitive names, "Angle" as an alias of "GantryAngle",
parameter values persisting from one projection to the next, the stored matrix
checked against the one recomputed from the parameters within 1e-3).
"""
import math
import os
import re
import xml.etree.ElementTree as ET
import zlib

import numpy as np

from .filters import ThreeDCircularProjectionGeometry

GEOMETRY_VERSION = 3
_NO_COLLIMATION = np.finfo(np.float64).max

# (XML name, geometry row index, written in degrees)
_ROW_PARAMS = [("SourceToIsocenterDistance", 0, False), ("SourceToDetectorDistance", 1, False),
               ("SourceOffsetX", 7, False), ("SourceOffsetY", 8, False),
               ("ProjectionOffsetX", 3, False), ("ProjectionOffsetY", 4, False),
               ("InPlaneAngle", 6, True), ("OutOfPlaneAngle", 5, True)]
_COLLIMATION = ["CollimationUInf", "CollimationUSup", "CollimationVInf", "CollimationVSup"]


def _deg_0_360(a_deg):
    """Angle as the reference stores it (radians in [0, 2 pi)), written back in degrees."""
    r = math.radians(a_deg)
    r -= 2 * math.pi * math.floor(r / (2 * math.pi))
    return r * 45. / math.atan(1.)


def _num(v):
    # 15 significant digits as the reference writes; the "no collimation" default
    # (largest double) with 17, since 15 digits round it above the largest double
    return f"{v:.17g}" if v == _NO_COLLIMATION else f"{v:.15g}"


def _parse(text):
    v = float(text)
    return _NO_COLLIMATION if v == math.inf else v


def write_geometry(geometry, path):
    """rtk::ThreeDCircularProjectionGeometryXMLFileWriter::WriteFile."""
    rows = np.asarray(geometry.rows(), dtype=np.float64)
    if rows.size == 0:
        raise ValueError("Geometry object is empty, cannot write it")
    coll = np.asarray(geometry.collimation, dtype=np.float64).reshape(-1, 4)
    ind = "  "
    lines = ['<?xml version="1.0"?>', "<!DOCTYPE RTKGEOMETRY>",
             f'<RTKThreeDCircularGeometry version="{GEOMETRY_VERSION}">']
    glob = {}
    for name, col, deg in _ROW_PARAMS:
        v = rows[:, col]
        if np.all(v == v[0]):
            glob[name] = True
            val = _deg_0_360(v[0]) if deg else v[0]
            if v[0] != 0.:
                lines.append(f"{ind * 2}<{name}>{_num(val)}</{name}>")
    for k, name in enumerate(_COLLIMATION):
        v = coll[:, k]
        if np.all(v == v[0]):
            glob[name] = True
            if v[0] != _NO_COLLIMATION:
                lines.append(f"{ind * 2}<{name}>{_num(v[0])}</{name}>")
    radius = geometry.GetRadiusCylindricalDetector()
    if radius != 0.:
        lines.append(f"{ind * 2}<RadiusCylindricalDetector>{_num(radius)}</RadiusCylindricalDetector>")
    for i, r in enumerate(rows):
        lines.append(f"{ind}<Projection>")
        lines.append(f"{ind * 2}<GantryAngle>{_num(_deg_0_360(r[2]))}</GantryAngle>")
        for name, col, deg in _ROW_PARAMS:
            if name not in glob:
                val = _deg_0_360(r[col]) if deg else r[col]
                lines.append(f"{ind * 2}<{name}>{_num(val)}</{name}>")
        for k, name in enumerate(_COLLIMATION):
            if name not in glob:
                lines.append(f"{ind * 2}<{name}>{_num(coll[i, k])}</{name}>")
        m = geometry.GetMatrix(i)
        lines.append(f"{ind * 2}<Matrix>")
        for j in range(3):
            lines.append(ind * 3 + " ".join(f"{m[j][k]:>19.15g}" for k in range(4)))
        lines.append(f"{ind * 2}</Matrix>")
        lines.append(f"{ind}</Projection>")
    lines.append("</RTKThreeDCircularGeometry>")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_geometry(path):
    """rtk::ReadGeometry: returns a ThreeDCircularProjectionGeometry."""
    root = ET.parse(path).getroot()
    if root.tag.lower() != "rtkthreedcirculargeometry":
        raise ValueError(f"{path}: not an RTKThreeDCircularGeometry file")
    version = int(root.attrib.get("version", 0))
    if version != GEOMETRY_VERSION and version != 2:
        raise ValueError(f"Incompatible version of input geometry (v{version}) with current geometry "
                         f"(v{GEOMETRY_VERSION}). You must re-generate your geometry file again.")
    state = {n.lower(): 0. for n, _, _ in _ROW_PARAMS}
    state["gantryangle"] = 0.
    for n in _COLLIMATION:
        state[n.lower()] = _NO_COLLIMATION
    g = ThreeDCircularProjectionGeometry()
    matrix = None

    def take(el):
        nonlocal matrix
        tag = el.tag.lower()
        text = (el.text or "").strip()
        if tag == "angle":
            tag = "gantryangle"
        if tag in state:
            state[tag] = _parse(text)
        elif tag == "radiuscylindricaldetector":
            g.SetRadiusCylindricalDetector(float(text))
        elif tag == "matrix":
            vals = [float(t) for t in text.split()]
            matrix = np.array(vals[:12]).reshape(3, 4) if len(vals) >= 12 else None

    for el in root:
        if el.tag.lower() != "projection":
            take(el)
            continue
        matrix = None
        for sub in el:
            take(sub)
        g.AddProjection(state["sourcetoisocenterdistance"], state["sourcetodetectordistance"], state["gantryangle"],
                        state["projectionoffsetx"], state["projectionoffsety"], state["outofplaneangle"],
                        state["inplaneangle"], state["sourceoffsetx"], state["sourceoffsety"])
        g.SetCollimationOfLastProjection(*[state[n.lower()] for n in _COLLIMATION])
        if matrix is not None:
            computed = np.asarray(g.GetMatrix(len(g) - 1))
            if np.any(np.abs(matrix - computed) > 0.001):
                raise ValueError("Matrix and parameters are not consistent.\nRead matrix from geometry file:\n"
                                 f"{matrix}\nComputed matrix from parameters:\n{computed}")
    return g


# ---------------------------------------------------------------- MetaImage

_MET_TYPES = {"MET_FLOAT": np.float32, "MET_DOUBLE": np.float64, "MET_CHAR": np.int8, "MET_UCHAR": np.uint8,
              "MET_SHORT": np.int16, "MET_USHORT": np.uint16, "MET_INT": np.int32, "MET_UINT": np.uint32,
              "MET_LONG_LONG": np.int64, "MET_ULONG_LONG": np.uint64}


def read_image(path):
    """MetaImage reader: returns (array [z][y][x] or [y][x], origin, spacing) with the
    file's element type.  Non-identity direction cosines and multi-component pixels
    are rejected (the projectors assume axis-aligned images)."""
    with open(path, "rb") as f:
        raw = f.read()
    header, pos = {}, 0
    while True:
        end = raw.index(b"\n", pos)
        line = raw[pos:end].decode("ascii", "replace").strip()
        pos = end + 1
        if "=" in line:
            k, v = (t.strip() for t in line.split("=", 1))
            header[k] = v
            if k == "ElementDataFile":
                break
    ndims = int(header["NDims"])
    dims = [int(t) for t in header["DimSize"].split()][:ndims]
    spacing = [float(t) for t in header.get("ElementSpacing", header.get("ElementSize", "1 " * ndims)).split()]
    origin = [float(t) for t in header.get("Offset", header.get("Origin", header.get("Position", "0 " * ndims))).split()]
    tm = header.get("TransformMatrix", header.get("Rotation", header.get("Orientation")))
    if tm is not None and not np.allclose(np.array([float(t) for t in tm.split()]).reshape(ndims, ndims),
                                          np.eye(ndims)):
        raise ValueError(f"{path}: direction cosines other than identity are not supported")
    if int(header.get("ElementNumberOfChannels", "1")) != 1:
        raise ValueError(f"{path}: multi-component pixels are not supported")
    etype = header["ElementType"]
    if etype not in _MET_TYPES:
        raise ValueError(f"{path}: unsupported element type {etype}")
    dt = np.dtype(_MET_TYPES[etype])
    msb = header.get("BinaryDataByteOrderMSB", header.get("ElementByteOrderMSB", "False")).lower() == "true"
    dt = dt.newbyteorder(">" if msb else "<")
    src = header["ElementDataFile"]
    if src == "LOCAL":
        data = raw[pos:]
    else:
        with open(os.path.join(os.path.dirname(path), src), "rb") as f:
            data = f.read()
        skip = int(header.get("HeaderSize", "0"))
        data = data[skip:] if skip > 0 else data
    if header.get("CompressedData", "False").lower() == "true":
        data = zlib.decompress(data)
    n = int(np.prod(dims))
    arr = np.frombuffer(data, dtype=dt, count=n).astype(dt.newbyteorder("="))
    return arr.reshape(dims[::-1]), tuple(origin[:ndims]), tuple(spacing[:ndims])


def write_image(path, array, origin, spacing):
    """MetaImage writer (.mha inline data; .mhd with a .raw file beside it)."""
    arr = np.ascontiguousarray(array)
    names = {np.dtype(v): k for k, v in _MET_TYPES.items()}
    if arr.dtype not in names:
        raise ValueError(f"unsupported dtype {arr.dtype}")
    nd = arr.ndim
    inline = not path.lower().endswith(".mhd")
    datafile = "LOCAL" if inline else os.path.splitext(os.path.basename(path))[0] + ".raw"
    hdr = ["ObjectType = Image", f"NDims = {nd}", "BinaryData = True", "BinaryDataByteOrderMSB = False",
           "CompressedData = False",
           "TransformMatrix = " + " ".join("1" if i == j else "0" for i in range(nd) for j in range(nd)),
           "Offset = " + " ".join(_num(float(o)) for o in origin),
           "CenterOfRotation = " + " ".join("0" for _ in range(nd)),
           "ElementSpacing = " + " ".join(_num(float(s)) for s in spacing),
           "DimSize = " + " ".join(str(d) for d in arr.shape[::-1]),
           f"ElementType = {names[arr.dtype]}", f"ElementDataFile = {datafile}"]
    body = arr.astype(arr.dtype.newbyteorder("<"), copy=False).tobytes()
    with open(path, "wb") as f:
        f.write(("\n".join(hdr) + "\n").encode("ascii"))
        if inline:
            f.write(body)
    if not inline:
        with open(os.path.join(os.path.dirname(path), datafile), "wb") as f:
            f.write(body)


def projection_files(path, regexp, nsort=False, submatch=0):
    """Files of `path` whose name matches `regexp` (searched anywhere in the name), in
    lexicographic order or, with nsort, by the numeric value of group `submatch`
    (itk::RegularExpressionSeriesFileNames as the reference's projections reader uses it)."""
    rx = re.compile(regexp)
    hits = []
    for name in os.listdir(path):
        m = rx.search(name)
        full = os.path.join(path, name)
        if m and os.path.isfile(full):
            hits.append((m, full))
    if nsort:
        def key(h):
            num = re.search(r"[-+]?\d+(\.\d+)?", h[0].group(submatch))
            return (float(num.group()) if num else float("inf"), h[1])
        hits.sort(key=key)
    else:
        hits.sort(key=lambda h: (h[0].group(submatch), h[1]))
    return [h[1] for h in hits]
