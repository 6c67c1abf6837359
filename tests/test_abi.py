"""The C ABI library loads, exports every symbol include/rtkhip.h declares, and
maps host-detectable failures to status codes (no GPU work is launched here)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "rtkhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rtkhip_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    from rtk_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 25
    bound = {s[0] for s in _lib.SIGNATURES}
    for n in names:
        assert hasattr(L, n), n
        assert n in bound, n


def test_version_and_device_count():
    from rtk_amd import _lib
    L = _lib.lib()
    assert L.rtkhip_version().startswith(b"rtkhip")
    assert L.rtkhip_device_count() >= 0


def test_null_arguments_are_statuses_not_crashes():
    from rtk_amd import _lib
    L = _lib.lib()
    pd = _lib.int3((4, 4, 1))
    assert L.rtkhip_fdk_backproject(pd, pd, None, None, None, None, None, 0, None) == _lib.RTKHIP_ERR_INVALID_ARGUMENT
    assert b"null" in L.rtkhip_last_error()
    assert L.rtkhip_voxel_backproject(None, None, None, None, None, None, None, 0, None) == 1
    assert L.rtkhip_joseph_forward_project(None, None, None, None, None, 0., 1., None, None, None, None, None) == 1
    assert L.rtkhip_ramp_filter(None, 8, None, 0, None, None, None, None, None, None, 0, None) == 1
    assert L.rtkhip_cg_step(0, None, None, None, None, None, None) == 1
    assert L.rtkhip_cg_step_noalloc(0, None, None, None, None, None, None) == 1
    m, mp = _lib.dbl(np.zeros(12))
    # unknown dtype is rejected before any device call
    assert L.rtkhip_fdk_backproject(pd, pd, mp, None, ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), 7,
                                    None) == 1


def test_filter_level_geometry_errors():
    import rtk_amd
    from rtk_amd import _lib
    L = _lib.lib()
    ctx = L.rtkhip_context_create()
    try:
        vol = _lib.Image()
        st = _lib.Image()
        for i in range(3):
            vol.size[i] = 4
            st.size[i] = 4
            vol.spacing[i] = st.spacing[i] = 1.
        g = rtk_amd.ThreeDCircularProjectionGeometry()
        g.AddProjection(1000., 1500., 0.)
        # stack of 4 views but geometry of 1: the reference's mismatch exception
        rc = L.rtkhip_back_project(ctx, g.handle, ctypes.byref(vol), None, ctypes.c_void_p(8), ctypes.byref(st),
                                   ctypes.c_void_p(8), 0, 0, None)
        assert rc == _lib.RTKHIP_ERR_GEOMETRY
        assert b"Mismatch between the number of projections" in L.rtkhip_last_error()
        # missing geometry
        rc = L.rtkhip_back_project(ctx, None, ctypes.byref(vol), None, ctypes.c_void_p(8), ctypes.byref(st),
                                   ctypes.c_void_p(8), 0, 0, None)
        assert rc == _lib.RTKHIP_ERR_GEOMETRY
        # cylindrical detector (include/rtkFDKBackProjectionImageFilter.hxx:38-41)
        g2 = rtk_amd.ThreeDCircularProjectionGeometry()
        for a in range(4):
            g2.AddProjection(1000., 1500., 90. * a)
        g2.SetRadiusCylindricalDetector(1500.)
        rc = L.rtkhip_back_project(ctx, g2.handle, ctypes.byref(vol), None, ctypes.c_void_p(8), ctypes.byref(st),
                                   ctypes.c_void_p(8), 0, 1, None)
        assert rc == _lib.RTKHIP_ERR_UNSUPPORTED
        # HannY (2-D kernel) not supported: a clear status
        ro = _lib.RampOptions()
        ro.hann_cut_y = 0.1
        rc = L.rtkhip_fdk_reconstruct(ctx, g2.handle, ctypes.byref(vol), None, ctypes.c_void_p(8), ctypes.byref(st),
                                      ctypes.c_void_p(8), ctypes.byref(ro), 0, 0, None, None)
        assert rc in (_lib.RTKHIP_ERR_UNSUPPORTED,)
    finally:
        L.rtkhip_context_destroy(ctx)


def test_no_device_is_a_status():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from rtk_amd import _lib
    L = _lib.lib()
    pd = _lib.int3((4, 4, 1))
    m, mp = _lib.dbl(np.zeros(12))
    rc = L.rtkhip_fdk_backproject(pd, pd, mp, None, None, ctypes.c_void_p(8), ctypes.c_void_p(8), 0, None)
    assert rc == _lib.RTKHIP_ERR_NO_DEVICE


def test_joseph_rejects_degenerate_volumes_and_too_many_views():
    """Host checks before any device work: a Joseph walk needs two voxels along every
    axis for its bilinear taps, and one launch takes at most 65535 views."""
    from rtk_amd import _lib
    L = _lib.lib()
    m, mp = _lib.dbl(np.zeros(12))
    src, sp = _lib.dbl(np.zeros(3))
    vs, vsp = _lib.dbl(np.ones(3))
    dummy = ctypes.c_void_p(256)
    for vol, views, msg in [((1, 8, 8), 2, b"at least 2"), ((8, 8, 1), 2, b"at least 2"),
                            ((8, 8, 8), 70000, b"65535")]:
        rc = L.rtkhip_joseph_forward_project(_lib.int3((4, 4, views)), _lib.int3(vol), mp, sp, vsp, 0., 1., dummy,
                                             dummy, dummy, None, None)
        assert rc == _lib.RTKHIP_ERR_UNSUPPORTED and msg in L.rtkhip_last_error()
        rc = L.rtkhip_joseph_backproject(_lib.int3((4, 4, views)), _lib.int3(vol), mp, sp, vsp, 0., 1., dummy,
                                         dummy, dummy, _lib.F32, None)
        assert rc == _lib.RTKHIP_ERR_UNSUPPORTED and msg in L.rtkhip_last_error()
