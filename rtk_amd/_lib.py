"""ctypes binding of ``librtkhip.so`` (the C ABI declared in ``include/rtkhip.h``).

The library is built in-tree by ``make -C rtk_amd`` (``__graft_entry__.build()``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtkhip.so")

RTKHIP_OK = 0
RTKHIP_ERR_INVALID_ARGUMENT = 1
RTKHIP_ERR_GEOMETRY = 2
RTKHIP_ERR_UNSUPPORTED = 3
RTKHIP_ERR_HIP = 4
RTKHIP_ERR_OUT_OF_MEMORY = 5
RTKHIP_ERR_NO_DEVICE = 6

F32, F16, BF16 = 0, 1, 2
BORDER_OPTIMIZED, BORDER_ITK_LINEAR = 0, 1
BP_VOXELBASED, BP_JOSEPH = 0, 1


class RtkHipError(RuntimeError):
    """Raised for a non-zero rtkhip status (RTK's itk::ExceptionObject counterpart)."""

    def __init__(self, code, message):
        super().__init__(f"rtkhip error {code}: {message}")
        self.code = code


class Image(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int * 3), ("index", ctypes.c_int * 3),
                ("origin", ctypes.c_double * 3), ("spacing", ctypes.c_double * 3)]


class RampOptions(ctypes.Structure):
    _fields_ = [("hann_cut", ctypes.c_double), ("hann_cut_y", ctypes.c_double),
                ("cosine_cut", ctypes.c_double), ("hamming", ctypes.c_double),
                ("ramlak_cut", ctypes.c_double), ("shepp_logan_cut", ctypes.c_double),
                ("truncation_correction", ctypes.c_double), ("zero_pad_x", ctypes.c_int),
                ("zero_pad_y", ctypes.c_int), ("greatest_prime_factor", ctypes.c_int)]


P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
PI3 = ctypes.POINTER(ctypes.c_int)
PD = ctypes.POINTER(ctypes.c_double)

# (name, restype, argtypes) for every symbol of include/rtkhip.h
SIGNATURES = [
    ("rtkhip_last_error", ctypes.c_char_p, []),
    ("rtkhip_version", ctypes.c_char_p, []),
    ("rtkhip_device_count", I, []),
    ("rtkhip_fdk_backproject", I, [PI3, PI3, PD, PI3, P, P, P, I, P]),
    ("rtkhip_voxel_backproject", I, [PI3, PI3, PD, PI3, P, P, P, I, P]),
    ("rtkhip_joseph_forward_project", I, [PI3, PI3, PD, PD, PD, D, D, P, P, P,
                                          ctypes.POINTER(ctypes.c_longlong), P]),
    ("rtkhip_joseph_backproject", I, [PI3, PI3, PD, PD, PD, D, D, P, P, P, I, P]),
    ("rtkhip_fdk_weight", I, [PI3, PD, PD, PD, P, P, P]),
    ("rtkhip_ramp_filter", I, [PI3, I, PD, I, PD, PD, PD, PD, P, P, I, P]),
    ("rtkhip_column_filter", I, [PI3, I, PD, P, P, I, P]),
    ("rtkhip_cg_step", I, [ctypes.c_size_t, P, P, P, P, P, P]),
    ("rtkhip_cg_step_noalloc", I, [ctypes.c_size_t, P, P, P, P, P, P]),
    ("rtkhip_cg_init", I, [ctypes.c_size_t, P, P, P, P, P, P]),
    ("rtkhip_dot", I, [ctypes.c_size_t, P, P, P, P]),
    ("rtkhip_multiply", I, [ctypes.c_size_t, P, P, P, P]),
    ("rtkhip_buffer_create", P, [ctypes.c_size_t]),
    ("rtkhip_buffer_destroy", None, [P]),
    ("rtkhip_buffer_set_host", I, [P, P]),
    ("rtkhip_buffer_device", I, [P, I, ctypes.POINTER(P), P]),
    ("rtkhip_buffer_host", I, [P, I, ctypes.POINTER(P), P]),
    ("rtkhip_buffer_transfers", I, [P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    ("rtkhip_geometry_create", P, []),
    ("rtkhip_geometry_destroy", None, [P]),
    ("rtkhip_geometry_add_projection", I, [P, D, D, D, D, D, D, D, D, D]),
    ("rtkhip_geometry_set_radius_cylindrical_detector", I, [P, D]),
    ("rtkhip_geometry_size", I, [P]),
    ("rtkhip_geometry_get", I, [P, I, PD, PD, PD]),
    ("rtkhip_context_create", P, []),
    ("rtkhip_context_destroy", None, [P]),
    ("rtkhip_fdk_reconstruct", I, [P, P, ctypes.POINTER(Image), P, P, ctypes.POINTER(Image), P,
                                   ctypes.POINTER(RampOptions), I, I, PD, P]),
    ("rtkhip_fft_ramp_filter", I, [P, ctypes.POINTER(Image), P, P, ctypes.POINTER(RampOptions), P]),
    ("rtkhip_forward_project", I, [P, P, ctypes.POINTER(Image), P, P, ctypes.POINTER(Image), P,
                                   ctypes.POINTER(ctypes.c_longlong), P]),
    ("rtkhip_back_project", I, [P, P, ctypes.POINTER(Image), P, P, ctypes.POINTER(Image), P, I, I, P]),
    ("rtkhip_joseph_back_project", I, [P, P, ctypes.POINTER(Image), P, P, ctypes.POINTER(Image), P, I, P]),
    ("rtkhip_cg_reconstruct", I, [P, P, ctypes.POINTER(Image), P, ctypes.POINTER(Image), P, P, I, D, I, I, PD,
                                  PI3, P]),
    ("rtkhip_parker_weight", I, [PI3, PD, D, D, D, D, P, P, P]),
    ("rtkhip_displaced_detector_weight", I, [PI3, I, I, PD, D, I, D, D, P, P, P]),
    ("rtkhip_parker_short_scan", I, [P, ctypes.POINTER(Image), P, P, D, PI3, PI3, P]),
    ("rtkhip_displaced_detector_region", I, [P, ctypes.POINTER(Image), I, I, I, D, D, ctypes.POINTER(Image), PD,
                                             PI3]),
    ("rtkhip_displaced_detector", I, [P, ctypes.POINTER(Image), P, ctypes.POINTER(Image), P, I, I, I, D, D, P]),
    ("rtkhip_draw_quadrics", I, [ctypes.POINTER(Image), P, P, I, PD, P]),
    ("rtkhip_project_quadrics", I, [P, P, ctypes.POINTER(Image), P, P, I, PD, P]),
    ("rtkhip_cnn_num_params", I, []),
    ("rtkhip_cnn_create", P, []),
    ("rtkhip_cnn_destroy", None, [P]),
    ("rtkhip_cnn_set_params", I, [P, P, P]),
    ("rtkhip_cnn_get_params", I, [P, P, P, P]),
    ("rtkhip_cnn_forward", I, [P, I, PI3, P, P, ctypes.c_float, P]),
    ("rtkhip_cnn_backward", I, [P, P, P, PD, P]),
    ("rtkhip_cnn_adam_step", I, [P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, I, P]),
]

_lib = None


def lib():
    """Loads librtkhip.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C rtk_amd` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != RTKHIP_OK:
        raise RtkHipError(rc, lib().rtkhip_last_error().decode())
    return rc


def int3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def dbl(v):
    import numpy as np
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float64).ravel())
    return a, a.ctypes.data_as(PD)
