"""ctypes binding of liblpa_ici.so (the C ABI in include/lpa_ici.h).

The library is loaded from the package directory (built in-tree by
``make -C lpa-ici-cuda_amd``).  There is no fallback: if the library is
missing, every call raises.  torch is imported before the library when it is
available so both share one HIP runtime in the process.
"""
import ctypes
import os

try:  # share torch's HIP runtime (same soname) when torch is present
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is part of this image
    torch = None

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "liblpa_ici.so")
KERNEL_DIR = os.path.join(PKG_DIR, "kernels")
DEFAULT_KERNELS = os.path.join(KERNEL_DIR, "lpa_kernels_m_1_0_d_16_h12_1_2_4_8_16_24_32_symmetric.txt.gz")
GAUSSIAN_M20_KERNELS = os.path.join(KERNEL_DIR, "lpa_kernels_m_2_0_d_16_h12_1_2_4_8_16_24_32_gaussian.txt.gz")

LPA_OK, LPA_EINVAL, LPA_ETOOBIG, LPA_ENOMEM, LPA_EHIP, LPA_EIO = range(6)
STATUS_NAMES = {0: "LPA_OK", 1: "LPA_EINVAL", 2: "LPA_ETOOBIG", 3: "LPA_ENOMEM", 4: "LPA_EHIP", 5: "LPA_EIO"}
FWD_PLAIN, FWD_EXACT, FWD_GRAPH, FWD_WEIGHTED, FWD_RING_FLOAT = 1, 2, 4, 8, 16


class LpaError(RuntimeError):
    """A non-LPA_OK status from the C ABI; .status holds the code."""

    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class LpaImage(ctypes.Structure):
    _fields_ = [
        ("x", ctypes.c_void_p),
        ("y", ctypes.c_void_p),
        ("scale_map", ctypes.c_void_p),
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("pitch", ctypes.c_int),
        ("sigma", ctypes.c_float),
        ("gamma", ctypes.c_float),
        ("variances", ctypes.c_void_p),
        ("estimates", ctypes.c_void_p),
    ]


class KsetInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("num_dirs", "num_scales", "num_runs", "num_weights", "num_nonzero", "max_radius", "ring_form",
                 "num_ring_taps", "ring_frac_bits")] + [("ring_range", ctypes.c_float)]


class TileConfig(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("tile_w", "tile_h", "px_per_lane", "body", "halo", "blocks_per_cu")]


_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_LL = ctypes.c_longlong

# name -> (restype, argtypes); the full list of entry points the header declares.
SIGNATURES = {
    "lpa_last_error": (ctypes.c_char_p, []),
    "lpa_kset_create": (_I, [_I, _I, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "lpa_kset_destroy": (None, [_P]),
    "lpa_kset_get_info": (_I, [_P, ctypes.POINTER(KsetInfo)]),
    "lpa_kset_get_runs": (_I, [_P] + [_P] * 9),
    "lpa_ici_forward": (_I, [_P, ctypes.POINTER(LpaImage), _I, ctypes.c_uint, _P]),
    "lpa_ici_forward_noalloc": (_I, [_P, ctypes.POINTER(LpaImage), _I, ctypes.c_uint, _P]),
    "lpa_plan_create": (_I, [_P, ctypes.POINTER(LpaImage), _I, ctypes.c_uint, ctypes.POINTER(_P)]),
    "lpa_plan_run": (_I, [_P, _P]),
    "lpa_plan_destroy": (None, [_P]),
    "lpa_num_tile_configs": (_I, []),
    "lpa_tile_config_get": (_I, [_I, ctypes.POINTER(TileConfig)]),
    "lpa_kset_set_tile_config": (_I, [_P, _I]),
    "lpa_kset_get_tile_config": (_I, [_P]),
    "lpa_ici_estimates": (_I, [_P, ctypes.POINTER(LpaImage), _P, _I, ctypes.c_uint, _P]),
    "lpa_ici_backward": (_I, [_P, ctypes.POINTER(LpaImage), _P, _P, _P, _I, _P]),
    "lpa_workspace_peak": (_I, [ctypes.POINTER(ctypes.c_ulonglong), _I]),
    "lpa_load_kernels": (_I, [ctypes.c_char_p, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I),
                              _P, _P, _P, _P]),
    "lpa_kernel_variances": (None, [_I, _P, _P, _P, _P, _F, _P]),
    "lpa_wedge_kernels": (_I, [_I, _I, _P, ctypes.POINTER(_I), _P, _P, _P, _P]),
    "lpa_make_kernels": (_I, [_I, _I, _P, _I, _I, ctypes.POINTER(_I), _P, _P, _P, _P]),
    "lpa_write_kernels": (_I, [ctypes.c_char_p, _I, _I, _P, _P, _P, _P]),
    "lpa_make_noisy": (None, [_P, _P, _LL, _F, ctypes.c_uint]),
    "lpa_psnr": (ctypes.c_double, [_P, _P, _LL]),
    "lpa_synth_image": (None, [_P, _I, _I, ctypes.c_uint]),
    "lpa_load_image": (_I, [ctypes.c_char_p, ctypes.POINTER(_I), ctypes.POINTER(_I), _P, _LL]),
    "lpa_write_hdr": (_I, [ctypes.c_char_p, _I, _I, _P]),
    "lpa_load_image_channels": (_I, [ctypes.c_char_p, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), _P, _LL]),
    "lpa_write_hdr_channels": (_I, [ctypes.c_char_p, _I, _I, _I, _P]),
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C {PKG_DIR}`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status):
    if status != LPA_OK:
        raise LpaError(status, lib().lpa_last_error().decode(errors="replace"))
