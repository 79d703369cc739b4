"""ctypes binding of ``libcucim_amd.so`` (the C ABI in ``include/cucim_amd.h``).

The library is loaded on first use.  If it is missing, or no GPU is present,
every compute entry point raises: there is no CPU fallback.
"""

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcucim_amd.so")

MAX_NDIM = 8

# hc_dtype
(BOOL, INT8, UINT8, INT16, UINT16, INT32, UINT32, INT64, UINT64, FLOAT16, FLOAT32, FLOAT64,
 BFLOAT16) = range(13)

# hc_mode
MODE_REFLECT, MODE_MIRROR, MODE_NEAREST, MODE_WRAP, MODE_CONSTANT = range(5)

# hc_algorithm
ALGO_AUTO, ALGO_TILED, ALGO_GATHER = range(3)

# hc_tune_knob
TUNE_CONTIG_SEG, TUNE_SEP2D_TY, TUNE_HESSIAN_VALU, TUNE_EDGE_ROWS, TUNE_MFMA_WAVES = range(5)
BIN_SUB, BIN_MUL, BIN_ADD = 0, 1, 2

# hc_status
OK, ERR_VALUE, ERR_RESOURCE, ERR_HIP, ERR_UNSUPPORTED = range(5)

ENTRY_POINTS = (
    "hc_last_error",
    "hc_version",
    "hc_max_inline_taps",
    "hc_max_resident_line",
    "hc_timing_enable",
    "hc_timing_collect",
    "hc_tune_set",
    "hc_tune_get",
    "hc_correlate1d",
    "hc_correlate_nd",
    "hc_separable2d",
    "hc_hessian_gaussian",
    "hc_sym_eigvals",
    "hc_frangi",
    "hc_frangi_from_eigvals",
    "hc_rgb2gray",
    "hc_to_float",
    "hc_gradient",
    "hc_edge_magnitude",
    "hc_mask_eroded",
    "hc_maximum_inplace",
    "hc_square_acc",
    "hc_binary",
    "hc_div_scalar",
    "hc_correlate_sparse",
    "hc_gaussian_kernel1d",
    "hc_gaussian_nd",
    "hc_hessian_matrix",
    "hc_frangi_sigmas",
    "hc_minmax",
    "hc_minmax_noalloc",
    "hc_histogram",
    "hc_li_means",
    "hc_label_work_elems",
    "hc_label",
    "hc_wsi_open",
    "hc_wsi_close",
    "hc_wsi_level_count",
    "hc_wsi_level_info",
    "hc_wsi_properties",
    "hc_wsi_read_region",
    "hc_wsi_read_region_device",
    "hc_wsi_associated_count",
    "hc_wsi_associated_info",
    "hc_wsi_associated_read",
    "hc_sqrt_inplace",
    "hc_copy_cast",
    "hc_batch2d_plan",
    "hc_batch2d_gray_gauss_sobel",
    "hc_sato_fold",
    "hc_meijering_fold",
    "hc_fill_le",
    "hc_cnn3d_packed_size",
    "hc_cnn3d_pack",
    "hc_cnn3d_conv",
    "hc_cnn3d_pad",
    "hc_cnn3d_gelu_bwd",
    "hc_cnn3d_dgrad",
    "hc_cnn3d_wgrad",
    "hc_cnn3d_mse_grad",
    "hc_cnn3d_adam",
    "hc_frangi_check",
)


class ResourceLimitError(RuntimeError):
    """The LDS-tiled kernels cannot take this kernel length / line length
    (the reference's ``filters/_separable_filtering.py:126-153``)."""


class HCArray(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("ndim", ctypes.c_int32),
        ("shape", ctypes.c_int64 * MAX_NDIM),
        ("strides", ctypes.c_int64 * MAX_NDIM),
    ]


class HCTaps(ctypes.Structure):
    _fields_ = [
        ("w", ctypes.POINTER(ctypes.c_double)),
        ("K", ctypes.c_int64),
        ("anchor", ctypes.c_int32),
        ("w_f64", ctypes.c_int32),
        ("w_dev", ctypes.c_void_p),
    ]


_P = ctypes.POINTER
_A = _P(HCArray)
_T = _P(HCTaps)
_v = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

_SIGNATURES = {
    "hc_last_error": ([], ctypes.c_char_p),
    "hc_version": ([], _i),
    "hc_max_inline_taps": ([_i], _i),
    "hc_max_resident_line": ([_i], _i),
    "hc_timing_enable": ([_i], _i),
    "hc_timing_collect": ([ctypes.c_char_p, _i, _P(_d), _i], _i),
    "hc_tune_set": ([_i, _i], _i),
    "hc_tune_get": ([_i], _i),
    "hc_correlate1d": ([_A, _A, _T, _i, _i, _d, _i, _v], _i),
    "hc_correlate_nd": ([_A, _A, _v, _P(ctypes.c_int64), _P(ctypes.c_int32),
                         _P(ctypes.c_int32), _i, _d, _v], _i),
    "hc_separable2d": ([_A, _A, _T, _T, _i, _i, _d, _v], _i),
    "hc_hessian_gaussian": ([_A, _A, _i, _T, _T, _i, _d, _i, _v, _v], _i),
    "hc_sym_eigvals": ([_A, _i, _A, _i, _i, _v], _i),
    "hc_frangi": ([_A, _A, _T, _T, _i, _d, _d, _d, _i, _i, _i, _d, _A, _i, _v, _v, _v], _i),
    "hc_frangi_from_eigvals": ([_A, _i, _A, _d, _d, _d, _i, _i, _v, _v], _i),
    "hc_rgb2gray": ([_A, _A, _v], _i),
    "hc_to_float": ([_A, _A, _v], _i),
    "hc_gradient": ([_A, _A, _i, _v], _i),
    "hc_edge_magnitude": ([_A, _A, _P(_d), _P(_d), _i, _i, _A, _v], _i),
    "hc_mask_eroded": ([_A, _A, _v], _i),
    "hc_maximum_inplace": ([_A, _A, _v], _i),
    "hc_square_acc": ([_A, _A, _i, _v], _i),
    "hc_binary": ([_A, _A, _A, _i, _v], _i),
    "hc_div_scalar": ([_A, _d, _v], _i),
    "hc_correlate_sparse": ([_A, _A, _v, _v, _v, _i, _i, _i, _v], _i),
    "hc_gaussian_kernel1d": ([_d, _i, ctypes.c_int64, _v], _i),
    "hc_gaussian_nd": ([_A, _A, _v, _v, _v, _v, _d, _d, _A, _v], _i),
    "hc_hessian_matrix": ([_A, _A, _i, _v, _i, _d, _i, _v, _v], _i),
    "hc_frangi_sigmas": ([_A, _A, _v, _i, _d, _d, _d, _i, _i, _i, _d, _A, _i, _v, _v], _i),
    "hc_minmax": ([_A, _v, _v, ctypes.c_int64, _v], _i),
    "hc_minmax_noalloc": ([_A, _v, _v], _i),
    "hc_histogram": ([_A, ctypes.c_int64, _v, ctypes.c_int64, _v, _v], _i),
    "hc_li_means": ([_A, _d, _d, _v, _v, ctypes.c_int64, _v], _i),
    "hc_label_work_elems": ([ctypes.c_int64], ctypes.c_int64),
    "hc_label": ([_A, ctypes.c_int64, _i, _A, _v, ctypes.c_int64, _v, _v], _i),
    "hc_wsi_open": ([ctypes.c_char_p, _P(_v)], _i),
    "hc_wsi_close": ([_v], None),
    "hc_wsi_level_count": ([_v], _i),
    "hc_wsi_level_info": ([_v, _i, _P(ctypes.c_int64)], _i),
    "hc_wsi_properties": ([_v, _P(_d), _P(_i), ctypes.c_char_p, _i, ctypes.c_char_p, _i], _i),
    "hc_wsi_read_region": ([_v, _i, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                            ctypes.c_int64, _v, _i], _i),
    "hc_wsi_read_region_device": ([_v, _i, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_int64, _v, _i, _v], _i),
    "hc_wsi_associated_count": ([_v], _i),
    "hc_wsi_associated_info": ([_v, _i, ctypes.c_char_p, _i, _P(ctypes.c_int64)], _i),
    "hc_wsi_associated_read": ([_v, _i, _v, _i], _i),
    "hc_sqrt_inplace": ([_A, _v], _i),
    "hc_copy_cast": ([_A, _A, _v], _i),
    "hc_batch2d_plan": ([_v, _i, _i, _i, _v, ctypes.c_int64], ctypes.c_int64),
    "hc_batch2d_gray_gauss_sobel": ([_v, _i, _v, ctypes.c_int64, _T, _i, _d, _P(_d), _P(_d),
                                     _i, _i, _v], _i),
    "hc_sato_fold": ([_A, _i, _A, _d, _i, _v], _i),
    "hc_meijering_fold": ([_A, _i, _A, _A, _d, _i, _v, _v], _i),
    "hc_fill_le": ([_A, _d, _d, _v], _i),
    "hc_cnn3d_packed_size": ([_i, _i], _i),
    "hc_cnn3d_pack": ([_v, _v, _i, _i, _v], _i),
    "hc_cnn3d_conv": ([_A, _A, _v, _v, _i, _v, _v], _i),
    "hc_cnn3d_pad": ([_A, _A, _i, _i, _d, _v], _i),
    "hc_cnn3d_gelu_bwd": ([_v, _v, ctypes.c_int64, _v], _i),
    "hc_cnn3d_dgrad": ([_v, _v, _v, _i, _i, _i, _i, _i, _v], _i),
    "hc_cnn3d_wgrad": ([_v, _v, _v, _v, _i, _i, _i, _i, _i, _v], _i),
    "hc_cnn3d_mse_grad": ([_v, _v, _v, ctypes.c_int64, _v, _v], _i),
    "hc_cnn3d_adam": ([_v, _v, _v, _v, ctypes.c_int64, _d, _d, _d, _d, _i, _v], _i),
    "hc_frangi_check": ([_A, _A, _v, _i, _v, _v, _P(ctypes.c_int32), _i, _i, _d, _d, _d, _v,
                         _v, _v], _i),
}

_lock = threading.Lock()
_lib = None


def load_library(path=LIB_PATH):
    """Load the shared library and declare every entry point (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f"cucim_amd HIP library not found at {path}; build it with "
                "`python -m cucim_amd._build`")
        lib = ctypes.CDLL(path)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
        return lib


def lib():
    """The loaded library, for compute calls: requires a visible GPU."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("cucim_amd needs a ROCm GPU; none is visible")
    return load_library()


def check(rc, what=""):
    if rc == OK:
        return
    msg = load_library().hc_last_error().decode(errors="replace")
    if rc == ERR_RESOURCE:
        raise ResourceLimitError(msg)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == ERR_VALUE:
        raise ValueError(f"{what}: {msg}" if what else msg)
    raise RuntimeError(f"{what}: {msg}" if what else msg)


def make_taps(w, anchor, w_f64, w_dev=None):
    """Build an ``hc_taps`` for host weights ``w`` (already in the weights
    dtype's precision).  Returns (struct, keepalive)."""
    wd = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
    t = HCTaps()
    t.w = wd.ctypes.data_as(_P(ctypes.c_double))
    t.K = wd.size
    t.anchor = int(anchor)
    t.w_f64 = 1 if w_f64 else 0
    t.w_dev = None if w_dev is None else ctypes.c_void_p(int(w_dev))
    return t, wd


class KernelTimer:
    """Context manager recording per-launch kernel times through the
    library's hipEvent tracing; ``records`` is a list of (name, ms)."""

    def __enter__(self):
        load_library().hc_timing_enable(1)
        self.records = []
        return self

    def collect(self):
        lib = load_library()
        cap = 1 << 16
        names = ctypes.create_string_buffer(cap * 24)
        ms = (ctypes.c_double * cap)()
        n = lib.hc_timing_collect(names, len(names), ms, cap)
        tags = names.value.decode().split("\n")
        self.records += [(tags[i], ms[i]) for i in range(n)]
        return self.records

    def __exit__(self, *exc):
        self.collect()
        load_library().hc_timing_enable(0)
        return False
