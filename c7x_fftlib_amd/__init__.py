"""Python binding of the MI355X (gfx950) FFTLIB C-ABI library.

The product is the C-ABI shared library ``libfftlib_amd.so`` built from
``csrc/`` by :mod:`c7x_fftlib_amd.build`; it exports the reference library's
``FFTLIB_<kernel>_init/_kernel/_checkParams`` entry points (declared in the
repository's ``include/`` headers) and runs them on the GPU. This module only
loads that library with ctypes, declares the prototypes, and offers thin
helpers that call the entry points the way the reference's test drivers do.

PyTorch, when present, is imported before the library is loaded so that the
library binds to the same HIP runtime as torch (device pointers and streams
are then shared).
"""
import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfftlib_amd.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

# --- ABI constants (include/FFTLIB_types.h, include/FFTLIB_bufParams.h) -------
FFTLIB_SUCCESS = 0
FFTLIB_ERR_FAILURE = 1
FFTLIB_ERR_INVALID_VALUE = 2
FFTLIB_ERR_INVALID_TYPE = 3
FFTLIB_ERR_INVALID_DIMENSION = 4
FFTLIB_ERR_NULL_POINTER = 5
FFTLIB_ERR_NOT_IMPLEMENTED = 6
FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES = 8
FFTLIB_ERR_INVALID_OUTMODE = 14

FFTLIB_INT16 = 1
FFTLIB_INT32 = 2
FFTLIB_UINT32 = 8
FFTLIB_FLOAT32 = 12
FFTLIB_ERR_INVALID_SHIFT = 12

FFTLIB_FUNCTION_NATC = 0
FFTLIB_FUNCTION_OPTIMIZED = 1
FFTLIB_FUNCTION_GPU = 2

FFTLIB_RAGGED_GRAPH = 0
FFTLIB_RAGGED_PERSISTENT = 1

SE_PARAM_SIZE = 64
PBLOCK_SIZE = {
    "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc": 31 * SE_PARAM_SIZE + 18 * 4,
    "FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc": 31 * SE_PARAM_SIZE + 18 * 4,
    "FFTLIB_fft1d_i32fc_c32fc_o32fc": 25 * SE_PARAM_SIZE + 16 * 4,
    "FFTLIB_ifft1d_i32fc_c32fc_o32fc": 27 * SE_PARAM_SIZE + 16 * 4,
    "FFTLIB_fft1dBatched_i16sc_c16sc_o16sc": 13 * SE_PARAM_SIZE,
    "FFTLIB_fft1d_i16sc_c16sc_o16sc": 9 * SE_PARAM_SIZE,
    "FFTLIB_fft1dBatched_i32f_c32fc_o32fc": 41 * SE_PARAM_SIZE + 18 * 4,
    "FFTLIB_fft1d_i32f_c32fc_o32fc": 35 * SE_PARAM_SIZE + 18 * 4,
    "FFTLIB_ifft1dBatched_i32fc_c32fc_o32f": 41 * SE_PARAM_SIZE + 18 * 4,
    "FFTLIB_ifft1d_i32fc_c32fc_o32f": 34 * SE_PARAM_SIZE + 16 * 4,
}


class BufParams1D(ctypes.Structure):
    """FFTLIB_bufParams1D_t: dim_x counts elements (2 per complex point)."""
    _fields_ = [("data_type", ctypes.c_uint32), ("dim_x", ctypes.c_uint32)]


class MmaInitArgs(ctypes.Structure):
    """InitArgs of FFTLIB_FFT_fft_ixX_cxX_oxX and FFTLIB_FFT_highRadixDecompositions_ixX_cxX_oxX
    (five shifts); the dftSmall / dftLarge InitArgs are the same with one shift."""
    _fields_ = [("funcStyle", ctypes.c_int8), ("fftSize", ctypes.c_uint32), ("batchSize", ctypes.c_uint32),
                ("interleave", ctypes.c_uint32), ("shiftVector", ctypes.c_uint32 * 5)]


class MmaInitArgs1(ctypes.Structure):
    _fields_ = [("funcStyle", ctypes.c_int8), ("fftSize", ctypes.c_uint32), ("batchSize", ctypes.c_uint32),
                ("interleave", ctypes.c_uint32), ("shiftVector", ctypes.c_uint32 * 1)]


MMA_KERNELS = {  # alias -> (C name prefix, InitArgs type)
    "fft": ("FFTLIB_FFT_fft_ixX_cxX_oxX", MmaInitArgs),
    "dftsmall": ("FFTLIB_FFT_dftSmall_ixX_cxX_oxX", MmaInitArgs1),
    "dftlarge": ("FFTLIB_FFT_dftLarge_ixX_cxX_oxX", MmaInitArgs1),
    "highradix": ("FFTLIB_FFT_highRadixDecompositions_ixX_cxX_oxX", MmaInitArgs),
}

_P = ctypes.c_void_p
_BP = ctypes.POINTER(BufParams1D)
_U32 = ctypes.c_uint32
_U8 = ctypes.c_uint8
_S = ctypes.c_int

# name suffix -> argtypes, per kernel signature family
_SIG_C2C_B = [_P, _BP, _P, _BP, _P, _BP, _U32, _U32, _P]
_SIG_C2C_1 = [_P, _BP, _P, _BP, _P, _BP, _P]
_SIG_I16_B = [_P, _BP, _P, _BP, _P, _BP, _P, _BP, _U32, _U32, _P]
_SIG_I16_1 = [_P, _BP, _P, _BP, _P, _BP, _P, _BP, _P]
_SIG_REAL_B = [_P, _BP, _P, _BP, _P, _BP, _P, _BP, _P, _BP, _U32, _U32, _U8, _P]
_SIG_REAL_1 = [_P, _BP, _P, _BP, _P, _BP, _P, _BP, _P, _BP, _P]
_SIG_REAL_B_CHECK = [_P, _BP, _P, _BP, _P, _BP, _U32, _U32, _U8, _P]
_SIG_REAL_1_CHECK = [_P, _BP, _P, _BP, _P, _BP, _P]

KERNELS = {
    "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc": (_SIG_C2C_B, _SIG_C2C_B),
    "FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc": (_SIG_C2C_B, _SIG_C2C_B),
    "FFTLIB_fft1d_i32fc_c32fc_o32fc": (_SIG_C2C_1, _SIG_C2C_1),
    "FFTLIB_ifft1d_i32fc_c32fc_o32fc": (_SIG_C2C_1, _SIG_C2C_1),
    "FFTLIB_fft1dBatched_i16sc_c16sc_o16sc": (_SIG_I16_B, _SIG_I16_B),
    "FFTLIB_fft1d_i16sc_c16sc_o16sc": (_SIG_I16_1, _SIG_I16_1),
    "FFTLIB_fft1dBatched_i32f_c32fc_o32fc": (_SIG_REAL_B, _SIG_REAL_B_CHECK),
    "FFTLIB_fft1d_i32f_c32fc_o32fc": (_SIG_REAL_1, _SIG_REAL_1_CHECK),
    "FFTLIB_ifft1dBatched_i32fc_c32fc_o32f": (_SIG_REAL_B, _SIG_REAL_B_CHECK),
    "FFTLIB_ifft1d_i32fc_c32fc_o32f": (_SIG_REAL_1, _SIG_REAL_1_CHECK),
}

_EXTRA = {
    "FFTLIB_setFunctionStyle": (_S, [_S]),
    "FFTLIB_getFunctionStyle": (_S, []),
    "FFTLIB_GPU_setStream": (_S, [_P]),
    "FFTLIB_GPU_getStream": (_P, []),
    "FFTLIB_GPU_setSynchronous": (_S, [ctypes.c_int32]),
    "FFTLIB_GPU_setSpecializedKernels": (_S, [ctypes.c_int32]),
    "FFTLIB_GPU_setC2CVariant": (_S, [ctypes.c_int32]),
    "FFTLIB_GPU_getC2CVariant": (ctypes.c_int32, []),
    "FFTLIB_GPU_raggedCreate": (_S, [_P, _P, ctypes.c_uint32]),
    "FFTLIB_GPU_raggedExec": (_S, [_P, _P, _P, _P, ctypes.c_int32]),
    "FFTLIB_GPU_raggedTune": (_S, [_P, _P, _P]),
    "FFTLIB_GPU_raggedDestroy": (_S, [_P]),
    "FFTLIB_GPU_raggedPoints": (ctypes.c_uint64, [_P]),
    "FFTLIB_GPU_raggedSetForm": (_S, [_P, ctypes.c_int32]),
    "FFTLIB_GPU_raggedGetForm": (ctypes.c_int32, [_P]),
    "FFTLIB_GPU_tuneSave": (_S, [ctypes.c_char_p]),
    "FFTLIB_GPU_tuneLoad": (_S, [ctypes.c_char_p]),
    "FFTLIB_GPU_synchronize": (_S, []),
    "FFTLIB_GPU_getLastError": (ctypes.c_char_p, []),
    "FFTLIB_GPU_version": (ctypes.c_char_p, []),
    "FFTLIB_tw_gen_c2c": (None, [_P, ctypes.c_int32]),
    "FFTLIB_tw_gen_i16sc": (None, [_P, _U32]),
    "tw_gen_r2c": (None, [_P, ctypes.c_int32]),
    "split_factor_gen_r2c": (None, [_P, ctypes.c_int32]),
    "ifft_c2r_tw_gen": (None, [_P, ctypes.c_int32]),
    "ifft_c2r_split_factor_gen": (None, [_P, ctypes.c_int32]),
}

_lib = None


def declared_symbols():
    """Names of every function declared in the repository's include/*.h."""
    pat = re.compile(r"^\s*(?:FFTLIB_STATUS|void|const char \*|FFTLIB_FUNCTION_STYLE|int32_t|uint32_t|uint64_t)\s*\**\s*"
                     r"(\w+)\s*\(", re.M)
    names = set()
    for fname in sorted(os.listdir(INCLUDE_DIR)):
        if fname.endswith(".h"):
            with open(os.path.join(INCLUDE_DIR, fname)) as fh:
                text = re.sub(r"/\*.*?\*/", "", fh.read(), flags=re.S)
            names.update(m.group(1) for m in pat.finditer(text))
    names.discard("FFTLIB_sizeof")  # static inline in FFTLIB_bufParams.h
    return sorted(names)


def _declare(lib):
    for name, (init_sig, check_sig) in KERNELS.items():
        for suffix, sig in (("_init", init_sig), ("_kernel", init_sig), ("_checkParams", check_sig)):
            fn = getattr(lib, name + suffix)
            fn.restype = _S
            fn.argtypes = sig
    for prefix, _args in MMA_KERNELS.values():
        for suffix, res, argt in (("_getHandleSize", ctypes.c_int32, [_P]), ("_getSizes", _S, [_P, _BP, _BP, _BP]),
                                  ("_twGen", _S, [_P, _P, _BP]), ("_init", _S, [_P, _BP, _BP, _BP, _P]),
                                  ("_init_checkParams", _S, [_P, _BP, _BP, _BP, _P]),
                                  ("_exec", _S, [_P, _P, _P, _P]), ("_exec_checkParams", _S, [_P, _P, _P, _P])):
            fn = getattr(lib, prefix + suffix)
            fn.restype = res
            fn.argtypes = argt
    for name, (res, args) in _EXTRA.items():
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.restype = res
            fn.argtypes = args
    from . import _extra_decls  # noqa: F401  (declares later-added entry points)
    _extra_decls.declare(lib)


def load(path=None):
    """Load (once) and return the C-ABI library. Raises if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    try:
        import torch  # noqa: F401  -- share torch's HIP runtime
    except ImportError:
        pass
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is not built; run `python -m c7x_fftlib_amd.build`")
    lib = ctypes.CDLL(p)
    _declare(lib)
    if path is None:
        _lib = lib
    return lib


def bufparams(data_type, dim_x):
    return BufParams1D(data_type, dim_x)


def last_error():
    return load().FFTLIB_GPU_getLastError().decode()
