"""ctypes prototypes of the entry points beyond the ten fft_c7x kernels."""
import ctypes

_P = ctypes.c_void_p
_S = ctypes.c_int
_U32 = ctypes.c_uint32

# name -> (restype, argtypes); entries are declared only when the library
# exports them.
DECLS = {}


def declare(lib):
    for name, (res, args) in DECLS.items():
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.restype = res
            fn.argtypes = args
