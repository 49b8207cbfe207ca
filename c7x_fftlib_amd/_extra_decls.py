"""ctypes prototypes of the entry points beyond the ten fft_c7x kernels."""
import ctypes

_P = ctypes.c_void_p
_S = ctypes.c_int
_U32 = ctypes.c_uint32

# name -> (restype, argtypes); entries are declared only when the library
# exports them.
DECLS = {
    "FFTLIB_GPU_kernelAsync": (_S, [_P, _P, _P, _P, _P, _P]),
    "FFTLIB_GPU_kernelAsyncNoAlloc": (_S, [_P, _P, _P, _P, _P, _P]),
    "FFTLIB_GPU_pipelineCreate": (_S, [_P, _U32]),
    "FFTLIB_GPU_pipelineSubmit": (_S, [_P, _P, _P, _P, _P, _P]),
    "FFTLIB_GPU_pipelineWait": (_S, [_P]),
    "FFTLIB_GPU_pipelineDestroy": (_S, [_P]),
}


def declare(lib):
    for name, (res, args) in DECLS.items():
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.restype = res
            fn.argtypes = args
