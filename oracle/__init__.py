"""ORACLE — TEST INFRASTRUCTURE ONLY.

Python access to the C restatement of FFTLIB's natural-C kernels
(oracle/fftlib_cn.c, built into oracle/liboracle_cn.so by oracle/Makefile).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use it, as the
checker; the product library never loads it.

All functions take and return numpy arrays; inputs are copied before the call
because the natural-C kernels use their input buffer as scratch.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_cn.so")

FFTLIB_INT16 = 1
FFTLIB_UINT32 = 8
FFTLIB_FLOAT32 = 12


class _BP(ctypes.Structure):
    _fields_ = [("data_type", ctypes.c_uint32), ("dim_x", ctypes.c_uint32)]


_lib = None


def build():
    """Compile liboracle_cn.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle_cn.so"], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _bp(dtype, n):
    return ctypes.byref(_BP(dtype, n))


# --- table generators ---------------------------------------------------------
def tw_c2c(n):
    w = np.zeros(2 * n, np.float32)
    load().oracle_tw_gen_c2c(_ptr(w), ctypes.c_int32(n))
    return w


def tw_i16(n):
    w = np.zeros(2 * n, np.int16)
    load().oracle_tw_gen_i16sc(_ptr(w), ctypes.c_uint32(n))
    return w


def tw_r2c(n):
    w = np.zeros(max(n, 2), np.float32)
    load().oracle_tw_gen_r2c(_ptr(w), ctypes.c_int32(n))
    return w[:n]


def tw_c2r(n):
    w = np.zeros(max(n, 2), np.float32)
    load().oracle_ifft_c2r_tw_gen(_ptr(w), ctypes.c_int32(n))
    return w[:n]


def sf_r2c(n):
    s = np.zeros(2 * n, np.float32)
    load().oracle_split_factor_gen_r2c(_ptr(s), ctypes.c_int32(n))
    return s


def sf_c2r(n):
    s = np.zeros(2 * n, np.float32)
    load().oracle_ifft_c2r_split_factor_gen(_ptr(s), ctypes.c_int32(n))
    return s


# --- transforms ---------------------------------------------------------------
def c2c(x, n, b, inverse=False, single=False, w=None):
    """fp32 complex FFT of b channels of n points. x: complex64 (b, n) or the
    interleaved float32 buffer. single=True selects the single-channel entry
    points (b must be 1), whose inverse divides by N at the store."""
    lib = load()
    xin = np.ascontiguousarray(np.asarray(x, np.complex64)).view(np.float32).reshape(-1).copy()
    y = np.zeros_like(xin)
    w = tw_c2c(n) if w is None else np.ascontiguousarray(w, np.float32)
    if single:
        assert b == 1
        fn = lib.FFTLIB_ifft1d_i32fc_c32fc_o32fc_cn if inverse else lib.FFTLIB_fft1d_i32fc_c32fc_o32fc_cn
        bp = _bp(FFTLIB_FLOAT32, 2 * n)
        fn(_ptr(xin), bp, _ptr(w), bp, _ptr(y), bp)
    else:
        fn = lib.FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_cn if inverse else lib.FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_cn
        bx = _bp(FFTLIB_FLOAT32, 2 * n * b)
        fn(_ptr(xin), bx, _ptr(w), _bp(FFTLIB_FLOAT32, 2 * n), _ptr(y), bx, ctypes.c_uint32(n), ctypes.c_uint32(b))
    return y.view(np.complex64).reshape(b, n)


def i16(x, n, b, shifts, w=None, single=False):
    """int16 FFT. x: int16 array of shape (b, n, 2) (re, im pairs)."""
    lib = load()
    xin = np.ascontiguousarray(x, np.int16).reshape(-1).copy()
    y = np.zeros_like(xin)
    w = tw_i16(n) if w is None else np.ascontiguousarray(w, np.int16)
    sh = np.ascontiguousarray(shifts, np.uint32)
    if single:
        assert b == 1
        bp = _bp(FFTLIB_INT16, 2 * n)
        lib.FFTLIB_fft1d_i16sc_c16sc_o16sc_cn(_ptr(xin), bp, _ptr(w), bp, _ptr(y), bp, _ptr(sh),
                                              _bp(FFTLIB_UINT32, sh.size))
    else:
        bx = _bp(FFTLIB_INT16, 2 * n * b)
        lib.FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_cn(_ptr(xin), bx, _ptr(w), _bp(FFTLIB_INT16, 2 * n), _ptr(y), bx,
                                                     _ptr(sh), _bp(FFTLIB_UINT32, sh.size), ctypes.c_uint32(n),
                                                     ctypes.c_uint32(b))
    return y.reshape(b, n, 2)


def r2c(x, n, b, outmode=0, single=False):
    """fp32 real FFT. x: float32 (b, n). Returns complex64 (b, nOut)."""
    lib = load()
    xin = np.ascontiguousarray(x, np.float32).reshape(-1).copy()
    nout = n if (outmode or single) else n // 2 + 1
    w, sf = tw_r2c(n), sf_r2c(n)
    xf = np.zeros(n * b, np.float32)
    y = np.zeros(2 * nout * b, np.float32)
    bp = _bp(FFTLIB_FLOAT32, n * b)
    if single:
        lib.FFTLIB_fft1d_i32f_c32fc_o32fc_cn(_ptr(xin), bp, _ptr(w), bp, _ptr(xf), bp, _ptr(sf), bp, _ptr(y), bp)
    else:
        lib.FFTLIB_fft1dBatched_i32f_c32fc_o32fc_cn(_ptr(xin), bp, _ptr(w), bp, _ptr(xf), bp, _ptr(sf), bp, _ptr(y),
                                                    bp, ctypes.c_uint32(n), ctypes.c_uint32(b),
                                                    ctypes.c_uint8(outmode))
    return y.view(np.complex64).reshape(b, nout)


def c2r(x, n, b, inmode=0, single=False):
    """fp32 complex-to-real inverse. x: complex64 (b, n//2+1) for inMode 0 or
    (b, n) for inMode 1 / single. Returns float32 (b, n)."""
    lib = load()
    xin = np.ascontiguousarray(np.asarray(x, np.complex64)).view(np.float32).reshape(-1).copy()
    w, sf = tw_c2r(n), sf_c2r(n)
    xf = np.zeros(n * b, np.float32)
    y = np.zeros(n * b, np.float32)
    bp = _bp(FFTLIB_FLOAT32, n * b)
    if single:
        lib.FFTLIB_ifft1d_i32fc_c32fc_o32f_cn(_ptr(xin), bp, _ptr(w), bp, _ptr(xf), bp, _ptr(sf), bp, _ptr(y), bp)
    else:
        lib.FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_cn(_ptr(xin), bp, _ptr(w), bp, _ptr(xf), bp, _ptr(sf), bp,
                                                     _ptr(y), bp, ctypes.c_uint32(n), ctypes.c_uint32(b),
                                                     ctypes.c_uint8(inmode))
    return y.reshape(b, n)


def c2c_threads(x, n, b, inverse=False, threads=1):
    """All-cores baseline: channel split over POSIX threads (same arithmetic)."""
    lib = load()
    xin = np.ascontiguousarray(np.asarray(x, np.complex64)).view(np.float32).reshape(-1).copy()
    y = np.zeros_like(xin)
    w = tw_c2c(n)
    lib.oracle_fft1dBatched_c2c_threads(_ptr(xin), _ptr(w), _ptr(y), ctypes.c_uint32(n), ctypes.c_uint32(b),
                                        ctypes.c_int(int(inverse)), ctypes.c_int(threads))
    return y.view(np.complex64).reshape(b, n)


def raw_c2c_batched(xin, w, y, n, b, inverse=False):
    """In-place call on caller-owned float32 buffers (for timing loops)."""
    lib = load()
    fn = lib.FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_cn if inverse else lib.FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_cn
    bx = _bp(FFTLIB_FLOAT32, 2 * n * b)
    return fn(_ptr(xin), bx, _ptr(w), _bp(FFTLIB_FLOAT32, 2 * n), _ptr(y), bx, ctypes.c_uint32(n),
              ctypes.c_uint32(b))


def raw_threads_c2c(xin, w, y, n, b, inverse, threads):
    return load().oracle_fft1dBatched_c2c_threads(_ptr(xin), _ptr(w), _ptr(y), ctypes.c_uint32(n),
                                                  ctypes.c_uint32(b), ctypes.c_int(int(inverse)),
                                                  ctypes.c_int(threads))


# --- fft_c7xmma kernels (fixed-point DFT as matrix products) -------------------
MMA_SIZE = {1: 32, 2: 16}          # FFTLIB_INT16 -> 32 elements per MMA row, FFTLIB_INT32 -> 16
_NP = {1: np.int16, 2: np.int32}
_SFX = {1: "i16", 2: "i32"}


def mma_variant(fft, batch, interleave, dtype):
    """0 dftSmall, 1 dftLarge, 2 highRadixDecompositions (the coordinator's choice)."""
    return load().oracle_mma_variant(ctypes.c_uint32(fft), ctypes.c_uint32(batch), ctypes.c_uint32(interleave),
                                     ctypes.c_uint32(MMA_SIZE[dtype]))


def mma_tables(kind, fft, batch, interleave, dtype):
    """Caller table of kernel `kind` ('dftsmall', 'dftlarge', 'highradix'), and
    for highradix also the internal decomposition matrices (w, w1)."""
    lib, t, mma, sfx = load(), _NP[dtype], MMA_SIZE[dtype], _SFX[dtype]
    if kind == "dftsmall":
        w = np.zeros(mma * mma, t)
        getattr(lib, "oracle_mma_dftsmall_tw_" + sfx)(_ptr(w), ctypes.c_uint32(fft), ctypes.c_uint32(batch),
                                                     ctypes.c_uint32(interleave))
        return w, None
    if kind == "dftlarge":
        w = np.zeros(4 * fft * fft, t)
        getattr(lib, "oracle_mma_dftlarge_tw_" + sfx)(_ptr(w), ctypes.c_uint32(fft), ctypes.c_uint32(interleave))
        return w, None
    w = np.zeros(lib.oracle_mma_hr_wsize(ctypes.c_uint32(fft), ctypes.c_uint32(mma)), t)
    w1 = np.zeros(mma * mma, t)
    getattr(lib, "oracle_mma_hr_tw_" + sfx)(_ptr(w1), _ptr(w), ctypes.c_uint32(fft))
    return w, w1


def mma(kind, x, fft, batch, interleave, shifts, dtype):
    """Natural-C result of kernel `kind` ('fft' = the coordinator's choice) on
    x (2*fft*batch values of the data type); returns the output array."""
    lib, t, mma_n, sfx = load(), _NP[dtype], MMA_SIZE[dtype], _SFX[dtype]
    if kind == "fft":
        kind = ("dftsmall", "dftlarge", "highradix")[mma_variant(fft, batch, interleave, dtype)]
    xin = np.ascontiguousarray(x, t).reshape(-1).copy()
    y = np.zeros_like(xin)
    w, w1 = mma_tables(kind, fft, batch, interleave, dtype)
    sh = np.zeros(5, np.uint32)
    sh[:len(shifts)] = shifts
    if kind in ("dftsmall", "dftlarge"):
        ldw = mma_n if kind == "dftsmall" else 2 * fft
        getattr(lib, "oracle_mma_matrix_" + sfx)(_ptr(xin), _ptr(w), _ptr(y), ctypes.c_uint32(batch),
                                                 ctypes.c_uint32(2 * fft), ctypes.c_uint32(ldw),
                                                 ctypes.c_uint32(int(sh[0])))
    else:
        getattr(lib, "oracle_mma_hr_exec_" + sfx)(_ptr(xin), _ptr(w1), _ptr(w), _ptr(y), ctypes.c_uint32(fft),
                                                  ctypes.c_uint32(batch), _ptr(sh), ctypes.c_uint32(interleave))
    return y
