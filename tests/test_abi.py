"""Host-side checks of the C-ABI library (no GPU needed).

  * the library loads and exports every function the include/*.h headers
    declare;
  * _checkParams reproduce the reference's optimized-kernel checks, case by
    case as the reference drivers' coverage tests (IDs 1000+) exercise them,
    and the oracle's _checkParams_cn reproduce the natural-C checks;
  * _init writes a plan (or rejects unsupported sizes) without touching the
    GPU; _kernel refuses an uninitialised pBlock, a size mismatch, NULL buffers
    and the natural-C style before any GPU work;
  * the library's table generators are bit-identical to the oracle's.
"""
import ctypes

import numpy as np
import pytest

import c7x_fftlib_amd as F

OK, NULLP, DIM, TYPE, ALIGN, OUTMODE, NOTIMPL, INVAL = (F.FFTLIB_SUCCESS, F.FFTLIB_ERR_NULL_POINTER,
                                                        F.FFTLIB_ERR_INVALID_DIMENSION, F.FFTLIB_ERR_INVALID_TYPE,
                                                        F.FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES,
                                                        F.FFTLIB_ERR_INVALID_OUTMODE, F.FFTLIB_ERR_NOT_IMPLEMENTED,
                                                        F.FFTLIB_ERR_INVALID_VALUE)


def bp(dt, n):
    return ctypes.byref(F.BufParams1D(dt, n))


class Buffers:
    """16-byte aligned host buffers (contents unused by the checks)."""

    def __init__(self):
        self.raw = np.zeros(1 << 16, np.float32)
        base = self.raw.ctypes.data
        self.base = base + (-base % 64)

    def p(self, i):
        return self.base + i * 4096


@pytest.fixture(scope="module")
def lib():
    return F.load()


@pytest.fixture(scope="module")
def mem():
    return Buffers()


def test_exports_every_declared_function(lib):
    names = F.declared_symbols()
    assert len(names) >= 44
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_style_selector(lib):
    assert b"gfx950" in lib.FFTLIB_GPU_version()
    assert lib.FFTLIB_setFunctionStyle(99) == INVAL
    for st in (F.FFTLIB_FUNCTION_NATC, F.FFTLIB_FUNCTION_GPU, F.FFTLIB_FUNCTION_OPTIMIZED):
        assert lib.FFTLIB_setFunctionStyle(st) == OK
        assert lib.FFTLIB_getFunctionStyle() == st


# ---- coverage cases of the reference drivers -------------------------------
# test/fft_c7x/FFTLIB_fft1d_i32fc_c32fc_o32fc/..._d.c:282-410 (N = 64)
def test_coverage_c2c_single(lib, oracle_lib, mem):
    n, dx = 64, 128
    X, W, Y, PB = mem.p(0), mem.p(1), mem.p(2), mem.p(3)
    f32 = F.FFTLIB_FLOAT32

    def both(px, bx, pw, bw, py, by):
        o = lib.FFTLIB_fft1d_i32fc_c32fc_o32fc_checkParams(px, bx, pw, bw, py, by, PB)
        c = oracle_lib.FFTLIB_fft1d_i32fc_c32fc_o32fc_checkParams_cn(ctypes.c_void_p(px), bx, ctypes.c_void_p(pw), bw,
                                                                     ctypes.c_void_p(py), by)
        return o, c

    assert both(None, bp(f32, dx), W, bp(f32, dx), Y, bp(f32, dx)) == (NULLP, NULLP)
    assert both(X, bp(f32, dx), None, bp(f32, dx), Y, bp(f32, dx)) == (NULLP, NULLP)
    assert both(X, bp(f32, dx), W, bp(f32, dx), None, bp(f32, dx)) == (NULLP, NULLP)
    assert both(X, bp(f32, dx), W, bp(f32, dx // 2), Y, bp(f32, dx)) == (DIM, DIM)
    assert both(X, bp(f32, dx), W, bp(f32, dx), Y, bp(f32, dx // 2)) == (DIM, DIM)
    assert both(X, bp(F.FFTLIB_UINT32, dx), W, bp(f32, dx), Y, bp(f32, dx)) == (TYPE, TYPE)
    assert both(X, bp(f32, dx), W, bp(F.FFTLIB_UINT32, dx), Y, bp(f32, dx)) == (TYPE, TYPE)
    assert both(X, bp(f32, dx), W, bp(f32, dx), Y, bp(F.FFTLIB_UINT32, dx)) == (TYPE, TYPE)
    assert both(X + 4, bp(f32, dx), W, bp(f32, dx), Y, bp(f32, dx)) == (ALIGN, OK)
    assert both(X, bp(f32, 14), W, bp(f32, 14), Y, bp(f32, 14)) == (DIM, DIM)   # N = 7
    assert both(X, bp(f32, 60), W, bp(f32, 60), Y, bp(f32, 60)) == (OK, OK)     # N = 30


# test/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/..._d.c:403-517 (and the
# inverse kernel, whose _cn reports a non-smooth N as NOT_IMPLEMENTED)
@pytest.mark.parametrize("inverse", [False, True])
def test_coverage_c2c_batched(lib, oracle_lib, mem, inverse):
    pre = "FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc" if inverse else "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc"
    opt = getattr(lib, pre + "_checkParams")
    nat = getattr(oracle_lib, pre + "_checkParams_cn")
    n, b = 64, 2
    X, W, Y, PB = mem.p(0), mem.p(1), mem.p(2), mem.p(3)
    f32 = F.FFTLIB_FLOAT32

    def both(px, bx, pw, bw, py, by, nn=n, pb=PB):
        o = opt(px, bx, pw, bw, py, by, nn, b, pb)
        c = nat(ctypes.c_void_p(px), bx, ctypes.c_void_p(pw), bw, ctypes.c_void_p(py), by, nn, b)
        return o, c

    d = 2 * n * b
    assert both(None, bp(f32, d), W, bp(f32, 2 * n), Y, bp(f32, d)) == (NULLP, NULLP)
    assert both(X, bp(f32, d), None, bp(f32, 2 * n), Y, bp(f32, d)) == (NULLP, NULLP)
    assert both(X, bp(f32, d), W, bp(f32, 2 * n), None, bp(f32, d)) == (NULLP, NULLP)
    assert both(X, bp(f32, d), W, bp(f32, 2 * n), Y, bp(f32, d), pb=None) == (NULLP, OK)
    assert both(X, bp(f32, d), W, bp(f32, 2 * n), Y, bp(f32, d // 2)) == (DIM, DIM)
    assert both(X, bp(f32, d), W, bp(f32, n), Y, bp(f32, d)) == (DIM, DIM)
    assert both(X, bp(F.FFTLIB_UINT32, d), W, bp(f32, 2 * n), Y, bp(f32, d)) == (TYPE, TYPE)
    assert both(X, bp(f32, d), W, bp(F.FFTLIB_UINT32, 2 * n), Y, bp(f32, d)) == (TYPE, TYPE)
    assert both(X, bp(f32, d), W, bp(f32, 2 * n), Y, bp(F.FFTLIB_UINT32, d)) == (TYPE, TYPE)
    assert both(X + 4, bp(f32, d), W, bp(f32, 2 * n), Y, bp(f32, d)) == (ALIGN, OK)
    assert both(X, bp(f32, d), W, bp(f32, 14), Y, bp(f32, d), nn=7) == (DIM, NOTIMPL if inverse else DIM)
    assert both(X, bp(f32, d), W, bp(f32, 60), Y, bp(f32, d), nn=30) == (OK, OK)


# test/fft_c7x/FFTLIB_fft1dBatched_i32f_c32fc_o32fc/..._d.c:300-455 (N = 60)
def test_coverage_r2c_batched(lib, oracle_lib, mem):
    n = 60
    X, W, Y, PB = mem.p(0), mem.p(1), mem.p(2), mem.p(3)
    f32 = F.FFTLIB_FLOAT32

    def both(px, bx, pw, bw, py, by, nn=n, om=0):
        o = lib.FFTLIB_fft1dBatched_i32f_c32fc_o32fc_checkParams(px, bx, pw, bw, py, by, nn, 1, om, PB)
        c = oracle_lib.FFTLIB_fft1dBatched_i32f_c32fc_o32fc_checkParams_cn(
            ctypes.c_void_p(px), bx, ctypes.c_void_p(pw), bw, ctypes.c_void_p(py), by, nn, 1, ctypes.c_uint8(om))
        return o, c

    assert both(None, bp(f32, n), W, bp(f32, n), Y, bp(f32, n)) == (NULLP, NULLP)
    assert both(X, bp(f32, n), None, bp(f32, n), Y, bp(f32, n)) == (NULLP, NULLP)
    assert both(X, bp(f32, n), W, bp(f32, n), None, bp(f32, n)) == (NULLP, NULLP)
    assert both(X, bp(f32, n), W, bp(f32, n // 2), Y, bp(f32, n)) == (DIM, DIM)
    assert both(X, bp(f32, n), W, bp(f32, n), Y, bp(f32, n // 2)) == (DIM, DIM)
    assert both(X, bp(F.FFTLIB_UINT32, n), W, bp(f32, n), Y, bp(f32, n)) == (TYPE, TYPE)
    assert both(X, bp(f32, n), W, bp(F.FFTLIB_UINT32, n), Y, bp(f32, n)) == (TYPE, TYPE)
    assert both(X, bp(f32, n), W, bp(f32, n), Y, bp(F.FFTLIB_UINT32, n)) == (TYPE, TYPE)
    assert both(X + 4, bp(f32, n), W, bp(f32, n), Y, bp(f32, n)) == (ALIGN, OK)
    assert both(X, bp(f32, n), W, bp(f32, n), Y, bp(f32, n), om=2) == (OUTMODE, OK)
    assert both(X, bp(f32, 61), W, bp(f32, 61), Y, bp(f32, 61)) == (OK, DIM)
    assert both(X, bp(f32, 254), W, bp(f32, 254), Y, bp(f32, 254), nn=254) == (DIM, DIM)
    assert both(X, bp(f32, n), W, bp(f32, n), Y, bp(f32, n)) == (OK, OK)


def test_coverage_c2r_batched(lib, oracle_lib, mem):
    n = 60
    X, W, Y, PB = mem.p(0), mem.p(1), mem.p(2), mem.p(3)
    f32 = F.FFTLIB_FLOAT32

    def both(px, bx, pw, bw, py, by, nn=n, im=0):
        o = lib.FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_checkParams(px, bx, pw, bw, py, by, nn, 1, im, PB)
        c = oracle_lib.FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_checkParams_cn(
            ctypes.c_void_p(px), bx, ctypes.c_void_p(pw), bw, ctypes.c_void_p(py), by, nn, 1, ctypes.c_uint8(im))
        return o, c

    assert both(None, bp(f32, n), W, bp(f32, n), Y, bp(f32, n)) == (NULLP, NULLP)
    assert both(X, bp(f32, n), W, bp(f32, n // 2), Y, bp(f32, n)) == (DIM, DIM)
    assert both(X, bp(F.FFTLIB_UINT32, n), W, bp(f32, n), Y, bp(f32, n)) == (TYPE, TYPE)
    assert both(X + 4, bp(f32, n), W, bp(f32, n), Y, bp(f32, n)) == (ALIGN, OK)
    assert both(X, bp(f32, n), W, bp(f32, n), Y, bp(f32, n), im=7) == (OK, OK)   # inMode is not checked
    assert both(X, bp(f32, 254), W, bp(f32, 254), Y, bp(f32, 254), nn=254) == (DIM, DIM)
    assert both(X, bp(f32, n), W, bp(f32, n), Y, bp(f32, n)) == (OK, OK)


@pytest.mark.parametrize("name", ["FFTLIB_fft1d_i32f_c32fc_o32fc", "FFTLIB_ifft1d_i32fc_c32fc_o32f",
                                  "FFTLIB_ifft1d_i32fc_c32fc_o32fc"])
def test_coverage_single_real_and_inverse(lib, oracle_lib, mem, name):
    X, W, Y, PB = mem.p(0), mem.p(1), mem.p(2), mem.p(3)
    f32 = F.FFTLIB_FLOAT32
    opt = getattr(lib, name + "_checkParams")
    nat = getattr(oracle_lib, name + "_checkParams_cn")

    def both(px, d, dw=None, dt=f32):
        dw = d if dw is None else dw
        o = opt(px, bp(dt, d), W, bp(f32, dw), Y, bp(f32, d), PB)
        c = nat(ctypes.c_void_p(px), bp(dt, d), ctypes.c_void_p(W), bp(f32, dw), ctypes.c_void_p(Y), bp(f32, d))
        return o, c

    assert both(None, 128) == (NULLP, NULLP)
    assert both(X, 128, 64) == (DIM, DIM)
    assert both(X, 128, dt=F.FFTLIB_UINT32) == (TYPE, TYPE)
    assert both(X + 4, 128) == (ALIGN, OK)
    assert both(X, 14) == (DIM, DIM)    # N/2 = 7
    assert both(X, 60) == (OK, OK)


# test/fft_c7x/FFTLIB_fft1dBatched_i16sc_c16sc_o16sc/..._d.c:330-410: the
# optimized check only tests NULL pointers (pBlock included).
def test_coverage_i16_batched(lib, oracle_lib, mem):
    n, b = 64, 2
    X, W, Y, S, PB = mem.p(0), mem.p(1), mem.p(2), mem.p(3), mem.p(4)
    i16, u32 = F.FFTLIB_INT16, F.FFTLIB_UINT32

    def both(px=X, dx=2 * n * b, dy=None, dw=2 * n, tx=i16, ts=u32, ns=3, nn=n, pb=PB, ps=S):
        dy = dx if dy is None else dy
        o = lib.FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_checkParams(px, bp(tx, dx), W, bp(i16, dw), Y, bp(i16, dy), ps,
                                                                  bp(ts, ns), nn, b, pb)
        c = oracle_lib.FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_checkParams_cn(
            ctypes.c_void_p(px), bp(tx, dx), ctypes.c_void_p(W), bp(i16, dw), ctypes.c_void_p(Y), bp(i16, dy),
            ctypes.c_void_p(ps), bp(ts, ns), nn, b)
        return o, c

    assert both(px=None) == (NULLP, NULLP)
    assert both(ps=None) == (NULLP, NULLP)
    assert both(pb=None) == (NULLP, OK)
    assert both(dy=2 * n) == (OK, DIM)
    assert both(dx=2 * n, dy=2 * n) == (OK, DIM)
    assert both(dw=n) == (OK, DIM)
    assert both(tx=F.FFTLIB_FLOAT32) == (OK, TYPE)
    assert both(ts=F.FFTLIB_INT32 if hasattr(F, "FFTLIB_INT32") else 2) == (OK, TYPE)
    assert both(nn=96, dx=2 * 96 * b, dw=192) == (OK, DIM)
    assert both() == (OK, OK)


# test/fft_c7x/FFTLIB_fft1d_i16sc_c16sc_o16sc/..._d.c:293-430
def test_coverage_i16_single(lib, oracle_lib, mem):
    X, W, Y, S, PB = mem.p(0), mem.p(1), mem.p(2), mem.p(3), mem.p(4)
    i16, u32 = F.FFTLIB_INT16, F.FFTLIB_UINT32

    def both(px=X, d=512, dw=None, dy=None, tx=i16, ts=u32, ns=4, pb=PB):
        dw = d if dw is None else dw
        dy = d if dy is None else dy
        o = lib.FFTLIB_fft1d_i16sc_c16sc_o16sc_checkParams(px, bp(tx, d), W, bp(i16, dw), Y, bp(i16, dy), S,
                                                           bp(ts, ns), pb)
        c = oracle_lib.FFTLIB_fft1d_i16sc_c16sc_o16sc_checkParams_cn(
            ctypes.c_void_p(px), bp(tx, d), ctypes.c_void_p(W), bp(i16, dw), ctypes.c_void_p(Y), bp(i16, dy),
            ctypes.c_void_p(S), bp(ts, ns))
        return o, c

    assert both(px=None) == (NULLP, NULLP)
    assert both(pb=None) == (NULLP, OK)
    assert both(dw=256) == (DIM, DIM)
    assert both(dy=256) == (DIM, DIM)
    assert both(d=64, ns=2) == (DIM, OK)        # below the 64-point minimum of the optimized kernel
    assert both(tx=F.FFTLIB_FLOAT32) == (TYPE, TYPE)
    assert both(ts=2) == (TYPE, TYPE)
    assert both(px=X + 4) == (ALIGN, OK)
    assert both(d=192, ns=3) == (DIM, DIM)      # 96 points: neither 4^k nor 2*4^k
    assert both() == (OK, OK)                    # 256 points, 4 shifts


# ---- _init / _kernel host-side behaviour --------------------------------------
def test_init_rejects_unsupported_sizes(lib, mem):
    X, W, Y = mem.p(0), mem.p(1), mem.p(2)
    pb = ctypes.create_string_buffer(F.PBLOCK_SIZE["FFTLIB_fft1dBatched_i32fc_c32fc_o32fc"])
    f32 = F.FFTLIB_FLOAT32
    init = lib.FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_init
    for n, want in ((960, OK), (1024, OK), (7, DIM), (14, DIM), (0, DIM), (1 << 21, DIM)):
        assert init(X, bp(f32, 2 * n), W, bp(f32, 2 * n), Y, bp(f32, 2 * n), n, 1, pb) == want, n
    assert init(X, bp(f32, 8), W, bp(f32, 8), Y, bp(f32, 8), 4, 1, None) == NULLP
    pb16 = ctypes.create_string_buffer(F.PBLOCK_SIZE["FFTLIB_fft1dBatched_i16sc_c16sc_o16sc"])
    i16, u32 = F.FFTLIB_INT16, F.FFTLIB_UINT32
    init16 = lib.FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_init
    assert init16(X, bp(i16, 512), W, bp(i16, 512), Y, bp(i16, 512), mem.p(3), bp(u32, 4), 256, 1, pb16) == OK
    assert init16(X, bp(i16, 480), W, bp(i16, 480), Y, bp(i16, 480), mem.p(3), bp(u32, 4), 240, 1, pb16) == DIM
    pbr = ctypes.create_string_buffer(F.PBLOCK_SIZE["FFTLIB_fft1dBatched_i32f_c32fc_o32fc"])
    initr = lib.FFTLIB_fft1dBatched_i32f_c32fc_o32fc_init
    d = bp(f32, 64)
    assert initr(X, d, W, d, mem.p(4), d, mem.p(5), d, Y, d, 64, 1, 2, pbr) == OUTMODE
    assert initr(X, d, W, d, mem.p(4), d, mem.p(5), d, Y, d, 63, 1, 0, pbr) == DIM


def test_kernel_refuses_bad_calls_before_gpu_work(lib, mem):
    X, W, Y = mem.p(0), mem.p(1), mem.p(2)
    f32 = F.FFTLIB_FLOAT32
    name = "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc"
    init, kern = getattr(lib, name + "_init"), getattr(lib, name + "_kernel")
    pb = ctypes.create_string_buffer(F.PBLOCK_SIZE[name])
    d, dw = bp(f32, 2 * 64 * 2), bp(f32, 128)
    assert kern(X, d, W, dw, Y, d, 64, 2, pb) == INVAL                 # never initialised
    assert lib.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_GPU) == OK
    assert init(X, d, W, dw, Y, d, 64, 2, pb) == OK
    assert kern(X, d, W, dw, Y, d, 128, 2, pb) == DIM                  # size differs from _init
    assert kern(None, d, W, dw, Y, d, 64, 2, pb) == NULLP
    assert kern(X, d, W, dw, Y + 4, d, 64, 2, pb) == ALIGN             # GPU style needs 8-byte alignment
    other = ctypes.create_string_buffer(F.PBLOCK_SIZE[name])
    assert getattr(lib, "FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_init")(X, d, W, dw, Y, d, 64, 2, other) == OK
    assert kern(X, d, W, dw, Y, d, 64, 2, other) == INVAL              # plan of another kernel
    assert lib.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_NATC) == OK
    assert init(X, d, W, dw, Y, d, 64, 2, pb) == OK
    assert kern(X, d, W, dw, Y, d, 64, 2, pb) == NOTIMPL                # natural C is not in the product
    assert lib.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_OPTIMIZED) == OK


def test_async_entry_points_refuse_bad_calls_before_gpu_work(lib, mem):
    """FFTLIB_GPU_kernelAsync and the host pipeline validate the plan and the
    buffers before anything touches the GPU, with _kernel's status codes."""
    X, W, Y, S = mem.p(0), mem.p(1), mem.p(2), mem.p(3)
    f32 = F.FFTLIB_FLOAT32
    name = "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc"
    pb = ctypes.create_string_buffer(F.PBLOCK_SIZE[name])
    d, dw = bp(f32, 2 * 64 * 2), bp(f32, 128)
    run = lib.FFTLIB_GPU_kernelAsync
    assert run(None, X, W, None, Y, None) == NULLP
    assert run(pb, X, W, None, Y, None) == INVAL                       # never initialised
    assert lib.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_GPU) == OK
    assert getattr(lib, name + "_init")(X, d, W, dw, Y, d, 64, 2, pb) == OK
    assert run(pb, None, W, None, Y, None) == NULLP
    assert run(pb, X, W, None, Y + 4, None) == ALIGN                   # device buffers: 8-byte aligned
    pbr = ctypes.create_string_buffer(F.PBLOCK_SIZE["FFTLIB_fft1dBatched_i32f_c32fc_o32fc"])
    assert lib.FFTLIB_fft1dBatched_i32f_c32fc_o32fc_init(X, d, W, d, X, d, S, d, Y, d, 64, 2, 0, pbr) == OK
    assert run(pbr, X, W, None, Y, None) == NULLP                      # real transforms need pSf
    pbi = ctypes.create_string_buffer(F.PBLOCK_SIZE["FFTLIB_fft1dBatched_i16sc_c16sc_o16sc"])
    i16, u32 = F.FFTLIB_INT16, F.FFTLIB_UINT32
    assert lib.FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_init(X, bp(i16, 512), W, bp(i16, 512), Y, bp(i16, 512), S,
                                                           bp(u32, 2), 256, 1, pbi) == OK
    assert run(pbi, X, W, S, Y, None) == DIM                           # 4 radix-4 stages, 2 shifts
    assert lib.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_NATC) == OK
    assert getattr(lib, name + "_init")(X, d, W, dw, Y, d, 64, 2, pb) == OK
    assert run(pb, X, W, None, Y, None) == NOTIMPL
    assert lib.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_OPTIMIZED) == OK
    handle = ctypes.c_void_p(1)
    assert lib.FFTLIB_GPU_pipelineCreate(ctypes.byref(handle), 0) == INVAL
    assert handle.value is None
    assert lib.FFTLIB_GPU_pipelineCreate(ctypes.byref(handle), 9) == INVAL
    assert lib.FFTLIB_GPU_pipelineCreate(None, 2) == NULLP
    assert lib.FFTLIB_GPU_pipelineSubmit(None, pb, X, W, None, Y) == NULLP
    assert lib.FFTLIB_GPU_pipelineWait(None) == NULLP
    assert lib.FFTLIB_GPU_pipelineDestroy(None) == NULLP


# ---- table generators -----------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 15, 16, 30, 60, 64, 120, 960, 1000, 1024, 3125, 4096, 6250, 15625,
                               16384])
def test_tw_gen_c2c_bit_identical_to_oracle(lib, n):
    import oracle
    w = np.zeros(2 * n, np.float32)
    lib.FFTLIB_tw_gen_c2c(w.ctypes.data, n)
    assert np.array_equal(w.view(np.uint32), oracle.tw_c2c(n).view(np.uint32))


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192])
def test_tw_gen_i16_bit_identical_to_oracle(lib, n):
    import oracle
    w = np.zeros(2 * n, np.int16)
    lib.FFTLIB_tw_gen_i16sc(w.ctypes.data, n)
    assert np.array_equal(w, oracle.tw_i16(n))


@pytest.mark.parametrize("n", [2, 4, 10, 64, 250, 1024, 4096, 6250, 16000])
def test_real_tables_bit_identical_to_oracle(lib, n):
    import oracle
    for fn, ref in ((lib.tw_gen_r2c, oracle.tw_r2c), (lib.ifft_c2r_tw_gen, oracle.tw_c2r)):
        w = np.zeros(max(n, 2), np.float32)
        fn(w.ctypes.data, n)
        assert np.array_equal(w[:n].view(np.uint32), ref(n).view(np.uint32))
    for fn, ref in ((lib.split_factor_gen_r2c, oracle.sf_r2c), (lib.ifft_c2r_split_factor_gen, oracle.sf_c2r)):
        s = np.zeros(2 * n, np.float32)
        fn(s.ctypes.data, n)
        assert np.array_equal(s.view(np.uint32), ref(n).view(np.uint32))
