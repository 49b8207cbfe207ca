"""CPU checks of the fixed-point DFT (fft_c7xmma) oracle and of the host side
of the product's FFTLIB_FFT_* API (sizes, tables, parameter checks).

The oracle (oracle/fftlib_mma_cn.c) restates the reference's natural-C
kernels. It is pinned by
  * the reference's committed vectors (test/fft_c7xmma/<kernel>/<kernel>_idat.c,
    converted to tests/golden/mma_*.npz by tests/golden/extract_mma_fixtures.py),
    within the reference driver's tolerance of 50;
  * closed forms at full scale: an impulse gives a flat spectrum and a tone a
    single bin, within a few units, for every kernel and data type.
The committed expected outputs come from a floating-point model (they agree
with the FP64 DFT to about 3 units); the natural-C algorithm rounds every
stage to the data type, and for 8192-point transforms of 8-bit-range inputs
its result differs from them by up to 87 units (highRadixDecompositions cases
16, 17, 49, 68). Those four cases are checked against a bound of 100 and
against the FP64 DFT within the same bound.
"""
import ctypes

import numpy as np
import pytest

import c7x_fftlib_amd as F
import oracle
from tests import fftlib_calls as G

FIXTURE_TOL = 50
WIDE = {("highradix", 16), ("highradix", 17), ("highradix", 49), ("highradix", 68)}
WIDE_TOL = 100


def fixture_cases(alias):
    d = np.load(f"tests/golden/mma_{alias}.npz")
    for cid, dt, b, n, il in d["params"]:
        yield int(cid), int(dt), int(b), int(n), int(il), d[f"in_{cid}"], d[f"out_{cid}"], d[f"shift_{cid}"]


def complex_of(a, n, b, il):
    a = np.asarray(a, np.float64).reshape(b, 2 * n)
    return a[:, 0::2] + 1j * a[:, 1::2] if il else a[:, :n] + 1j * a[:, n:]


@pytest.mark.parametrize("alias", ["fft", "dftsmall", "dftlarge", "highradix"])
def test_oracle_matches_reference_vectors(alias):
    n_cases = 0
    for cid, dt, b, n, il, x, ref, sh in fixture_cases(alias):
        y = oracle.mma(alias, x, n, b, il, sh, dt)
        err = int(np.max(np.abs(y.astype(np.int64) - ref.astype(np.int64))))
        tol = WIDE_TOL if (alias, cid) in WIDE else FIXTURE_TOL
        assert err <= tol, (alias, cid, err)
        n_cases += 1
    assert n_cases >= 10


def test_wide_cases_are_the_natural_c_rounding():
    """For the four 8192-point cases the oracle stays within 100 of the FP64
    DFT (the committed outputs are within 3 of it)."""
    d = np.load("tests/golden/mma_highradix.npz")
    for cid, dt, b, n, il in d["params"]:
        if ("highradix", int(cid)) not in WIDE:
            continue
        x = d[f"in_{cid}"]
        y = oracle.mma("highradix", x, int(n), int(b), int(il), d[f"shift_{cid}"], int(dt))
        X = np.fft.fft(complex_of(x, int(n), int(b), int(il)), axis=1)
        Y = complex_of(y, int(n), int(b), int(il))
        R = complex_of(d[f"out_{cid}"], int(n), int(b), int(il))
        assert np.max(np.abs(R - X)) < 5
        assert np.max(np.abs(Y - X)) < WIDE_TOL


SHIFT = {F.FFTLIB_INT16: 15, F.FFTLIB_INT32: 31}
AMP = {F.FFTLIB_INT16: 1000, F.FFTLIB_INT32: 1000000}


@pytest.mark.parametrize("alias,n,dt", [("dftsmall", 8, F.FFTLIB_INT16), ("dftsmall", 5, F.FFTLIB_INT32),
                                        ("dftlarge", 32, F.FFTLIB_INT16), ("dftlarge", 40, F.FFTLIB_INT32),
                                        ("highradix", 256, F.FFTLIB_INT16), ("highradix", 2048, F.FFTLIB_INT16),
                                        ("highradix", 64, F.FFTLIB_INT32), ("highradix", 4096, F.FFTLIB_INT32)])
@pytest.mark.parametrize("il", [0, 1])
def test_oracle_closed_forms(alias, n, dt, il):
    """Impulse -> flat spectrum, tone -> one bin, at gain 1 (shift = data
    width - 1 for every stage). Every stage rounds to the data type, so the
    error grows like sqrt(N) units (5 units at 256 points, 36 at 2048 for an
    INT16 tone of amplitude 14); the quantised tables add 1e-4 relative. A
    wrong twiddle, index or sign moves whole bins by the signal's size."""
    b = 2
    amp = AMP[dt]
    sh = [SHIFT[dt]] * 5
    t = np.arange(n)
    tone = 3 % n
    z = np.zeros((b, n), np.complex128)
    z[0, 0] = amp
    peak = 0.9 * (32767 if dt == F.FFTLIB_INT16 else 2 ** 31) / n   # the tone's bin stays in range
    z[1] = min(amp / 8, peak) * np.exp(2j * np.pi * tone * t / n)
    zr = np.round(z)
    if il:
        x = np.stack([zr.real, zr.imag], axis=2).reshape(-1)
    else:
        x = np.concatenate([zr.real, zr.imag], axis=1).reshape(-1)
    y = oracle.mma(alias, x.astype(np.int64), n, b, il, sh, dt)
    Y = complex_of(y, n, b, il)
    X = np.fft.fft(zr, axis=1)
    assert np.max(np.abs(Y - X)) <= 1.5 * np.sqrt(n) + 2e-4 * np.max(np.abs(X)), (alias, n, dt, il)


def test_kernel_choice():
    """The coordinator's choice (FFTLIB_FFT_fft_ixX_cxX_oxX.cpp:13-49)."""
    small, large, high = 0, 1, 2
    i16, i32 = F.FFTLIB_INT16, F.FFTLIB_INT32
    cases = [((16, 1, 1, i16), small), ((17, 1, 1, i16), high), ((32, 4, 0, i16), large),
             ((64, 8, 1, i16), high), ((64, 16, 1, i16), high), ((64, 16, 0, i16), large), ((64, 32, 1, i16), large),
             ((128, 64, 1, i16), high), ((8, 3, 1, i32), small), ((16, 32, 1, i32), large), ((32, 4, 1, i32), high),
             ((32, 8, 0, i32), large), ((32, 8, 1, i32), high), ((32, 9, 1, i32), large)]
    for (n, b, il, dt), want in cases:
        assert oracle.mma_variant(n, b, il, dt) == want, (n, b, il, dt)


# ---- host side of the product API (no GPU work) ------------------------------------
def _args(alias, n, b, il, shifts, style=F.FFTLIB_FUNCTION_OPTIMIZED):
    a = F.MMA_KERNELS[alias][1]()
    a.funcStyle, a.fftSize, a.batchSize, a.interleave = style, n, b, il
    for i in range(len(a.shiftVector)):
        a.shiftVector[i] = shifts
    return a


@pytest.mark.parametrize("alias,n,b,il,dt", [("dftsmall", 8, 4, 1, 1), ("dftsmall", 7, 3, 0, 2), ("dftsmall", 16, 1, 0, 1),
                                             ("dftlarge", 40, 2, 0, 1), ("dftlarge", 64, 3, 1, 2),
                                             ("highradix", 1024, 1, 1, 1), ("highradix", 8192, 1, 0, 2),
                                             ("fft", 128, 2, 1, 1), ("fft", 32, 16, 1, 2), ("fft", 8, 5, 0, 2)])
def test_product_tables_bit_identical_to_oracle(alias, n, b, il, dt):
    lib = G.lib()
    prefix = F.MMA_KERNELS[alias][0]
    a = _args(alias, n, b, il, SHIFT[dt])
    bx, bw, by = F.BufParams1D(dt, 0), F.BufParams1D(dt, 0), F.BufParams1D(dt, 0)
    assert getattr(lib, prefix + "_getSizes")(ctypes.byref(a), ctypes.byref(bx), ctypes.byref(bw),
                                              ctypes.byref(by)) == 0
    assert bx.dim_x == by.dim_x == 2 * n * b
    t = np.int16 if dt == 1 else np.int32
    w = np.zeros(bw.dim_x, t)
    assert getattr(lib, prefix + "_twGen")(ctypes.byref(a), w.ctypes.data, ctypes.byref(bw)) == 0
    kind = alias if alias != "fft" else ("dftsmall", "dftlarge", "highradix")[oracle.mma_variant(n, b, il, dt)]
    ref, _ = oracle.mma_tables(kind, n, b, il, dt)
    assert w.shape == ref.shape and np.array_equal(w, ref)
    bad = F.BufParams1D(dt, bw.dim_x + 1)
    assert getattr(lib, prefix + "_twGen")(ctypes.byref(a), w.ctypes.data, ctypes.byref(bad)) == \
        F.FFTLIB_ERR_INVALID_DIMENSION


def test_product_parameter_checks():
    lib = G.lib()
    i16, i32 = F.FFTLIB_INT16, F.FFTLIB_INT32

    def check(alias, a, dt, dims=None, types=(None, None, None), handle=True):
        prefix = F.MMA_KERNELS[alias][0]
        bx, bw, by = F.BufParams1D(dt, 0), F.BufParams1D(dt, 0), F.BufParams1D(dt, 0)
        getattr(lib, prefix + "_getSizes")(ctypes.byref(a), ctypes.byref(bx), ctypes.byref(bw), ctypes.byref(by))
        if dims:
            bx.dim_x, bw.dim_x, by.dim_x = [d if d is not None else o for d, o in zip(dims, (bx.dim_x, bw.dim_x,
                                                                                             by.dim_x))]
        for bp, ty in zip((bx, bw, by), types):
            if ty is not None:
                bp.data_type = ty
        h = ctypes.create_string_buffer(getattr(lib, prefix + "_getHandleSize")(ctypes.byref(a))) if handle else None
        return getattr(lib, prefix + "_init_checkParams")(h, ctypes.byref(bx), ctypes.byref(bw), ctypes.byref(by),
                                                          ctypes.byref(a))

    a = _args("fft", 64, 4, 1, 15)
    assert check("fft", a, i16) == 0
    assert check("fft", a, i16, handle=False) == F.FFTLIB_ERR_NULL_POINTER
    assert check("fft", a, F.FFTLIB_FLOAT32) == F.FFTLIB_ERR_NOT_IMPLEMENTED
    assert check("fft", a, i16, dims=(100, None, None)) == F.FFTLIB_ERR_INVALID_DIMENSION
    assert check("fft", a, i16, dims=(None, 7, None)) == F.FFTLIB_ERR_INVALID_DIMENSION
    assert check("fft", a, i16, types=(None, i32, None)) == F.FFTLIB_ERR_INVALID_TYPE
    # highRadixDecompositions called directly below 2*MMA points
    assert check("highradix", _args("highradix", 32, 1, 1, 15), i16) == F.FFTLIB_ERR_INVALID_VALUE
    # dftSmall above MMA/2 points and dftLarge at or below it: getSizes refuses
    assert check("dftsmall", _args("dftsmall", 17, 1, 1, 15), i16) == F.FFTLIB_ERR_INVALID_DIMENSION
    assert check("dftlarge", _args("dftlarge", 16, 1, 1, 15), i16) == F.FFTLIB_ERR_INVALID_DIMENSION
    # the matrix product's shift limit (an int8_t against 64 bits for INT16)
    assert check("dftlarge", _args("dftlarge", 32, 1, 1, 70), i16) == F.FFTLIB_ERR_INVALID_SHIFT
    assert check("dftlarge", _args("dftlarge", 32, 1, 1, 64), i16) == 0
    # exec_checkParams: NULL buffers
    h = ctypes.create_string_buffer(64)
    assert lib.FFTLIB_FFT_fft_ixX_cxX_oxX_exec_checkParams(h, None, None, None) == F.FFTLIB_ERR_NULL_POINTER
    assert lib.FFTLIB_FFT_fft_ixX_cxX_oxX_exec_checkParams(h, 8, 8, 8) == 0
    # natural-C style has no implementation in the product
    a = _args("fft", 64, 1, 1, 15, style=F.FFTLIB_FUNCTION_NATC)
    bx, bw, by = F.BufParams1D(i16, 128), F.BufParams1D(i16, 0), F.BufParams1D(i16, 128)
    lib.FFTLIB_FFT_fft_ixX_cxX_oxX_getSizes(ctypes.byref(a), ctypes.byref(bx), ctypes.byref(bw), ctypes.byref(by))
    h = ctypes.create_string_buffer(lib.FFTLIB_FFT_fft_ixX_cxX_oxX_getHandleSize(ctypes.byref(a)))
    assert lib.FFTLIB_FFT_fft_ixX_cxX_oxX_init(h, ctypes.byref(bx), ctypes.byref(bw), ctypes.byref(by),
                                               ctypes.byref(a)) == F.FFTLIB_ERR_NOT_IMPLEMENTED
    # exec on a handle that init never filled
    assert lib.FFTLIB_FFT_fft_ixX_cxX_oxX_exec(ctypes.create_string_buffer(64), 8, 8, 8) == F.FFTLIB_ERR_INVALID_VALUE
