"""GPU parity: every fft_c7x entry point, run on the MI355X through the C-ABI,
against the oracle (CPU restatement of the reference's natural-C kernels) on
the same seeded inputs.

Bars: int16 bit-exact; fp32 relative L2 <= 1e-6 against the oracle and, in
addition, the reference drivers' own absolute tolerances (2^-10 for the
single-channel C2C, 2^-5 for batched C2C/IFFT with +-256 inputs, 2^-3 for
R2C, 1e-3 for C2R).
"""
import numpy as np
import pytest

import oracle
from tests import fftlib_calls as G

pytestmark = pytest.mark.gpu

REL_TOL = 1e-6


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def cplx(r, b, n, scale=1.0):
    return (r.uniform(-scale, scale, (b, n)) + 1j * r.uniform(-scale, scale, (b, n))).astype(np.complex64)


C2C_SIZES = [1, 2, 3, 4, 5, 8, 15, 16, 30, 60, 64, 100, 120, 128, 243, 256, 960, 1000, 1024, 1536, 3125, 4096, 6250,
             8192]


@pytest.mark.parametrize("style", ["gpu", "host"])
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("n", C2C_SIZES)
def test_c2c_batched(n, inverse, style):
    b = 3 if n < 4096 else 2
    x = cplx(rng(n * 7 + inverse), b, n, 256.0)
    s, y = G.c2c(x, n, b, inverse=inverse, style=style)
    assert s == 0, G.F.last_error()
    ref = oracle.c2c(x, n, b, inverse=inverse)
    assert G.rel_l2(y, ref) <= REL_TOL
    tol = 2.0 ** -5 if not inverse else 2.0 ** -5
    assert np.max(np.abs(y - ref)) <= max(tol, 1e-6 * np.max(np.abs(ref)))


# Lengths served by the compile-time specialised complex kernels; the batch
# counts leave a partly filled last workgroup.
FIXED_SIZES = [16, 32, 48, 64, 96, 128, 192, 240, 256, 480, 512, 960, 1000, 1024, 1200, 1536, 1920, 2000, 2048,
               3200, 3840, 4000, 4096, 6400, 8192, 16384]


@pytest.fixture
def generic_kernels():
    lib = G.lib()
    assert lib.FFTLIB_GPU_setSpecializedKernels(0) == 0
    yield
    assert lib.FFTLIB_GPU_setSpecializedKernels(1) == 0


@pytest.mark.parametrize("kind", ["fft", "ifft", "ifft_single"])
@pytest.mark.parametrize("n", FIXED_SIZES)
def test_c2c_specialised_lengths(n, kind):
    single = kind == "ifft_single"
    b = 1 if single else (3 if n >= 8192 else (5 if n >= 2048 else 37))
    x = cplx(rng(n * 13 + len(kind)), b, n, 1.0)
    s, y = G.c2c(x, n, b, inverse=kind != "fft", single=single)
    assert s == 0, G.F.last_error()
    ref = oracle.c2c(x, n, b, inverse=kind != "fft", single=single)
    assert G.rel_l2(y, ref) <= REL_TOL


@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("n", [16, 48, 960, 1024, 4096])
def test_c2c_generic_kernels_at_specialised_lengths(n, inverse, generic_kernels):
    b = 9
    x = cplx(rng(n * 17 + inverse), b, n, 1.0)
    s, y = G.c2c(x, n, b, inverse=inverse)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(y, oracle.c2c(x, n, b, inverse=inverse)) <= REL_TOL


@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("n", [16, 64, 128, 256, 512, 1024, 2048, 4096, 30, 60, 120, 240, 480, 960, 1920, 3840,
                               15625])
def test_c2c_single(n, inverse):
    x = cplx(rng(n + 11 * inverse), 1, n, 1.0)
    s, y = G.c2c(x, n, 1, inverse=inverse, single=True)
    assert s == 0, G.F.last_error()
    ref = oracle.c2c(x, n, 1, inverse=inverse, single=True)
    assert G.rel_l2(y, ref) <= REL_TOL
    assert np.max(np.abs(y - ref)) <= 2.0 ** -10


def _i16_input(r, b, n, lo, hi):
    return r.integers(lo, hi + 1, size=(b, n, 2)).astype(np.int16)


I16_CASES = [
    # (n, b, shifts) -- shapes and shift vectors of the reference test list plus edge shapes
    (2, 4, [0]), (4, 4, [1]), (8, 3, [1, 1]), (16, 3, [0, 1]), (32, 3, [1, 0, 1]),
    (64, 32, [0, 1, 1]), (64, 64, [1, 0, 1]), (128, 16, [0, 1, 1]), (128, 256, [1, 0, 0]),
    (256, 8, [0, 1, 1, 1]), (256, 128, [1, 0, 0, 1]), (512, 4, [0, 1, 1, 1]), (1024, 2, [0, 1, 1, 1, 1]),
    (2048, 4, [0, 0, 1, 1, 0]), (4096, 2, [0, 1, 1, 0, 0, 1]), (8192, 1, [1, 0, 0, 0, 1, 1]),
]


@pytest.mark.parametrize("style", ["gpu", "host"])
@pytest.mark.parametrize("n,b,shifts", I16_CASES)
def test_i16_batched_bit_exact(n, b, shifts, style):
    r = rng(n * 31 + b)
    for lo, hi in ((-2048, 2047), (-32768, 32767)):  # config-2 range, then full range (wraps)
        x = _i16_input(r, b, n, lo, hi)
        s, y = G.i16(x, n, b, shifts, style=style)
        assert s == 0, G.F.last_error()
        ref = oracle.i16(x, n, b, shifts)
        assert np.array_equal(y, ref), f"range [{lo},{hi}]: {np.count_nonzero(y != ref)} int16 differ"


def test_i16_saturation_edges():
    n, b = 64, 4
    x = np.full((b, n, 2), -32768, np.int16)
    x[1, ::2, 1] = 32767
    x[2] = np.int16(32767)
    x[3, :, 0] = -32768
    x[3, :, 1] = 0
    for shifts in ([0, 0, 0], [1, 1, 1], [2, 0, 3], [16, 17, 31]):
        s, y = G.i16(x, n, b, shifts)
        assert s == 0
        assert np.array_equal(y, oracle.i16(x, n, b, shifts))


@pytest.mark.parametrize("n,shifts", [(64, [0, 1, 1]), (256, [1, 0, 1, 1]), (1024, [0, 1, 1, 1, 1]),
                                      (2048, [1, 0, 1, 0, 1])])
def test_i16_single_bit_exact(n, shifts):
    x = _i16_input(rng(n), 1, n, -32768, 32767)
    s, y = G.i16(x, n, 1, shifts, single=True)
    assert s == 0, G.F.last_error()
    assert np.array_equal(y, oracle.i16(x, n, 1, shifts, single=True))


@pytest.mark.parametrize("n", [16, 64, 256, 1024, 4096])
def test_i16_specialised_and_generic_kernels_agree(n):
    """N = 4^K runs the specialised kernel on 16-byte aligned buffers and the
    one-stage-per-pass kernel on buffers 4 bytes off that; both bit-exact."""
    import ctypes
    import torch
    F = G.F
    k = int(round(np.log(n) / np.log(4)))
    shifts = [1, 0, 1, 1, 0, 1][:k]
    b = 37
    x = _i16_input(rng(n + 5), b, n, -32768, 32767)
    ref = oracle.i16(x, n, b, shifts)
    L = G.lib()
    assert L.FFTLIB_setFunctionStyle(F.FFTLIB_FUNCTION_GPU) == 0
    L.FFTLIB_GPU_setSynchronous(1)
    name = "FFTLIB_fft1dBatched_i16sc_c16sc_o16sc"
    pb = G.pblock(name)
    bx, bw = ctypes.byref(F.BufParams1D(F.FFTLIB_INT16, 2 * n * b)), ctypes.byref(F.BufParams1D(F.FFTLIB_INT16, 2 * n))
    bs = ctypes.byref(F.BufParams1D(F.FFTLIB_UINT32, len(shifts)))
    dw = torch.from_numpy(G.tw_i16(n)).cuda()
    ds = torch.tensor(shifts, dtype=torch.int32, device="cuda")
    for off in (0, 2):                      # int16 elements: 0 -> 16-byte aligned, 2 -> 4 bytes off
        dx = torch.zeros(x.size + 8, dtype=torch.int16, device="cuda")
        dx[off:off + x.size] = torch.from_numpy(x.reshape(-1)).cuda()
        dy = torch.zeros_like(dx)
        xp, yp = dx.data_ptr() + 2 * off, dy.data_ptr() + 2 * off
        args = (xp, bx, dw.data_ptr(), bw, yp, bx, ds.data_ptr(), bs, n, b, pb)
        assert getattr(L, name + "_init")(*args) == 0
        assert getattr(L, name + "_kernel")(*args) == 0, F.last_error()
        y = dy[off:off + x.size].cpu().numpy().reshape(b, n, 2)
        assert np.array_equal(y, ref), (n, off)
        assert not dy[:off].any() and not dy[off + x.size:].any()   # nothing written outside the rows


R2C_SIZES = [(10, 16), (50, 4), (64, 1), (128, 2), (250, 4), (256, 4), (512, 8), (1024, 4), (1250, 4), (1600, 2),
             (4096, 3), (6250, 1), (8000, 2), (12800, 1), (16000, 1)]


@pytest.mark.parametrize("outmode", [0, 1])
@pytest.mark.parametrize("n,b", R2C_SIZES)
def test_r2c_batched(n, b, outmode):
    x = rng(n + outmode).uniform(-128, 128, (b, n)).astype(np.float32)
    s, y = G.r2c(x, n, b, outmode=outmode)
    assert s == 0, G.F.last_error()
    ref = oracle.r2c(x, n, b, outmode=outmode)
    assert G.rel_l2(y, ref) <= REL_TOL
    assert np.max(np.abs(y - ref)) <= 2.0 ** -3


@pytest.mark.parametrize("inmode", [0, 1])
@pytest.mark.parametrize("n,b", R2C_SIZES)
def test_c2r_batched(n, b, inmode):
    r = rng(n * 3 + inmode)
    bins = n // 2 + 1 if inmode == 0 else n
    x = cplx(r, b, bins, 100.0)
    s, y = G.c2r(x, n, b, inmode=inmode)
    assert s == 0, G.F.last_error()
    ref = oracle.c2r(x, n, b, inmode=inmode)
    assert G.rel_l2(y, ref) <= REL_TOL
    assert np.max(np.abs(y - ref)) <= 1e-3


# Real lengths N whose inner complex length N/2 has specialised kernels.
REAL_FIXED_SIZES = [64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 960, 1920]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n", REAL_FIXED_SIZES)
def test_real_specialised_lengths(n, mode):
    b = 3 if n >= 8192 else 37
    x = rng(n * 5 + mode).uniform(-1, 1, (b, n)).astype(np.float32)
    s, y = G.r2c(x, n, b, outmode=mode)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(y, oracle.r2c(x, n, b, outmode=mode)) <= REL_TOL
    bins = n // 2 + 1 if mode == 0 else n
    spec = cplx(rng(n * 7 + mode), b, bins, 1.0)
    s, z = G.c2r(spec, n, b, inmode=mode)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(z, oracle.c2r(spec, n, b, inmode=mode)) <= REL_TOL


@pytest.mark.parametrize("n", [64, 960, 4096])
def test_real_generic_kernels_at_specialised_lengths(n, generic_kernels):
    b = 5
    x = rng(n * 11).uniform(-1, 1, (b, n)).astype(np.float32)
    s, y = G.r2c(x, n, b, outmode=0)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(y, oracle.r2c(x, n, b, outmode=0)) <= REL_TOL
    s, z = G.c2r(y, n, b, inmode=0)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(z, oracle.c2r(y, n, b, inmode=0)) <= REL_TOL


@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 2048, 4096, 8192, 10, 50, 250, 1250])
def test_real_single(n):
    x = rng(n).uniform(-1, 1, (1, n)).astype(np.float32)
    s, y = G.r2c(x, n, 1, single=True)
    assert s == 0, G.F.last_error()
    ref = oracle.r2c(x, n, 1, single=True)
    assert G.rel_l2(y, ref) <= REL_TOL
    s, z = G.c2r(y, n, 1, single=True)
    assert s == 0, G.F.last_error()
    refz = oracle.c2r(y, n, 1, single=True)
    assert G.rel_l2(z, refz) <= REL_TOL


# Split-precision MFMA variants of the batched / single C2C: error against an
# FP64 DFT. FP16x2 and BF16x3 meet the 2e-6 bar; BF16x2 (3 bf16 products) is
# held to its own, looser bound. Every length with an MFMA kernel (the three
# shapes R3 = 1, 2, 4 of N = R1 x 16 x R3, R1 = 2 .. 16), batch counts that
# leave a partly filled last workgroup.
MFMA_VARIANTS = {1: 2e-6, 2: 2e-6, 3: 1e-5}
MFMA_LENGTHS = [32, 48, 64, 80, 96, 128, 144, 160, 192, 240, 256, 288, 320, 384, 480, 512, 576, 640, 768, 960, 1024]


@pytest.fixture
def c2c_variant():
    lib = G.lib()
    yield lambda v: lib.FFTLIB_GPU_setC2CVariant(v)
    assert lib.FFTLIB_GPU_setC2CVariant(0) == 0


@pytest.mark.parametrize("kind", ["fft", "ifft", "ifft_single"])
@pytest.mark.parametrize("variant", sorted(MFMA_VARIANTS))
@pytest.mark.parametrize("n", MFMA_LENGTHS)
def test_c2c_mfma_variants(n, variant, kind, c2c_variant):
    assert c2c_variant(variant) == 0
    single = kind == "ifft_single"
    b = 1 if single else 77
    x = cplx(rng(variant * 31 + len(kind) + n), b, n, 1.0)
    s, y = G.c2c(x, n, b, inverse=kind != "fft", single=single)
    assert s == 0, G.F.last_error()
    x64 = x.astype(np.complex128)
    ref = np.fft.ifft(x64, axis=1) if kind != "fft" else np.fft.fft(x64, axis=1)
    assert G.rel_l2(y, ref) <= MFMA_VARIANTS[variant]


@pytest.mark.parametrize("n", [240, 480, 960])
@pytest.mark.parametrize("variant", [1, 2])
def test_c2c_mfma_scaled_inputs(variant, n, c2c_variant):
    """Inputs far from unit magnitude (fp16 parts need the power-of-two scaling)."""
    assert c2c_variant(variant) == 0
    b = 6
    x = cplx(rng(99 + variant), b, n, 1.0) * np.array([1e-6, 1e-3, 1.0, 3e2, 6e4, 2e7], np.float32)[:, None]
    s, y = G.c2c(x.astype(np.complex64), n, b)
    assert s == 0, G.F.last_error()
    ref = np.fft.fft(x.astype(np.complex128), axis=1)
    for i in range(b):
        assert G.rel_l2(y[i], ref[i]) <= 2e-6, i


def test_c2c_variant_other_lengths_use_fp32(c2c_variant):
    """Lengths without an MFMA kernel keep the fp32 path (bit-identical results)."""
    n, b = 1000, 3
    x = cplx(rng(5), b, n, 1.0)
    s0, y0 = G.c2c(x, n, b)
    assert c2c_variant(2) == 0
    s1, y1 = G.c2c(x, n, b)
    assert s0 == 0 and s1 == 0
    assert np.array_equal(y0.view(np.uint32), y1.view(np.uint32))


# Lengths between 4096 and 16384 points without a dedicated kernel: the
# octave kernels, one transform per workgroup of N / 16 work items (512 up to
# 8192, 1024 beyond); odd lengths included (15625 = 5^6, 6561 = 3^8).
CLASS_LONG_SIZES = [4320, 5000, 6250, 6561, 7776, 8640, 10000, 12000, 12500, 15625, 16200]


@pytest.mark.parametrize("kind", ["fft", "ifft", "ifft_single"])
@pytest.mark.parametrize("n", CLASS_LONG_SIZES)
def test_c2c_class_long(n, kind):
    single = kind == "ifft_single"
    b = 1 if single else 3
    x = cplx(rng(n + 3 * len(kind)), b, n, 1.0)
    s, y = G.c2c(x, n, b, inverse=kind != "fft", single=single)
    assert s == 0, G.F.last_error()
    ref = oracle.c2c(x, n, b, inverse=kind != "fft", single=single)
    assert G.rel_l2(y, ref) <= REL_TOL


# Complex lengths longer than 16384 points: a column pass and a row pass on
# the multi-length kernels (n = n1 * m2: 20000 = 160 * 125, 16875 = 135 * 125,
# 65536 = 256 * 256, 2^20 = 1024 * 1024); with the generic kernels, outer radix
# passes over HBM plus sub-blocks finished in LDS.
LARGE_SIZES = [16875, 20000, 30720, 32768, 40000, 50000, 59049, 65536]


@pytest.mark.parametrize("kind", ["fft", "ifft", "ifft_single"])
@pytest.mark.parametrize("n", LARGE_SIZES)
def test_c2c_large(n, kind):
    single = kind == "ifft_single"
    b = 1 if single else 3
    x = cplx(rng(n + len(kind)), b, n, 1.0)
    s, y = G.c2c(x, n, b, inverse=kind != "fft", single=single)
    assert s == 0, G.F.last_error()
    ref = oracle.c2c(x, n, b, inverse=kind != "fft", single=single)
    assert G.rel_l2(y, ref) <= REL_TOL


def test_c2c_large_two_pass_longest():
    n = 1 << 20
    x = cplx(rng(20), 2, n, 1.0)
    s, y = G.c2c(x, n, 2)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(y, oracle.c2c(x, n, 2)) <= REL_TOL


@pytest.mark.parametrize("n", [20480, 20000, 48600, 16875, 65536])
def test_c2c_two_pass_output_offsets(n):
    """Two-pass transforms with input and output 8, 64 and 136 bytes off a
    128-byte line: the rows of the intermediate copy are rotated by the output
    address where N1 is a multiple of 16 or 8 (N1 x M2: 20480 = 128 x 160,
    20000 = 160 x 125, 65536 = 256 x 256; 48600 = 216 x 225 to 64-byte
    pieces; 16875 = 135 x 125 keeps them in order), which moves points but
    not their arithmetic, so every offset gives the aligned call's bits."""
    import torch
    b = 5
    x = cplx(rng(n + 11), b, n, 1.0)
    L = G.lib()
    G._set_style("gpu")
    xin = np.ascontiguousarray(x.astype(np.complex64)).view(np.float32).reshape(-1)
    w = torch.from_numpy(G.tw_c2c(n)).cuda()
    name = "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc"
    bx, bw = G._bp(G.F.FFTLIB_FLOAT32, 2 * n * b), G._bp(G.F.FFTLIB_FLOAT32, 2 * n)
    outs = {}
    for off in (0, 8, 64, 136):
        xt = torch.zeros(xin.size + 64, dtype=torch.float32, device="cuda:0")
        yt = torch.full((xin.size + 64,), float("nan"), dtype=torch.float32, device="cuda:0")
        k = off // 4
        xt[k:k + xin.size] = torch.from_numpy(xin).cuda()
        xp, yp = xt.data_ptr() + off, yt.data_ptr() + off
        print(f"[parity] two-pass n={n} b={b} offset {off} B", flush=True)
        pb = G.pblock(name)
        assert getattr(L, name + "_init")(xp, bx, w.data_ptr(), bw, yp, bx, n, b, pb) == 0
        assert getattr(L, name + "_kernel")(xp, bx, w.data_ptr(), bw, yp, bx, n, b, pb) == 0, G.F.last_error()
        torch.cuda.synchronize()
        y = yt.cpu().numpy()
        assert np.isnan(y[:k]).all() and np.isnan(y[k + xin.size:]).all(), "wrote outside the output"
        outs[off] = y[k:k + xin.size].view(np.complex64).reshape(b, n)
    assert G.rel_l2(outs[0], oracle.c2c(x, n, b)) <= REL_TOL
    for off in (8, 64, 136):
        assert np.array_equal(outs[off].view(np.uint32), outs[0].view(np.uint32)), off


@pytest.mark.parametrize("kind", ["fft", "ifft"])
@pytest.mark.parametrize("n", [20000, 65536])
def test_c2c_large_generic(n, kind, generic_kernels):
    x = cplx(rng(n + 7), 3, n, 1.0)
    s, y = G.c2c(x, n, 3, inverse=kind != "fft")
    assert s == 0, G.F.last_error()
    assert G.rel_l2(y, oracle.c2c(x, n, 3, inverse=kind != "fft")) <= REL_TOL


# Real transforms whose half length does not fit one workgroup's LDS: the
# half-length complex FFT and the split run as two passes through library
# scratch (the reference's test lists go up to N = 62500 batched and 156250
# single-channel). Half lengths: 20000, 31250, 32768, 65536.
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n,b", [(40000, 3), (62500, 2), (65536, 2), (131072, 1)])
def test_real_long_batched(n, b, mode):
    x = rng(n + mode).uniform(-1, 1, (b, n)).astype(np.float32)
    s, y = G.r2c(x, n, b, outmode=mode)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(y, oracle.r2c(x, n, b, outmode=mode)) <= REL_TOL
    s, z = G.c2r(y, n, b, inmode=mode)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(z, oracle.c2r(y, n, b, inmode=mode)) <= REL_TOL


@pytest.mark.parametrize("n", [65610, 156250])
def test_real_long_single(n):
    x = rng(n).uniform(-128, 128, (1, n)).astype(np.float32)
    s, y = G.r2c(x, n, 1, single=True)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(y, oracle.r2c(x, n, 1, single=True)) <= REL_TOL
    s, z = G.c2r(y, n, 1, single=True)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(z, oracle.c2r(y, n, 1, single=True)) <= REL_TOL


def test_real_long_generic(generic_kernels):
    """The same split passes around the generic outer-pass complex kernels."""
    n, b = 40000, 2
    x = rng(9).uniform(-1, 1, (b, n)).astype(np.float32)
    s, y = G.r2c(x, n, b, outmode=0)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(y, oracle.r2c(x, n, b, outmode=0)) <= REL_TOL
    s, z = G.c2r(y, n, b, inmode=0)
    assert s == 0, G.F.last_error()
    assert G.rel_l2(z, oracle.c2r(y, n, b, inmode=0)) <= REL_TOL


def test_i16_too_long_is_refused():
    n = 32768
    x = np.zeros((1, n, 2), np.int16)
    s, _ = G.i16(x, n, 1, [0] * 8, single=True)
    assert s == G.F.FFTLIB_ERR_NOT_IMPLEMENTED
