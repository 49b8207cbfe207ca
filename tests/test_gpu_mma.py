"""Fixed-point DFT kernels (FFTLIB_FFT_fft / dftSmall / dftLarge /
highRadixDecompositions _ixX_cxX_oxX) on the GPU through the C-ABI: every
committed reference vector (tests/golden/mma_*.npz) bit-exact against the
oracle's natural-C restatement and within the reference driver's tolerance of
50 of the committed expected output (100 for the four 8192-point cases whose
natural-C rounding the oracle tests document), in device and host memory;
full-range random inputs (saturation, negation of -2^(w-1)) bit-exact."""
import numpy as np
import pytest

import c7x_fftlib_amd as F
import oracle
from tests import fftlib_calls as G

pytestmark = pytest.mark.gpu

WIDE = {("highradix", 16), ("highradix", 17), ("highradix", 49), ("highradix", 68)}


def cases(alias):
    d = np.load(f"tests/golden/mma_{alias}.npz")
    return [(alias, int(r[0])) for r in d["params"]]


ALL = cases("fft") + cases("dftsmall") + cases("dftlarge") + cases("highradix")


@pytest.mark.parametrize("alias,cid", ALL)
def test_reference_vectors(alias, cid):
    d = np.load(f"tests/golden/mma_{alias}.npz")
    row = [r for r in d["params"] if int(r[0]) == cid][0]
    _, dt, b, n, il = [int(v) for v in row]
    x, ref, sh = d[f"in_{cid}"], d[f"out_{cid}"], d[f"shift_{cid}"]
    s, y = G.mma(alias, x, n, b, il, sh, dt, style="gpu")
    assert s == 0, G.F.last_error()
    want = oracle.mma(alias, x, n, b, il, sh, dt)
    assert np.array_equal(y, want), (alias, cid)
    tol = 100 if (alias, cid) in WIDE else 50
    assert int(np.max(np.abs(y.astype(np.int64) - ref.astype(np.int64)))) <= tol


@pytest.mark.parametrize("alias,cid", [("fft", 3), ("fft", 16), ("dftsmall", 9), ("dftlarge", 1006),
                                       ("highradix", 47), ("highradix", 70)])
def test_host_memory(alias, cid):
    d = np.load(f"tests/golden/mma_{alias}.npz")
    row = [r for r in d["params"] if int(r[0]) == cid][0]
    _, dt, b, n, il = [int(v) for v in row]
    x, sh = d[f"in_{cid}"], d[f"shift_{cid}"]
    s, y = G.mma(alias, x, n, b, il, sh, dt, style="host")
    assert s == 0, G.F.last_error()
    assert np.array_equal(y, oracle.mma(alias, x, n, b, il, sh, dt))


FULL = [("dftsmall", 16, 3, 1, F.FFTLIB_INT16, [12]), ("dftsmall", 7, 5, 0, F.FFTLIB_INT32, [28]),
        ("dftlarge", 32, 4, 1, F.FFTLIB_INT16, [13]), ("dftlarge", 40, 3, 0, F.FFTLIB_INT32, [30]),
        ("highradix", 256, 3, 1, F.FFTLIB_INT16, [16, 17, 17, 17, 16]),
        ("highradix", 1024, 2, 0, F.FFTLIB_INT16, [15, 15, 14, 15, 15]),
        ("highradix", 8192, 1, 1, F.FFTLIB_INT16, [16, 16, 16, 16, 16]),
        ("highradix", 64, 4, 1, F.FFTLIB_INT32, [32, 32, 32, 32, 32]),
        ("highradix", 4096, 2, 0, F.FFTLIB_INT32, [31, 33, 31, 31, 33]),
        ("fft", 128, 8, 1, F.FFTLIB_INT16, [0, 15, 15, 15, 15]), ("fft", 32, 8, 0, F.FFTLIB_INT32, [31, 31, 0, 0, 0])]


@pytest.mark.parametrize("alias,n,b,il,dt,sh", FULL)
def test_full_range_inputs(alias, n, b, il, dt, sh):
    """Full-range data including the most negative value: saturating
    rounding shifts and negation are exercised; results bit-exact."""
    r = np.random.Generator(np.random.Philox(n * 7 + b))
    t = np.int16 if dt == F.FFTLIB_INT16 else np.int32
    info = np.iinfo(t)
    x = r.integers(info.min, info.max, 2 * n * b, dtype=np.int64, endpoint=True).astype(t)
    x[:4] = info.min
    shifts = (sh + [0] * 5)[:5]
    s, y = G.mma(alias, x, n, b, il, shifts, dt, style="gpu")
    assert s == 0, G.F.last_error()
    assert np.array_equal(y, oracle.mma(alias, x, n, b, il, shifts, dt))
